"""Facets of one major cycle batched into one HIP graph (wsclean_amd.facets,
hws_graph_capture_*_batch).

- Every replay of the batch graph reproduces the eager per-facet results bit
  for bit, for invert and predict, with eager work between the replays.
- Each facet image matches the oracle's w-stacking gridder at the facet's
  shift, with the visibilities rotated as WSMSGridder rotates them
  (gridding/wsmsgridder.cpp:185-190).
- Facet decomposition: the predicted visibilities of the facets of a model,
  summed, match the FP64 direct predict of the whole model, and the facet
  dirty images pasted together match the direct invert of the whole field.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

C_LIGHT = 299792458.0


def _samples(seed, dl, nrow=400, nchan=3, frac=0.42, wmax_m=60.0, nimg=1):
    rng = np.random.default_rng(seed)
    wl = C_LIGHT / np.linspace(120e6, 150e6, nchan)
    umax = frac / dl * wl.min()
    uvw = np.column_stack([rng.uniform(-umax, umax, nrow), rng.uniform(-umax, umax, nrow),
                           rng.uniform(-wmax_m, wmax_m, nrow)])
    vis = (rng.standard_normal((nrow, nchan, nimg)) +
           1j * rng.standard_normal((nrow, nchan, nimg))).astype(np.complex64)
    return uvw, wl, vis


def _rel(a, b):
    a = np.asarray(a, np.complex128)
    b = np.asarray(b, np.complex128)
    return np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel())


def _uneven_facets(dl):
    from wsclean_amd.facets import Facet
    return [Facet(96, 64, 0.021, -0.013), Facet(128, 128, -0.017, 0.02),
            Facet(64, 96, 0.0, 0.0), Facet(160, 96, 0.03, 0.025)]


@pytest.mark.parametrize("precision", ["mixed", "reference"])
def test_batch_graph_replays_match_eager(precision):
    from wsclean_amd.facets import FacetBatch
    dl = 0.003
    uvw, wl, vis = _samples(31, dl, nimg=2)
    dev = torch.device("cuda:0")
    d_uvw, d_wl = torch.from_numpy(uvw).to(dev), torch.from_numpy(wl).to(dev)
    d_vis = torch.from_numpy(vis).to(dev)
    fb = FacetBatch(_uneven_facets(dl), dl, dl, d_uvw, d_wl, n_images=2, complex_images=(1,),
                    precision=precision, n_w_layers=8, work_chunk=64)
    assert fb.passes == [1] * 4
    re, im = fb.images(), fb.images()
    fb.invert(d_vis, 0.5, re, im)
    outs = [torch.empty_like(d_vis) for _ in fb.gridders]
    fb.predict(re, im, outs)
    torch.cuda.synchronize()
    re_e = [t.clone() for t in re]
    im_e = [t.clone() for t in im]
    out_e = [t.clone() for t in outs]

    gre = [torch.full_like(t, 7.0) for t in re]
    gim = [torch.full_like(t, 7.0) for t in im]
    gout = [torch.full_like(t, 3.0) for t in outs]
    ginv = fb.capture_invert(d_vis, 0.5, gre, gim)
    gpred = fb.capture_predict(re_e, im_e, gout)
    scratch = [torch.empty_like(t) for t in outs]
    for k in range(3):
        for t in gre + gim + gout:
            t.fill_(-1.0)
        ginv.launch()
        gpred.launch()
        torch.cuda.synchronize()
        for i in range(len(re)):
            assert torch.equal(gre[i], re_e[i]), (k, i)
            assert torch.equal(gim[i][1], im_e[i][1]), (k, i)
            ok = ~torch.isnan(out_e[i])
            assert torch.equal(torch.isnan(gout[i]), ~ok), (k, i)
            assert torch.equal(gout[i][ok], out_e[i][ok]), (k, i)
        fb.predict(re_e, im_e, scratch)  # eager work between replays
        torch.cuda.synchronize()
    ginv.close()
    gpred.close()
    fb.close()


def test_batch_capture_rejects_repeated_and_multipass_contexts():
    import ctypes as C
    from wsclean_amd._lib import HwsError
    from wsclean_amd.facets import Facet, FacetBatch
    dl = 0.003
    uvw, wl, vis = _samples(32, dl, nrow=60)
    dev = torch.device("cuda:0")
    d_uvw, d_wl = torch.from_numpy(uvw).to(dev), torch.from_numpy(wl).to(dev)
    d_vis = torch.from_numpy(vis).to(dev)
    fb = FacetBatch([Facet(64, 64, 0.01, 0.0)], dl, dl, d_uvw, d_wl, n_w_layers=4)
    g = fb.gridders[0]
    re = fb.images()[0]
    hs = (C.c_void_p * 2)(g.h.value, g.h.value)
    vs = (C.c_void_p * 2)(d_vis.data_ptr(), d_vis.data_ptr())
    mult = (C.c_double * 2)(1.0, 1.0)
    rs = (C.c_void_p * 2)(re.data_ptr(), re.data_ptr())
    gr = C.c_void_p()
    st = g.L.hws_graph_capture_invert_batch(hs, 2, vs, mult, rs, None, C.byref(gr))
    assert st != 0 and b"twice" in g.L.hws_last_error(g.h)
    fb.close()
    # a facet that needs two passes cannot be captured
    fb = FacetBatch([Facet(64, 64, 0.0, 0.0), Facet(64, 64, 0.01, 0.0)], dl, dl, d_uvw, d_wl,
                    n_w_layers=4)
    fb.gridders[1].prepare_w_layers(8, 0.0, 200.0, max_mem=64 * 64 * 8 * 3)
    fb.gridders[1].plan(d_uvw, d_wl)
    with pytest.raises(HwsError, match="single pass"):
        fb.capture_invert(d_vis, 1.0, fb.images())
    fb.close()


def test_facet_images_match_oracle(oracle):
    from wsclean_amd.facets import FacetBatch
    dl = 0.003
    uvw, wl, vis = _samples(33, dl)
    uvw_l = (uvw[:, None, :] / wl[None, :, None]).reshape(-1, 3)
    minw, maxw = float(np.abs(uvw_l[:, 2]).min()), float(np.abs(uvw_l[:, 2]).max())
    dev = torch.device("cuda:0")
    d_uvw, d_wl = torch.from_numpy(uvw).to(dev), torch.from_numpy(wl).to(dev)
    facets = _uneven_facets(dl)
    fb = FacetBatch(facets, dl, dl, d_uvw, d_wl, n_w_layers=12, w_limits=(minw, maxw))
    re = fb.images()
    graph = fb.capture_invert(torch.from_numpy(vis).to(dev), 1.0, re)
    graph.launch()
    torch.cuda.synchronize()
    for f, img in zip(facets, re):
        ph = -2 * np.pi * (uvw_l[:, 0] * f.l_shift + uvw_l[:, 1] * f.m_shift)
        rot = (vis.reshape(-1).astype(np.complex128) * np.exp(1j * ph)).astype(np.complex64)
        g = oracle.RefGridder(f.width, f.height, dl, dl, 2, 7, 63, "float")
        g.set_denormal_phase_centre(f.l_shift, f.m_shift)
        g.prepare_w_layers(12, 1e12, minw, maxw)
        g.start_inversion_pass(0)
        g.add_samples(uvw_l, rot)
        g.finish_inversion_pass()
        g.finalize_image(1.0)
        err = _rel(img[0].cpu().numpy(), g.image())
        print(f"facet {f.width}x{f.height} shift ({f.l_shift}, {f.m_shift}): rel {err:.2e}")
        assert err < 3e-5
    graph.close()
    fb.close()


def test_facet_decomposition_matches_direct_transform():
    """A 3 x 3 facet grid over a 288^2 field, each 96^2 box gridded at 160^2
    (padding 1.5), 32 w-layers; two point sources in every box. The error
    bounds are those of the oracle's FP64 w-stacking on the same facets
    (predict 1.1e-2 relative RMS, invert 1.9e-3 RMS of the peak: the kernel's
    accuracy across the box at this padding), with a margin; a wrong shift or
    rotation sign gives errors of order one."""
    from wsclean_amd import direct
    from wsclean_amd.facets import FacetBatch, cut_facet, facets_of_grid, paste_facet
    dl = 0.0005
    W = 288
    uvw, wl, _ = _samples(34, dl, nrow=500, nchan=2, wmax_m=40.0)
    dev = torch.device("cuda:0")
    d_uvw, d_wl = torch.from_numpy(uvw).to(dev), torch.from_numpy(wl).to(dev)
    facets = facets_of_grid(W, W, 3, 3, dl, dl, padding=1.5)
    assert all(f.width == 160 and f.height == 160 for f in facets)
    rng = np.random.default_rng(35)
    model = np.zeros((W, W), np.float32)
    for f in facets:
        x0, y0, x1, y1 = f.box
        for _ in range(2):
            model[rng.integers(y0 + 8, y1 - 8), rng.integers(x0 + 8, x1 - 8)] = \
                rng.uniform(0.5, 2.0)
    d_model = torch.from_numpy(model).to(dev)
    fb = FacetBatch(facets, dl, dl, d_uvw, d_wl, n_w_layers=32)
    # predict: the facet models' visibilities summed over the facets
    sub = [cut_facet(d_model, f.box, f.width, f.height)[None].contiguous() for f in facets]
    outs = [torch.empty((len(uvw), len(wl), 1), dtype=torch.complex64, device=dev)
            for _ in facets]
    gp = fb.capture_predict(sub, None, outs)
    gp.launch()
    torch.cuda.synchronize()
    total = sum(o.to(torch.complex128) for o in outs)[..., 0]
    ref = direct.direct_predict(d_model, d_uvw, d_wl, dl, dl)
    err = _rel(total.cpu().numpy(), ref.cpu().numpy())
    print(f"sum of facet predicts vs direct predict: rel rms {err:.2e}")
    assert err < 2e-2
    # invert the direct visibilities of the model: pasted facet dirty images
    # against the direct invert of the whole field
    d_vis = ref[..., None].contiguous()
    re = fb.images()
    gi = fb.capture_invert(d_vis, 1.0, re)
    gi.launch()
    torch.cuda.synchronize()
    pasted = torch.zeros((W, W), dtype=torch.float32, device=dev)
    for f, img in zip(facets, re):
        paste_facet(pasted, img[0], f.box)
    dref = direct.direct_invert(d_uvw, d_wl, d_vis, W, W, dl, dl).cpu().numpy()
    err = np.sqrt(np.mean((pasted.cpu().numpy() - dref) ** 2)) / np.abs(dref).max()
    print(f"pasted facet images vs direct invert: rms / peak {err:.2e}")
    assert err < 4e-3
    gp.close()
    gi.close()
    fb.close()
