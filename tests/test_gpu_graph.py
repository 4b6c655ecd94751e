"""HIP graph capture of the device-resident invert and predict
(hws_graph_capture_*): every replay must reproduce the eager result
bit for bit (the batched path sums in a fixed order), with the program's own
eager work and synchronisations between the replays."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

C_LIGHT = 299792458.0


def _case(w, nrow=600, nchan=4, seed=21):
    rng = np.random.default_rng(seed)
    dl = 0.003
    wl = C_LIGHT / np.linspace(120e6, 150e6, nchan)
    umax = 0.42 / dl * wl.min()
    uvw = np.column_stack([rng.uniform(-umax, umax, nrow), rng.uniform(-umax, umax, nrow),
                           rng.uniform(-60, 60, nrow)])
    vis = (rng.standard_normal((nrow, nchan, 2)) + 1j * rng.standard_normal((nrow, nchan, 2))
           ).astype(np.complex64)
    maxw = float(np.abs(uvw[:, 2]).max() / wl.min())
    return uvw, wl, vis, dl, maxw


def _rel(a, b):
    return float(torch.linalg.vector_norm((a - b).double()) /
                 torch.linalg.vector_norm(b.double()))


@pytest.mark.parametrize("precision", ["mixed", "reference"])
def test_graph_replays_match_eager(precision):
    import wsclean_amd
    w = 256
    uvw, wl, vis, dl, maxw = _case(w)
    dev = torch.device("cuda:0")
    g = wsclean_amd.DeviceGridder(w, w, dl, dl, 7, 63, precision=precision, n_images=2,
                                  complex_images=(1,), work_chunk=64)
    g.prepare_w_layers(8, 0.0, maxw)
    d_uvw, d_wl = torch.from_numpy(uvw).to(dev), torch.from_numpy(wl).to(dev)
    d_vis = torch.from_numpy(vis).to(dev)
    g.plan(d_uvw, d_wl)
    re = torch.zeros((2, w, w), dtype=torch.float32, device=dev)
    im = torch.zeros_like(re)
    g.invert(d_vis, 0.5, re, im)
    out = torch.empty_like(d_vis)
    g.predict(re, im, out)
    torch.cuda.synchronize()
    re_e, im_e, out_e = re.clone(), im.clone(), out.clone()

    gre, gim = torch.full_like(re, 7.0), torch.full_like(im, 7.0)
    gout = torch.full_like(out, 3.0)
    ginv = g.capture_invert(d_vis, 0.5, gre, gim)
    gpred = g.capture_predict(re_e, im_e, gout)
    scratch = torch.empty_like(out)
    for k in range(3):
        gre.fill_(-1.0)
        gim.fill_(-1.0)
        gout.fill_(-1.0)
        ginv.launch()
        gpred.launch()
        torch.cuda.synchronize()
        assert torch.equal(gre, re_e), k
        assert torch.equal(gim[1], im_e[1]), k
        assert torch.equal(torch.isnan(gout), torch.isnan(out_e)), k
        ok = ~torch.isnan(out_e)
        assert torch.equal(gout[ok], out_e[ok]), k
        # eager work of the program between replays
        g.predict(re_e, im_e, scratch)
        torch.cuda.synchronize()
    ginv.close()
    gpred.close()
    g.close()


def test_graph_capture_rejects_multipass_and_plain():
    import wsclean_amd
    from wsclean_amd._lib import HwsError
    w = 64
    uvw, wl, vis, dl, maxw = _case(w, nrow=50)
    dev = torch.device("cuda:0")
    d_uvw, d_wl = torch.from_numpy(uvw).to(dev), torch.from_numpy(wl).to(dev)
    d_vis = torch.from_numpy(vis[:, :, :1].copy()).to(dev)
    img = torch.zeros((1, w, w), dtype=torch.float32, device=dev)
    g = wsclean_amd.DeviceGridder(w, w, dl, dl, 7, 63, precision="mixed")
    assert g.prepare_w_layers(8, 0.0, maxw, max_mem=w * w * 8 * 3) > 1
    g.plan(d_uvw, d_wl)
    with pytest.raises(HwsError, match="single pass"):
        g.capture_invert(d_vis, 1.0, img)
    g.close()
    p = wsclean_amd.DeviceGridder(w, w, dl, dl, 7, 63, precision="plain")
    p.prepare_w_layers(8, 0.0, maxw)
    p.plan(d_uvw, d_wl)
    with pytest.raises(HwsError, match="batched"):
        p.capture_invert(d_vis, 1.0, img)
    p.close()


def test_stale_graphs_and_replaced_plans_are_refused():
    """A graph bakes in the plan, the layer buffers and the weighting of its
    context: after any of them changes, or the context is destroyed, a
    launch is refused instead of replaying stale pointers. A plan made by
    the host protocol (hws_add_data) replaces the device plan, and the
    planned calls refuse it until the next plan()."""
    import ctypes as C
    import wsclean_amd
    from wsclean_amd._lib import HwsError
    from wsclean_amd.weights import DeviceImageWeights
    w = 64
    uvw, wl, vis, dl, maxw = _case(w, nrow=80)
    dev = torch.device("cuda:0")
    d_uvw, d_wl = torch.from_numpy(uvw).to(dev), torch.from_numpy(wl).to(dev)
    d_vis = torch.from_numpy(vis[:, :, :1].copy()).to(dev)
    img = torch.zeros((1, w, w), dtype=torch.float32, device=dev)
    g = wsclean_amd.DeviceGridder(w, w, dl, dl, 7, 63, precision="mixed")
    g.prepare_w_layers(8, 0.0, maxw)
    g.plan(d_uvw, d_wl)
    ref = torch.zeros_like(img)
    g.invert(d_vis, 1.0, ref)
    torch.cuda.synchronize()

    # re-planned: stale; captured again: replays the eager result
    gr = g.capture_invert(d_vis, 1.0, img)
    gr.launch()
    torch.cuda.synchronize()
    assert torch.equal(img, ref)
    g.plan(d_uvw, d_wl)
    with pytest.raises(HwsError, match="capture again"):
        gr.launch()
    gr.close()
    gr = g.capture_invert(d_vis, 1.0, img)
    img.zero_()
    gr.launch()
    torch.cuda.synchronize()
    # a new plan may hold a bin's records in another order (the binning
    # scatter places them with atomics), which changes the FP32 sums at
    # rounding level: equal to the eager image within that
    assert _rel(img, ref) < 1e-6

    # another weighting: stale
    iw = DeviceImageWeights("uniform", w, w, dl, dl)
    iw.grid(d_uvw, d_wl)
    iw.finish()
    g.set_weighting(iw)
    with pytest.raises(HwsError, match="capture again"):
        gr.launch()
    gr.close()
    g.set_weighting(None)

    # re-prepared: the plan is gone and the graph stale
    gr = g.capture_invert(d_vis, 1.0, img)
    g.prepare_w_layers(8, 0.0, maxw)
    with pytest.raises(HwsError, match="capture again"):
        gr.launch()
    with pytest.raises(HwsError, match="hws_plan_device has not been called"):
        g.invert(d_vis, 1.0, img)
    gr.close()

    # the host protocol replaces the device plan
    g.plan(d_uvw, d_wl)
    g._check(g.L.hws_start_inversion_pass(g.h, 0))
    host_uvw = np.ascontiguousarray(uvw[:10])
    host_vis = np.ascontiguousarray(vis[:10, :, :1])
    g._check(g.L.hws_add_data(g.h, 10, len(wl), host_uvw.ctypes.data_as(C.POINTER(C.c_double)),
                              wl.ctypes.data_as(C.POINTER(C.c_double)),
                              host_vis.ctypes.data_as(C.POINTER(C.c_float))))
    g._check(g.L.hws_finish_inversion_pass(g.h))
    with pytest.raises(HwsError, match="replaced by hws_add_data"):
        g.invert(d_vis, 1.0, img)
    with pytest.raises(HwsError, match="replaced by hws_add_data"):
        g.capture_invert(d_vis, 1.0, img)
    g.plan(d_uvw, d_wl)
    img.zero_()
    g.invert(d_vis, 1.0, img)
    torch.cuda.synchronize()
    assert _rel(img, ref) < 1e-6

    # a destroyed context: its graphs refuse to launch
    gr = g.capture_invert(d_vis, 1.0, img)
    g.close()
    with pytest.raises(HwsError, match="destroyed"):
        gr.launch()
    gr.close()
    iw.close()
