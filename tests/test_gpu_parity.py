"""Parity of the HIP w-stacking core (C ABI) with the oracle: the CPU
restatement of the reference WStackingGridder<float|double> on the same
seeded inputs, and the FP64 direct FT.

Tolerances (relative L2 unless stated), set from the oracle's own
float-vs-double spread (~1e-6) and the precision of each form:
  layers:  reference/mixed 2e-6; plain 2e-3 (FP32 sample positions)
  images:  reference 1e-5, mixed 3e-5, plain 3e-3
  predict: reference 1e-5, mixed 3e-5, plain 3e-3
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL_IMG = {"reference": 1e-5, "mixed": 3e-5, "plain": 3e-3}
TOL_LAYER = {"reference": 2e-6, "mixed": 2e-6, "plain": 2e-3}
C_LIGHT = 299792458.0


def rel(a, b):
    a = np.asarray(a, np.complex128)
    b = np.asarray(b, np.complex128)
    return np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel())


@pytest.fixture(scope="module")
def hws():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    import wsclean_amd
    return wsclean_amd


def case(seed, width, nrow=300, nchan=3, wmax_m=60.0, frac=0.42, dl=0.004, nimg=1):
    """Random rows (metres) and channels whose uv cover `frac` of the grid."""
    rng = np.random.default_rng(seed)
    freqs = np.linspace(120e6, 150e6, nchan)
    wl = C_LIGHT / freqs
    umax = frac / dl * wl.min()
    uvw = np.column_stack([rng.uniform(-umax, umax, nrow), rng.uniform(-umax, umax, nrow),
                           rng.uniform(-wmax_m, wmax_m, nrow)])
    vis = (rng.standard_normal((nrow, nchan, nimg)) +
           1j * rng.standard_normal((nrow, nchan, nimg))).astype(np.complex64)
    return uvw, wl, vis, dl


def lambda_samples(uvw, wl):
    return (uvw[:, None, :] / wl[None, :, None]).reshape(-1, 3)


def oracle_invert(oracle, w, h, dl, dm, uvw_l, vis, nw, minw, maxw, cplx=False, mode="kb",
                  prec="float", shift=(0.0, 0.0), K=7, OS=63):
    g = oracle.RefGridder(w, h, dl, dm, 2, K, OS, prec)
    if mode != "kb":
        g.set_grid_mode(mode)
    g.set_is_complex(cplx)
    if shift != (0.0, 0.0):
        g.set_denormal_phase_centre(*shift)
    g.prepare_w_layers(nw, 1e12, minw, maxw)
    g.start_inversion_pass(0)
    g.add_samples(uvw_l, vis)
    layers = [g.layer(i) for i in range(nw)]
    g.finish_inversion_pass()
    g.finalize_image(1.0)
    return g.image(imag=cplx), layers


def oracle_predict(oracle, w, h, dl, dm, uvw_l, real, imag, nw, minw, maxw, cplx=False,
                   mode="kb", prec="float", shift=(0.0, 0.0), K=7, OS=63):
    g = oracle.RefGridder(w, h, dl, dm, 2, K, OS, prec)
    if mode != "kb":
        g.set_grid_mode(mode)
    g.set_is_complex(cplx)
    if shift != (0.0, 0.0):
        g.set_denormal_phase_centre(*shift)
    g.prepare_w_layers(nw, 1e12, minw, maxw)
    g.initialize_prediction(real, imag)
    g.start_prediction_pass(0)
    return g.sample(uvw_l)


def gpu_invert(hws, w, h, dl, dm, uvw, wl, vis, nw, minw, maxw, precision, complex_images=(),
               mode="kb", layers=False, max_mem=0.0, K=7, OS=63, **kw):
    nimg = vis.shape[2]
    g = hws.WStackingGridder(w, h, dl, dm, K, OS, kernel_mode=mode, precision=precision,
                             n_images=nimg, complex_images=complex_images, **kw)
    npass = g.prepare_w_layers(nw, max_mem, minw, maxw)
    lay = None
    for p in range(npass):
        g.start_inversion_pass(p)
        g.add_data(uvw, wl, vis)
        if layers and p == 0:
            lay = [[g.get_gridded_uv_layer(i, image=k) for i in range(nw)] for k in range(nimg)]
        g.finish_inversion_pass()
    re, im = g.finalize_image(1.0)
    return re, im, lay, npass


@pytest.mark.parametrize("precision", ["reference", "mixed", "plain"])
def test_invert_layers_and_image_match_oracle(hws, oracle, precision):
    w = 64
    uvw, wl, vis, dl = case(1, w)
    uvw_l = lambda_samples(uvw, wl)
    maxw = np.abs(uvw_l[:, 2]).max()
    nw = 8
    (img_o, lay_o) = oracle_invert(oracle, w, w, dl, dl, uvw_l, vis.reshape(-1), nw, 0.0, maxw)
    re, _, lay, _ = gpu_invert(hws, w, w, dl, dl, uvw, wl, vis, nw, 0.0, maxw, precision,
                               layers=True)
    for i in range(nw):
        if np.abs(lay_o[i]).sum() == 0:
            assert np.abs(lay[0][i]).sum() == 0
            continue
        assert rel(lay[0][i], lay_o[i]) < TOL_LAYER[precision], (precision, i)
    e = rel(re[0], img_o)
    print(f"invert {precision}: image rel L2 vs oracle float = {e:.2e}")
    assert e < TOL_IMG[precision]


@pytest.mark.parametrize("precision", ["reference", "mixed", "plain"])
def test_predict_matches_oracle(hws, oracle, precision):
    w = 64
    uvw, wl, _, dl = case(2, w)
    uvw_l = lambda_samples(uvw, wl)
    maxw = np.abs(uvw_l[:, 2]).max()
    rng = np.random.default_rng(22)
    model = rng.standard_normal((w, w)).astype(np.float32)
    v_o = oracle_predict(oracle, w, w, dl, dl, uvw_l, model, None, 8, 0.0, maxw)
    g = hws.WStackingGridder(w, w, dl, dl, 7, 63, precision=precision)
    g.prepare_w_layers(8, 0.0, 0.0, maxw)
    g.initialize_prediction(model)
    g.start_prediction_pass(0)
    v = g.sample_data(uvw, wl).reshape(-1)
    assert np.array_equal(np.isnan(v), np.isnan(v_o))
    ok = ~np.isnan(v_o)
    e = rel(v[ok], v_o[ok])
    print(f"predict {precision}: rel L2 vs oracle float = {e:.2e}")
    assert e < TOL_IMG[precision]


def test_complex_and_real_images_together(hws, oracle):
    # XX (real) and XY (complex) sharing uvw, gridded in one context
    w = 64
    uvw, wl, vis, dl = case(3, w, nimg=2)
    uvw_l = lambda_samples(uvw, wl)
    maxw = np.abs(uvw_l[:, 2]).max()
    re, im, _, _ = gpu_invert(hws, w, w, dl, dl, uvw, wl, vis, 8, 0.0, maxw, "reference",
                              complex_images=(1,))
    img0, _ = oracle_invert(oracle, w, w, dl, dl, uvw_l, vis[:, :, 0].reshape(-1), 8, 0.0, maxw)
    (img1r, img1i), _ = oracle_invert(oracle, w, w, dl, dl, uvw_l, vis[:, :, 1].reshape(-1), 8,
                                      0.0, maxw, cplx=True)
    assert rel(re[0], img0) < TOL_IMG["reference"]
    assert rel(re[1], img1r) < TOL_IMG["reference"]
    assert rel(im[1], img1i) < TOL_IMG["reference"]
    # prediction of both
    rng = np.random.default_rng(33)
    mre = rng.standard_normal((2, w, w)).astype(np.float32)
    mim = rng.standard_normal((2, w, w)).astype(np.float32)
    g = hws.WStackingGridder(w, w, dl, dl, 7, 63, precision="reference", n_images=2,
                             complex_images=(1,))
    g.prepare_w_layers(8, 0.0, 0.0, maxw)
    g.initialize_prediction(mre, mim)
    g.start_prediction_pass(0)
    v = g.sample_data(uvw, wl)
    v0 = oracle_predict(oracle, w, w, dl, dl, uvw_l, mre[0], None, 8, 0.0, maxw)
    v1 = oracle_predict(oracle, w, w, dl, dl, uvw_l, mre[1], mim[1], 8, 0.0, maxw, cplx=True)
    assert rel(v[:, :, 0].reshape(-1), v0) < TOL_IMG["reference"]
    assert rel(v[:, :, 1].reshape(-1), v1) < TOL_IMG["reference"]


def test_four_images_beyond_lds_split_per_launch(hws, oracle):
    # 56-pixel uv tiles with a 9-tap kernel give 64 x 64 degridding subgrids
    # (32 KiB per image); beside the 1023-times oversampled kernel table
    # (36.8 KB) four images need 172 KB of LDS, more than a workgroup's
    # 160 KiB, so the degridder runs them as three and one (and the gridder
    # splits its private tiles the same way)
    w, K, OS = 224, 9, 1023
    uvw, wl, vis, dl = case(12, w, nimg=4)
    uvw_l = lambda_samples(uvw, wl)
    maxw = np.abs(uvw_l[:, 2]).max()
    re, _, _, _ = gpu_invert(hws, w, w, dl, dl, uvw, wl, vis, 8, 0.0, maxw, "reference",
                             K=K, OS=OS, tile_size=56)
    rng = np.random.default_rng(121)
    models = rng.standard_normal((4, w, w)).astype(np.float32)
    g = hws.WStackingGridder(w, w, dl, dl, K, OS, precision="reference", n_images=4,
                             tile_size=56)
    g.prepare_w_layers(8, 0.0, 0.0, maxw)
    g.initialize_prediction(models)
    g.start_prediction_pass(0)
    v = g.sample_data(uvw, wl)
    for k in range(4):
        img_o, _ = oracle_invert(oracle, w, w, dl, dl, uvw_l, vis[:, :, k].reshape(-1), 8, 0.0,
                                 maxw, K=K, OS=OS)
        ei = rel(re[k], img_o)
        v_o = oracle_predict(oracle, w, w, dl, dl, uvw_l, models[k], None, 8, 0.0, maxw, K=K,
                             OS=OS)
        vk = v[:, :, k].reshape(-1)
        assert np.array_equal(np.isnan(vk), np.isnan(v_o))
        ok = ~np.isnan(v_o)
        ep = rel(vk[ok], v_o[ok])
        print(f"K={K} OS={OS} T=56 image {k}: invert rel {ei:.2e}, predict rel {ep:.2e}")
        assert ei < TOL_IMG["reference"] and ep < TOL_IMG["reference"]


def test_nearest_neighbour_mode(hws, oracle):
    w = 64
    uvw, wl, vis, dl = case(4, w)
    uvw_l = lambda_samples(uvw, wl)
    maxw = np.abs(uvw_l[:, 2]).max()
    img_o, lay_o = oracle_invert(oracle, w, w, dl, dl, uvw_l, vis.reshape(-1), 4, 0.0, maxw,
                                 mode="nn")
    re, _, lay, _ = gpu_invert(hws, w, w, dl, dl, uvw, wl, vis, 4, 0.0, maxw, "reference",
                               mode="nn", layers=True)
    for i in range(4):
        np.testing.assert_allclose(lay[0][i], lay_o[i], rtol=1e-6, atol=1e-5)
    assert rel(re[0], img_o) < TOL_IMG["reference"]


def test_shifted_phase_centre_with_rotation(hws, oracle):
    # WSMSGridder rotates samples by exp(-2 pi i (u dl + v dm)/lambda) before
    # gridding (gridding/wsmsgridder.cpp:185-190) and n' uses the shifted l, m
    w = 64
    uvw, wl, vis, dl = case(5, w)
    shift = (0.013, -0.008)
    uvw_l = lambda_samples(uvw, wl)
    maxw = np.abs(uvw_l[:, 2]).max()
    ph = -2 * np.pi * (uvw_l[:, 0] * shift[0] + uvw_l[:, 1] * shift[1])
    rot = (vis.reshape(-1).astype(np.complex128) * np.exp(1j * ph)).astype(np.complex64)
    img_o, _ = oracle_invert(oracle, w, w, dl, dl, uvw_l, rot, 8, 0.0, maxw, shift=shift)
    re, _, _, _ = gpu_invert(hws, w, w, dl, dl, uvw, wl, vis, 8, 0.0, maxw, "reference",
                             l_shift=shift[0], m_shift=shift[1], rotate_visibilities=True)
    assert rel(re[0], img_o) < TOL_IMG["reference"]


@pytest.mark.parametrize("precision", ["reference", "mixed"])
def test_shifted_phase_centre_predict_with_rotation(hws, oracle, precision):
    # WSMSGridder rotates predicted samples by exp(+2 pi i (u dl + v dm)/lambda)
    # after sampling (gridding/wsmsgridder.cpp:326-332); the sampling n'
    # table uses the shifted l, m (gridding/wstackinggridder.cpp:909-932)
    w = 64
    uvw, wl, _, dl = case(7, w)
    shift = (-0.011, 0.017)
    uvw_l = lambda_samples(uvw, wl)
    maxw = np.abs(uvw_l[:, 2]).max()
    model = np.random.default_rng(23).standard_normal((w, w)).astype(np.float32)
    v_o = oracle_predict(oracle, w, w, dl, dl, uvw_l, model, None, 8, 0.0, maxw, shift=shift)
    ph = 2 * np.pi * (uvw_l[:, 0] * shift[0] + uvw_l[:, 1] * shift[1])
    v_o = v_o * np.exp(1j * ph)
    g = hws.WStackingGridder(w, w, dl, dl, 7, 63, precision=precision, l_shift=shift[0],
                             m_shift=shift[1], rotate_visibilities=True)
    g.prepare_w_layers(8, 0.0, 0.0, maxw)
    g.initialize_prediction(model)
    g.start_prediction_pass(0)
    v = g.sample_data(uvw, wl).reshape(-1)
    assert np.array_equal(np.isnan(v), np.isnan(v_o))
    ok = ~np.isnan(v_o)
    e = rel(v[ok], v_o[ok])
    print(f"shifted predict {precision}: rel L2 vs oracle float = {e:.2e}")
    assert e < TOL_IMG[precision]


def test_multipass_equals_single_pass(hws):
    w = 64
    uvw, wl, vis, dl = case(6, w)
    maxw = np.abs(lambda_samples(uvw, wl)[:, 2]).max()
    one, _, _, n1 = gpu_invert(hws, w, w, dl, dl, uvw, wl, vis, 12, 0.0, maxw, "reference")
    many, _, _, n2 = gpu_invert(hws, w, w, dl, dl, uvw, wl, vis, 12, 0.0, maxw, "reference",
                                max_mem=w * w * 8 * 5)
    assert n1 == 1 and n2 == 3
    assert rel(many, one) < 1e-6


def test_nan_contract_off_grid_and_other_pass(hws):
    w = 32
    dl = 0.01
    g = hws.WStackingGridder(w, w, dl, dl, 7, 63, precision="reference")
    npass = g.prepare_w_layers(8, w * w * 8 * 4, 1.0, 100.0)
    assert npass == 2
    g.initialize_prediction(np.ones((w, w), np.float32))
    uvw = np.array([[0.0, 0.0, 1.0], [0.0, 0.0, 99.0], [1e6, 0.0, 1.0]])
    g.start_prediction_pass(0)
    v0 = g.sample_samples(uvw)[:, 0]
    assert np.isfinite(v0[0]) and np.isnan(v0[1]) and np.isnan(v0[2])
    g.start_prediction_pass(1)
    v1 = g.sample_samples(uvw)[:, 0]
    assert np.isnan(v1[0]) and np.isfinite(v1[1]) and np.isnan(v1[2])


@pytest.mark.parametrize("shape,precision", [((96, 96), "reference"), ((80, 96), "reference"),
                                             ((144, 112), "mixed"), ((90, 90), "plain")])
def test_non_power_of_two_sizes(hws, oracle, shape, precision):
    # multiples of 4: the blocked layout with the runtime-radix transforms;
    # 90 x 90 (plain form only: no tile edge divides it) the row-major layout
    w, h = shape
    uvw, wl, vis, dl = case(7, min(w, h))
    uvw_l = lambda_samples(uvw, wl)
    maxw = np.abs(uvw_l[:, 2]).max()
    img_o, _ = oracle_invert(oracle, w, h, dl, dl, uvw_l, vis.reshape(-1), 6, 0.0, maxw)
    re, _, _, _ = gpu_invert(hws, w, h, dl, dl, uvw, wl, vis, 6, 0.0, maxw, precision)
    assert rel(re[0], img_o) < TOL_IMG[precision]
    # predict through the same transforms
    model = np.random.default_rng(71).standard_normal((h, w)).astype(np.float32)
    v_o = oracle_predict(oracle, w, h, dl, dl, uvw_l, model, None, 6, 0.0, maxw)
    g = hws.WStackingGridder(w, h, dl, dl, 7, 63, precision=precision)
    g.prepare_w_layers(6, 0.0, 0.0, maxw)
    g.initialize_prediction(model)
    g.start_prediction_pass(0)
    v = g.sample_data(uvw, wl).reshape(-1)
    ok = ~np.isnan(v_o)
    assert np.array_equal(np.isnan(v), ~ok)
    assert rel(v[ok], v_o[ok]) < TOL_IMG[precision]


def test_device_forms_match_host_protocol(hws):
    import torch
    w = 64
    uvw, wl, vis, dl = case(8, w, nimg=4)
    maxw = np.abs(lambda_samples(uvw, wl)[:, 2]).max()
    cimg = (1, 2)
    re_h, im_h, _, _ = gpu_invert(hws, w, w, dl, dl, uvw, wl, vis, 8, 0.0, maxw, "mixed",
                                  complex_images=cimg)
    dev = torch.device("cuda:0")
    g = hws.DeviceGridder(w, w, dl, dl, 7, 63, precision="mixed", n_images=4,
                          complex_images=cimg)
    g.prepare_w_layers(8, 0.0, maxw)
    d_uvw = torch.from_numpy(uvw).to(dev)
    d_wl = torch.from_numpy(wl).to(dev)
    d_vis = torch.from_numpy(vis).to(dev)
    re = torch.zeros((4, w, w), dtype=torch.float32, device=dev)
    im = torch.zeros_like(re)
    g.plan(d_uvw, d_wl)
    g.invert(d_vis, 1.0, re, im)
    g.synchronize()
    assert rel(re.cpu().numpy(), re_h) < 1e-6
    assert rel(im.cpu().numpy()[1:3], im_h[1:3]) < 1e-6
    out = torch.empty_like(d_vis)
    g.predict(re, im, out)
    g.synchronize()
    h = hws.WStackingGridder(w, w, dl, dl, 7, 63, precision="mixed", n_images=4,
                             complex_images=cimg)
    h.prepare_w_layers(8, 0.0, 0.0, maxw)
    h.initialize_prediction(re.cpu().numpy(), im.cpu().numpy())
    h.start_prediction_pass(0)
    v_h = h.sample_data(uvw, wl)
    v_d = out.cpu().numpy()
    assert np.array_equal(np.isnan(v_h), np.isnan(v_d))
    ok = ~np.isnan(v_h)
    assert rel(v_d[ok], v_h[ok]) < 1e-6
    st = g.stats()
    assert st["samples_in"] == uvw.shape[0] * wl.shape[0]


@pytest.mark.parametrize("precision", ["reference", "mixed"])
def test_gpu_predict_is_adjoint_of_invert(hws, precision):
    w = 64
    uvw, wl, vis, dl = case(9, w, nimg=2)
    maxw = np.abs(lambda_samples(uvw, wl)[:, 2]).max()
    re, im, _, _ = gpu_invert(hws, w, w, dl, dl, uvw, wl, vis, 8, 0.0, maxw, precision,
                              complex_images=(1,))
    rng = np.random.default_rng(99)
    mre = rng.standard_normal((2, w, w)).astype(np.float32)
    mim = rng.standard_normal((2, w, w)).astype(np.float32)
    g = hws.WStackingGridder(w, w, dl, dl, 7, 63, precision=precision, n_images=2,
                             complex_images=(1,))
    g.prepare_w_layers(8, 0.0, 0.0, maxw)
    g.initialize_prediction(mre, mim)
    g.start_prediction_pass(0)
    a = g.sample_data(uvw, wl)
    for k in range(2):
        lhs = np.sum(re[k].astype(np.float64) * mre[k])
        if k == 1:
            lhs += np.sum(im[k].astype(np.float64) * mim[k])
        rhs = np.real(np.vdot(vis[:, :, k].astype(np.complex128).ravel(),
                              a[:, :, k].astype(np.complex128).ravel()))
        assert abs(lhs - rhs) < 1e-5 * abs(lhs), (k, lhs, rhs)


@pytest.mark.parametrize("shape,precision", [((256, 256), "reference"), ((512, 256), "mixed"),
                                             ((256, 256), "plain")])
def test_blocked_layer_layout_sizes(hws, oracle, shape, precision):
    # power-of-two sizes 256..4096 store layers with four rows interleaved and
    # use the coalesced radix-16 transforms; layers read back in natural order
    w, h = shape
    uvw, wl, vis, dl = case(11, min(w, h), nrow=400, dl=0.003)
    uvw_l = lambda_samples(uvw, wl)
    maxw = np.abs(uvw_l[:, 2]).max()
    img_o, lay_o = oracle_invert(oracle, w, h, dl, dl, uvw_l, vis.reshape(-1), 4, 0.0, maxw)
    re, _, lay, _ = gpu_invert(hws, w, h, dl, dl, uvw, wl, vis, 4, 0.0, maxw, precision,
                               layers=True)
    for i in range(4):
        if np.abs(lay_o[i]).sum() > 0:
            assert rel(lay[0][i], lay_o[i]) < TOL_LAYER[precision]
    assert rel(re[0], img_o) < TOL_IMG[precision]
    rng = np.random.default_rng(5)
    model = rng.standard_normal((h, w)).astype(np.float32)
    v_o = oracle_predict(oracle, w, h, dl, dl, uvw_l, model, None, 4, 0.0, maxw)
    g = hws.WStackingGridder(w, h, dl, dl, 7, 63, precision=precision)
    g.prepare_w_layers(4, 0.0, 0.0, maxw)
    g.initialize_prediction(model)
    g.start_prediction_pass(0)
    v = g.sample_data(uvw, wl).reshape(-1)
    ok = ~np.isnan(v_o)
    assert np.array_equal(np.isnan(v), ~ok)
    assert rel(v[ok], v_o[ok]) < TOL_IMG[precision]


@pytest.mark.parametrize("precision", ["reference", "mixed"])
def test_split_bins_match_oracle(hws, oracle, precision):
    # a dense uv centre: every sample falls in a few (layer, tile) bins, and a
    # work chunk of 37 records splits each of them into many work items whose
    # tiles are summed afterwards (grid) or which gather separately (degrid)
    import torch
    w = 128
    uvw, wl, vis, dl = case(12, w, nrow=1500, nchan=4, frac=0.06, nimg=2)
    uvw_l = lambda_samples(uvw, wl)
    maxw = np.abs(uvw_l[:, 2]).max()
    nw = 4
    img_o, lay_o = oracle_invert(oracle, w, w, dl, dl, uvw_l, vis[:, :, 0].reshape(-1), nw, 0.0,
                                 maxw)
    # host protocol, two add_data batches: the accumulate path
    g = hws.WStackingGridder(w, w, dl, dl, 7, 63, precision=precision, n_images=2,
                             work_chunk=37)
    g.prepare_w_layers(nw, 0.0, 0.0, maxw)
    g.start_inversion_pass(0)
    g.add_data(uvw[:700], wl, vis[:700])
    g.add_data(uvw[700:], wl, vis[700:])
    lay = [g.get_gridded_uv_layer(i, image=0) for i in range(nw)]
    g.finish_inversion_pass()
    re, _ = g.finalize_image(1.0)
    for i in range(nw):
        if np.abs(lay_o[i]).sum() > 0:
            assert rel(lay[i], lay_o[i]) < TOL_LAYER[precision], i
    assert rel(re[0], img_o) < TOL_IMG[precision]
    # device form: the tile-writing path; both images, then predict
    dev = torch.device("cuda:0")
    d = hws.DeviceGridder(w, w, dl, dl, 7, 63, precision=precision, n_images=2, work_chunk=37)
    d.prepare_w_layers(nw, 0.0, maxw)
    d_uvw, d_wl = torch.from_numpy(uvw).to(dev), torch.from_numpy(wl).to(dev)
    d_vis = torch.from_numpy(vis).to(dev)
    out = torch.zeros((2, w, w), dtype=torch.float32, device=dev)
    d.plan(d_uvw, d_wl)
    d.invert(d_vis, 1.0, out)
    st = d.stats()
    assert st["chunks"] > st["bins_used"] + 50  # large bins were split into extra items
    # tile counters against the oracle's gridded layers: a (layer, uv tile)
    # bin holds records exactly when the oracle grids a non-zero cell into
    # that tile (random visibilities, kernel taps all non-zero)
    tiles_per_layer = st["tiles_total"] // nw
    t_edge = w // int(round(np.sqrt(tiles_per_layer)))
    nz = np.stack([np.abs(np.asarray(lo)) > 0 for lo in lay_o])  # [layer, y, x]
    nt = w // t_edge
    tiles = nz.reshape(nw, nt, t_edge, nt, t_edge).any(axis=(2, 4))  # [layer, ty, tx]
    assert st["tiles_occupied"] == [int(tiles.sum()), 0]
    assert st["tile_rows_total"] == nw * nt
    assert st["tile_rows_live"] == [int(tiles.any(axis=2).sum()), 0]
    # needed tiles: a tile is read by the subgrids of its own bin and of the
    # bins to its left, above and above-left (the subgrid reaches K - 1 cells
    # into the next tiles, wrapping round the grid)
    need = (tiles | np.roll(tiles, 1, axis=2) | np.roll(tiles, 1, axis=1) |
            np.roll(np.roll(tiles, 1, axis=1), 1, axis=2))
    assert st["tiles_needed"] == [int(need.sum()), 0]
    assert st["tile_cols_total"] == nw * nt
    assert st["tile_cols_needed"] == [int(need.any(axis=1).sum()), 0]
    assert st["tiles_occupied"][0] <= st["tiles_needed"][0] <= st["tiles_total"]
    assert rel(out.cpu().numpy(), re) < 1e-6
    rng = np.random.default_rng(3)
    model = rng.standard_normal((2, w, w)).astype(np.float32)
    pred = torch.empty_like(d_vis)
    d.predict(torch.from_numpy(model).to(dev), None, pred)
    v_o = oracle_predict(oracle, w, w, dl, dl, uvw_l, model[1], None, nw, 0.0, maxw)
    v = pred.cpu().numpy()[:, :, 1].reshape(-1)
    ok = ~np.isnan(v_o)
    assert np.array_equal(np.isnan(v), ~ok)
    assert rel(v[ok], v_o[ok]) < TOL_IMG[precision]


@pytest.mark.parametrize("tile", [8, 16, 32, 24])
def test_tile_sizes_match_oracle(hws, oracle, tile):
    # every grid kernel form: tiles 8 and 16 take the cell-owning kernel
    # (sums in registers, records in any order), 24 and 32 the private-tile
    # kernel over origin-sorted records; two real images and one complex
    # one, a shifted and rotated phase centre, bins split into work items
    # (work_chunk 37) on a dense uv centre, through the host protocol (two
    # batches: the accumulate path) and the device form (the tile-writing
    # path), then predict
    import torch
    w = 96
    uvw, wl, vis, dl = case(40, w, nrow=900, nchan=3, frac=0.1, nimg=3)
    shift = (0.006, -0.004)
    uvw_l = lambda_samples(uvw, wl)
    maxw = np.abs(uvw_l[:, 2]).max()
    nw = 6
    ph = np.exp(1j * -2 * np.pi * (uvw_l[:, 0] * shift[0] + uvw_l[:, 1] * shift[1]))
    rot = [(vis[:, :, k].reshape(-1).astype(np.complex128) * ph).astype(np.complex64)
           for k in range(3)]
    img0, _ = oracle_invert(oracle, w, w, dl, dl, uvw_l, rot[0], nw, 0.0, maxw, shift=shift)
    img1, _ = oracle_invert(oracle, w, w, dl, dl, uvw_l, rot[1], nw, 0.0, maxw, shift=shift)
    (img2r, img2i), _ = oracle_invert(oracle, w, w, dl, dl, uvw_l, rot[2], nw, 0.0, maxw,
                                      cplx=True, shift=shift)
    kw = dict(precision="mixed", n_images=3, complex_images=(2,), l_shift=shift[0],
              m_shift=shift[1], rotate_visibilities=True, tile_size=tile, work_chunk=37)
    g = hws.WStackingGridder(w, w, dl, dl, 7, 63, **kw)
    g.prepare_w_layers(nw, 0.0, 0.0, maxw)
    g.start_inversion_pass(0)
    g.add_data(uvw[:400], wl, vis[:400])
    g.add_data(uvw[400:], wl, vis[400:])
    g.finish_inversion_pass()
    re, im = g.finalize_image(1.0)
    for got, want in ((re[0], img0), (re[1], img1), (re[2], img2r), (im[2], img2i)):
        assert rel(got, want) < TOL_IMG["mixed"]
    dev = torch.device("cuda:0")
    d = hws.DeviceGridder(w, w, dl, dl, 7, 63, **kw)
    d.prepare_w_layers(nw, 0.0, maxw)
    d_uvw, d_wl = torch.from_numpy(uvw).to(dev), torch.from_numpy(wl).to(dev)
    d_vis = torch.from_numpy(vis).to(dev)
    out = torch.zeros((3, w, w), dtype=torch.float32, device=dev)
    out_i = torch.zeros_like(out)
    d.plan(d_uvw, d_wl)
    d.invert(d_vis, 1.0, out, out_i)
    st = d.stats()
    assert st["chunks"] > st["bins_used"] + 20  # split bins
    o, oi = out.cpu().numpy(), out_i.cpu().numpy()
    for got, want in ((o[0], img0), (o[1], img1), (o[2], img2r), (oi[2], img2i)):
        assert rel(got, want) < TOL_IMG["mixed"]
    rng = np.random.default_rng(41)
    mre = rng.standard_normal((3, w, w)).astype(np.float32)
    mim = rng.standard_normal((3, w, w)).astype(np.float32)
    pred = torch.empty_like(d_vis)
    d.predict(torch.from_numpy(mre).to(dev), torch.from_numpy(mim).to(dev), pred)
    v_o = oracle_predict(oracle, w, w, dl, dl, uvw_l, mre[2], mim[2], nw, 0.0, maxw, cplx=True,
                         shift=shift)
    v_o = v_o * np.conj(ph)
    v = pred.cpu().numpy()[:, :, 2].reshape(-1)
    ok = ~np.isnan(v_o)
    assert np.array_equal(np.isnan(v), ~ok)
    assert rel(v[ok], v_o[ok]) < TOL_IMG["mixed"]
    d.close()


@pytest.mark.parametrize("precision", ["reference", "mixed"])
def test_complex_only_with_skipped_samples(hws, oracle, precision):
    # only a complex image (the real-image kind is absent, so the rows are
    # ordered by the complex kind's layer map), uv reaching past the grid
    # edge and w past the layer range: a large share of the samples is
    # gridded by no record; the invert must match the oracle, and the
    # predict must return NaN exactly for those samples (listed by the plan)
    import torch
    w = 128
    uvw, wl, vis, dl = case(77, w, nrow=1200, nchan=4, frac=0.7, nimg=1)
    uvw_l = lambda_samples(uvw, wl)
    maxw = 0.7 * np.abs(uvw_l[:, 2]).max()
    nw = 6
    (img_r, img_i), _ = oracle_invert(oracle, w, w, dl, dl, uvw_l, vis[:, :, 0].reshape(-1), nw,
                                      0.0, maxw, cplx=True)
    dev = torch.device("cuda:0")
    d = hws.DeviceGridder(w, w, dl, dl, 7, 63, precision=precision, n_images=1,
                          complex_images=(0,))
    d.prepare_w_layers(nw, 0.0, maxw)
    d_uvw, d_wl = torch.from_numpy(uvw).to(dev), torch.from_numpy(wl).to(dev)
    d_vis = torch.from_numpy(vis).to(dev)
    out = torch.zeros((1, w, w), dtype=torch.float32, device=dev)
    out_i = torch.zeros_like(out)
    d.plan(d_uvw, d_wl)
    d.invert(d_vis, 1.0, out, out_i)
    st = d.stats()
    skipped = st["samples_off_grid"] + st["samples_off_layer"]
    assert st["samples_off_grid"] > 0 and st["samples_off_layer"] > 0
    assert st["samples_gridded"] + skipped == uvw_l.shape[0]
    assert st["tiles_occupied"][0] == 0 and st["tiles_occupied"][1] > 0
    assert rel(out.cpu().numpy()[0], img_r) < TOL_IMG[precision]
    assert rel(out_i.cpu().numpy()[0], img_i) < TOL_IMG[precision]
    rng = np.random.default_rng(78)
    mre = rng.standard_normal((1, w, w)).astype(np.float32)
    mim = rng.standard_normal((1, w, w)).astype(np.float32)
    # the output starts as a finite value everywhere: every entry must be
    # written, NaN where no record grids the sample
    pred = torch.full_like(d_vis, 7.0 + 7.0j)
    d.predict(torch.from_numpy(mre).to(dev), torch.from_numpy(mim).to(dev), pred)
    v_o = oracle_predict(oracle, w, w, dl, dl, uvw_l, mre[0], mim[0], nw, 0.0, maxw, cplx=True)
    v = pred.cpu().numpy()[:, :, 0].reshape(-1)
    ok = ~np.isnan(v_o)
    assert int((~ok).sum()) == skipped
    assert np.array_equal(np.isnan(v), ~ok)
    assert rel(v[ok], v_o[ok]) < TOL_IMG[precision]
