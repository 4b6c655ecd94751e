"""Pins the oracle (oracle/: CPU restatement of the reference w-stacking
gridder + FP64 direct FT) against golden fixtures, closed forms and
brute force, so that it can serve as the checker of the HIP path.

Each test names the reference behaviour it pins (file:line in the
reference tree).
"""
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def rel(a, b):
    return np.linalg.norm(np.ravel(a - b)) / np.linalg.norm(np.ravel(b))


# ---------------------------------------------------------------- kernels
def test_kaiser_bessel_matches_scipy_fixture(oracle):
    # gridding/wstackinggridder.cpp:425-447, 504-515 against scipy.special.i0
    f = np.load(os.path.join(GOLDEN, "kernel_kb.npz"))
    for key, (mode, os_) in {"kb_63_7": ("kb", 63), "kb_1023_7": ("kb", 1023),
                             "kb_nosinc_63_7": ("kb_nosinc", 63)}.items():
        k = oracle.make_kernel(mode, os_, 7)
        assert np.max(np.abs(k - f[key])) < 1e-9


def test_kernel_known_answers(oracle):
    k = oracle.make_kernel("kb", 63, 7)
    mid = 7 * 63 // 2
    assert abs(k[mid] - 1.0) < 1e-12          # sinc(0) * I0(a)/I0(a) * OS/OS
    assert np.array_equal(k, k[::-1])         # mirrored (:446)
    nn = oracle.make_kernel("nn", 63, 7)
    assert nn[mid] == 1.0 and np.count_nonzero(nn) == 1  # GetKernel NN delta (:409-411)
    with pytest.raises(RuntimeError):
        oracle.make_kernel("bh", 63, 7)       # rejected (:418-420)
    assert abs(oracle.lib().orc_bessel0(0.0, 1e-8) - 1.0) < 1e-15


def test_kernel_correction_is_redft01(oracle):
    # correctImageForKernel's FFTW REDFT01 (cpp:784-789) == scipy dct type 3
    d = np.load(os.path.join(GOLDEN, "dct3.npz"))
    for w in (64, 96):
        g = oracle.RefGridder(w, w, 0.01, 0.01, 1, 7, 63, "double")
        assert rel(g.kernel_correction(w), d[f"d_{w}"]) < 1e-9


# ------------------------------------------------------------ layer map
def test_layer_map_round_trip_and_passes(oracle):
    for cplx in (False, True):
        g = oracle.RefGridder(64, 64, 0.01, 0.01, 1, 7, 63, "double")
        g.set_is_complex(cplx)
        g.prepare_w_layers(17, 64 * 64 * 4 * 5 * 1 + 64 * 64 * 4 * 2 * 5, 10.0, 500.0)
        nw = 17
        lo = -500.0 if cplx else 10.0
        dw = (500.0 - lo) / (nw - 1)
        for w in np.linspace(lo, 500.0, 301):
            layer = g.w_to_layer(w)
            assert 0 <= layer < nw
            assert abs(g.layer_to_w(layer) - w) <= dw / 2 + 1e-9
        if not cplx:
            # |w| folds: -w maps to the layer of w (h:190-191)
            assert g.w_to_layer(-123.0) == g.w_to_layer(123.0)
            # round(-0.33) is -0.0, which the size_t cast maps to layer 0
            assert g.w_to_layer(0.0) == 0
            # below minW - dw/2: negative index, reported as SIZE_MAX
            g2 = oracle.RefGridder(64, 64, 0.01, 0.01, 1, 7, 63, "double")
            g2.prepare_w_layers(17, 1e9, 100.0, 500.0)
            assert g2.w_to_layer(0.0) == 2**64 - 1
        # passes partition [0, nW) (h:566-568)
        npass = g.n_passes
        assert npass > 1
        starts = [(nw * p) // npass for p in range(npass + 1)]
        assert starts[0] == 0 and starts[-1] == nw
        assert all(b >= a for a, b in zip(starts, starts[1:]))


def test_prepare_w_layers_min_equals_max(oracle):
    # gridding/wstackinggridder.cpp:70-75: maxW += 1 keeps layer 0 at minW
    g = oracle.RefGridder(32, 32, 0.01, 0.01, 1, 7, 63, "double")
    g.prepare_w_layers(4, 1e9, 7.0, 7.0)
    assert g.layer_to_w(0) == 7.0 and g.layer_to_w(3) == 8.0


# ----------------------------------------------------- gridding footprint
def _single_sample_layer(oracle, u, v, w=0.0, vis=1 + 0j, mode="kb", cplx=False,
                         size=32, prec="double"):
    dl = 0.01
    g = oracle.RefGridder(size, size, dl, dl, 1, 7, 63, prec)
    if mode != "kb":
        g.set_grid_mode(mode)
    g.set_is_complex(cplx)
    g.prepare_w_layers(1, 1e9, 0.0, 1.0)
    g.start_inversion_pass(0)
    g.add_samples(np.array([[u / (dl * size), v / (dl * size), w]]), np.array([vis], np.complex64))
    return g.layer(0)


@pytest.mark.parametrize("fx,fy", [(0.0, 0.0), (0.2, -0.3), (-0.49, 0.41), (0.31, 0.11)])
def test_footprint_is_kernel_centred_on_sample(oracle, fx, fy):
    # AddDataSample (cpp:548-593): the deposit at cell offset d from the
    # nearest cell is K(d - f) sampled on the oversampled grid.
    k = oracle.make_kernel("kb", 63, 7)
    mid = len(k) // 2
    x0, y0 = 5, 7
    lay = _single_sample_layer(oracle, x0 + fx, y0 + fy)
    kx = int(np.round(fx * 63))
    ky = int(np.round(fy * 63))
    expect = np.zeros((32, 32))
    for j in range(-3, 4):
        for i in range(-3, 4):
            expect[y0 + j, x0 + i] = k[mid + j * 63 - ky] * k[mid + i * 63 - kx]
    assert np.allclose(lay.real, expect, rtol=1e-12, atol=1e-15)
    assert np.abs(lay.imag).max() == 0.0


def test_footprint_wraps_at_edges_and_negative_uv(oracle):
    # cyclic wrap (cpp:566-578) and x < 0 -> x + W (cpp:563-564)
    lay = _single_sample_layer(oracle, -1.0, 0.0)
    nz = np.argwhere(np.abs(lay) > 0)
    assert set(nz[:, 1]) == {28, 29, 30, 31, 0, 1, 2}
    assert set(nz[:, 0]) == {29, 30, 31, 0, 1, 2, 3}


def test_acceptance_window_is_half_open(oracle):
    # -W/2 < x <= W/2 (cpp:561-562)
    assert np.abs(_single_sample_layer(oracle, 16.0, 0.0)).sum() > 0
    assert np.abs(_single_sample_layer(oracle, -16.0, 0.0)).sum() == 0


def test_negative_w_folds_to_conjugate(oracle):
    # real images: (u, v, w, V) -> (-u, -v, -w, V*) (cpp:528-533)
    a = _single_sample_layer(oracle, 3.2, -1.7, w=-0.5, vis=1 + 2j)
    b = _single_sample_layer(oracle, -3.2, 1.7, w=0.5, vis=1 - 2j)
    assert np.array_equal(a, b)
    c = _single_sample_layer(oracle, 3.2, -1.7, w=-0.5, vis=1 + 2j, cplx=True)
    assert not np.array_equal(a, c)


def test_nearest_neighbour_deposits_delta(oracle):
    lay = _single_sample_layer(oracle, 4.4, -2.6, mode="nn", vis=2 - 1j)
    nz = np.argwhere(np.abs(lay) > 0)
    assert nz.tolist() == [[32 - 3, 4]]
    assert lay[29, 4] == 2 - 1j


# --------------------------------------------------------------------- FFT
@pytest.mark.parametrize("shape", [(16, 16), (12, 20), (30, 42), (64, 28)])
def test_fft_matches_numpy(oracle, shape):
    rng = np.random.default_rng(1)
    a = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    assert rel(oracle.fft2d(a, +1), np.fft.ifft2(a) * a.size) < 1e-13
    assert rel(oracle.fft2d(a, -1), np.fft.fft2(a)) < 1e-13
    # round trip = N * identity
    assert rel(oracle.fft2d(oracle.fft2d(a, -1), +1), a * a.size) < 1e-13


# --------------------------------------------------------------------- DFT
def test_dft_matches_brute_force_fixture(oracle):
    s = np.load(os.path.join(GOLDEN, "dft_small.npz"))
    w, h, dl, dm = int(s["width"]), int(s["height"]), float(s["dl"]), float(s["dm"])
    img = oracle.dft_invert(w, h, dl, dm, s["uvw"], s["vis"], nthreads=3)
    assert rel(img, s["image"]) < 1e-12
    pred = oracle.dft_predict(w, h, dl, dm, s["uvw"], s["model"], nthreads=2)
    assert rel(pred, s["predicted"]) < 1e-12


def test_dft_predict_is_adjoint_of_invert(oracle):
    rng = np.random.default_rng(3)
    w = h = 16
    uvw = rng.uniform(-20, 20, (25, 3))
    v = (rng.standard_normal(25) + 1j * rng.standard_normal(25)).astype(np.complex64)
    img = rng.standard_normal((h, w))
    a = np.sum(oracle.dft_invert(w, h, 0.02, 0.02, uvw, v) * img)
    b = np.real(np.vdot(v.astype(np.complex128), oracle.dft_predict(w, h, 0.02, 0.02, uvw, img)))
    assert abs(a - b) < 1e-10 * abs(a)


def test_dft_psf_peak_is_one(oracle):
    # V == 1, natural weighting, factor 1/sum(w): PSF peak 1 at (W/2, H/2)
    rng = np.random.default_rng(4)
    uvw = rng.uniform(-30, 30, (50, 3))
    img = oracle.dft_invert(32, 32, 0.01, 0.01, uvw, np.ones(50, np.complex64), factor=1 / 50)
    assert abs(img[16, 16] - 1.0) < 1e-12 and img.max() <= 1.0 + 1e-12


# ------------------------------------------------ restatement end to end
def _random_case(rng, w, n, wmax, field_frac=0.45, dl=0.004):
    uvw = np.column_stack([rng.uniform(-1, 1, n) * field_frac / dl,
                           rng.uniform(-1, 1, n) * field_frac / dl,
                           rng.uniform(-wmax, wmax, n)])
    vis = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    return uvw, vis, dl


def _invert(oracle, w, dl, uvw, vis, nw, prec, cplx=False, mem=1e10, nthreads=2):
    g = oracle.RefGridder(w, w, dl, dl, nthreads, 7, 63, prec)
    g.set_is_complex(cplx)
    g.prepare_w_layers(nw, mem, 0.0, np.abs(uvw[:, 2]).max())
    for p in range(g.n_passes):
        g.start_inversion_pass(p)
        g.add_samples(uvw, vis)
        g.finish_inversion_pass()
    g.finalize_image(1.0)
    return g.image(imag=cplx), g.n_passes


def test_float_and_double_restatements_agree(oracle):
    rng = np.random.default_rng(6)
    uvw, vis, dl = _random_case(rng, 64, 400, 150)
    (a, _), ((b, _)) = (_invert(oracle, 64, dl, uvw, vis, 16, "float"),
                        _invert(oracle, 64, dl, uvw, vis, 16, "double"))
    assert rel(a, b) < 2e-6


def test_multipass_equals_single_pass(oracle):
    # PrepareWLayers memory planning (cpp:77-113) must not change the image
    rng = np.random.default_rng(7)
    uvw, vis, dl = _random_case(rng, 32, 200, 100)
    single, n1 = _invert(oracle, 32, dl, uvw, vis, 12, "double")
    multi, n2 = _invert(oracle, 32, dl, uvw, vis, 12, "double",
                        mem=32 * 32 * 8 * 5 * 2 + 32 * 32 * 8 * 2 * 3)
    assert n1 == 1 and n2 > 1
    assert rel(multi, single) < 1e-12


def test_invert_converges_to_dft_with_layers(oracle):
    # the w-stacking image approaches the FP64 DFT in the image centre as the
    # number of w-layers grows (w-quantisation error ~ pi dw |n'|)
    rng = np.random.default_rng(8)
    uvw, vis, dl = _random_case(rng, 64, 150, 400, dl=0.002)
    ref = oracle.dft_invert(64, 64, dl, dl, uvw, vis)
    c = slice(20, 44)
    errs = []
    for nw in (2, 8, 32, 128):
        img, _ = _invert(oracle, 64, dl, uvw, vis, nw, "double")
        errs.append(np.abs(img - ref)[c, c].max() / np.abs(ref).max())
    assert all(b < a for a, b in zip(errs, errs[1:]))
    assert errs[-1] < 1e-2 and errs[0] > 10 * errs[-1]


def test_predict_matches_dft_predict_near_centre(oracle):
    rng = np.random.default_rng(9)
    w, dl = 64, 0.006
    uvw, _, _ = _random_case(rng, w, 120, 200, dl=dl)
    model = np.zeros((w, w), np.float32)
    model[30, 35] = 1.0
    model[36, 29] = -0.5
    g = oracle.RefGridder(w, w, dl, dl, 2, 7, 63, "double")
    g.prepare_w_layers(64, 1e10, 0.0, np.abs(uvw[:, 2]).max())
    g.initialize_prediction(model)
    g.start_prediction_pass(0)
    v = g.sample(uvw)
    ref = oracle.dft_predict(w, w, dl, dl, uvw, model.astype(np.float64))
    assert np.abs(v - ref).max() / np.abs(ref).max() < 5e-3


@pytest.mark.parametrize("cplx", [False, True])
def test_restatement_predict_is_adjoint_of_invert(oracle, cplx):
    rng = np.random.default_rng(10)
    w = 48
    uvw, vis, dl = _random_case(rng, w, 250, 120)
    img_re = rng.standard_normal((w, w)).astype(np.float32)
    img_im = rng.standard_normal((w, w)).astype(np.float32) if cplx else None
    ah = _invert(oracle, w, dl, uvw, vis, 16, "double", cplx=cplx)[0]
    g = oracle.RefGridder(w, w, dl, dl, 1, 7, 63, "double")
    g.set_is_complex(cplx)
    g.prepare_w_layers(16, 1e10, 0.0, np.abs(uvw[:, 2]).max())
    g.initialize_prediction(img_re, img_im)
    g.start_prediction_pass(0)
    a_img = g.sample(uvw).astype(np.complex128)
    ah_re, ah_im = ah if cplx else (ah, None)
    lhs = np.sum(ah_re * img_re) + (np.sum(ah_im * img_im) if cplx else 0.0)
    rhs = np.real(np.vdot(vis.astype(np.complex128), a_img))
    assert abs(lhs - rhs) < 2e-6 * abs(lhs)


def test_off_grid_and_off_pass_samples_are_nan(oracle):
    # SampleDataSample NaN contract (cpp:626-627, 677-678, 685-687)
    g = oracle.RefGridder(32, 32, 0.01, 0.01, 1, 7, 63, "double")
    g.prepare_w_layers(8, 32 * 32 * 8 * 5 + 32 * 32 * 8 * 2 * 2, 1.0, 100.0)
    assert g.n_passes > 1
    g.initialize_prediction(np.ones((32, 32), np.float32))
    g.start_prediction_pass(0)
    v = g.sample(np.array([[0.0, 0.0, 1.0], [0.0, 0.0, 99.0], [1e6, 0.0, 1.0]]))
    assert np.isfinite(v[0]) and np.isnan(v[1]) and np.isnan(v[2])


# ------------------------------------------------------------ ImageWeights
def test_image_weights_modes(oracle):
    # structures/imageweights.cpp:113-144 on a hand-countable grid
    uvw = np.array([[0.0, 0.5, 0], [0.0, 0.5, 0], [0.1, 0.7, 0], [0.0, -0.5, 0]])
    wl = np.array([1.0])
    w = np.array([1.0, 2.0, 1.0, 3.0], np.float32)
    dl = 1.0 / 8  # cell = 1 / (W dl) = 1 wavelength for W = 8
    res = {}
    for mode in ("natural", "uniform", "briggs"):
        iw = oracle.ImageWeights(mode, 8, 8, dl, dl, robust=0.0)
        iw.grid(uvw, wl, w)
        iw.finish()
        res[mode] = iw.lookup(uvw, wl)[:, 0]
    # all four samples land in cell (x=4, y=0) (v<0 folds to v>0)
    assert np.all(res["natural"] == 1.0)
    assert np.allclose(res["uniform"], 1.0 / 7.0)
    s2 = 25.0 / (49.0 / 7.0)  # (5 * 10^-0)^2 / (sum w^2 / sum w)
    assert np.allclose(res["briggs"], 1.0 / (1.0 + 7.0 * s2))


def test_suggested_w_grid_size_and_fft_size(oracle):
    s = oracle.ms_settings(1024, 1024, np.deg2rad(20 / 3600), np.deg2rad(20 / 3600), ncpus=1)
    # wsmsgridder.cpp:86-96 closed form
    maxl = 1024 * np.deg2rad(20 / 3600) * 0.5
    expect = int(np.ceil(2 * np.pi * (500.0 - 10.0) * (1 - np.sqrt(1 - 2 * maxl * maxl))))
    assert oracle.suggested_w_grid_size(s, 10.0, 500.0) == expect
    # math/calculatefftsize.h: 7-smooth multiple of 4
    mn, opt = oracle.calculate_fft_size(4000, 1e-5, 1e-3)
    assert mn == 80 and opt == 80
    mn, opt = oracle.calculate_fft_size(4000, 1.1e-5, 1e-3)
    assert opt % 4 == 0 and opt >= mn
