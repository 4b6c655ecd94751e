"""Parity at the benchmarked workloads' full image sizes.

The benchmark times the 4096^2, 64-w-layer, 4-correlation workload (62
stations on a 40 km log spiral, 1891 baselines, 64 channels 115-165 MHz,
4 arcsec pixels) and the 16-facet batch (boxes 512-2048 px). These tests run
the same kernels at the same image sizes, layer counts, tile sizes and image
kinds on a row subsample of the same observation, so that the oracle (the CPU
restatement of WStackingGridder<float>) finishes in tens of seconds:

- 4096^2 x 64 layers, XX,XY,YX,YY (XX/YY real, XY/YX complex), device path
  (plan + invert + predict), mixed and plain precision: images, predicted
  visibilities and NaN positions against the oracle; gridded uv layers of the
  host protocol against the oracle's layers;
- 2048^2 x 32 layers (the compile-time 2048 transforms), mixed;
- the 16-facet set, autotuned tile, one captured graph per direction: every
  facet's image and predicted visibilities against the oracle at the facet's
  shift and rotation (gridding/wsmsgridder.cpp:185-190, 328-332).

The oracle runs on two of the four images (XX: slot 0 of the real images;
YX: slot 1 of the complex images). The other two carry exact multiples of
them (YY = -2 XX, XY = 0.5 YX, visibilities and models): every floating-point
operation commutes exactly with a power-of-two scale, so in the deterministic
batched precisions the GPU's YY and XY outputs must equal -2 and 0.5 times its
XX and YX outputs bit for bit, which checks that each image slot reads and
writes only its own data. One exception: the FP16 matrix cores that grid the
mixed precision round their sums asymmetrically about zero, so the YY image
(a negative factor) matches -2 XX to relative L2 1e-6 rather than bit for bit;
its predicted visibilities and the XY outputs stay bit-exact. The plain form
sums with global atomics in run-dependent order, so there every image is
compared with the scaled oracle instead.

Tolerances (relative L2 against the oracle float restatement): mixed 3e-5,
as for the small cases (tests/test_gpu_parity.py). The plain form computes
sample positions, layer indices and phases in FP32: at 4096^2 a position
u dl W reaches 2048 pixels, where an FP32 rounding of 1.2e-4 pixel is 0.8%
of the kernel table's 1/63-pixel step, so a share of the samples takes a
neighbouring kernel offset (or layer) that the FP64 oracle does not. That
costs about 6e-3 here (2e-7 at 64^2, where positions stay small), hence 1e-2 for
the plain form at these sizes. The batched precisions keep those quantities in
FP64 for this reason.
"""
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

TOL = {"mixed": 3e-5, "reference": 1e-5, "plain": 1e-2}
IMAGES = ("XX", "XY", "YX", "YY")
COMPLEX = (1, 2)
ARCSEC = np.pi / (180.0 * 3600.0)
SCALE = {3: -2.0, 1: 0.5}      # image -> factor times its source image
SOURCE = {3: 0, 1: 2}
NEG_MULTIPLE_TOL = 1e-6
ORACLE_THREADS = max(1, min(16, os.cpu_count() or 1))


def _rel(a, b):
    a = np.asarray(a, np.complex128).ravel()
    b = np.asarray(b, np.complex128).ravel()
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _say(msg):
    print(msg, flush=True)


class Case:
    """A row subsample of the 4096^2 observation with four correlations whose
    YY and XY are exact multiples of XX and YX; seeded values and models."""

    def __init__(self, width, pixel, n_w, row_step, seed):
        from wsclean_amd import synth
        obs = synth.config2()
        self.uvw = np.ascontiguousarray(obs.uvw_m[::row_step])
        self.wl = obs.wavelengths
        self.nrow, self.nchan = self.uvw.shape[0], self.wl.shape[0]
        self.width, self.pixel, self.n_w = width, pixel, n_w
        aw = np.abs(self.uvw[:, 2])
        self.min_w = float(aw.min() / self.wl.max())
        self.max_w = float(aw.max() / self.wl.min())
        self.uvw_l = (self.uvw[:, None, :] / self.wl[None, :, None]).reshape(-1, 3)
        rng = np.random.default_rng(seed)
        vis = np.zeros((self.nrow, self.nchan, 4), np.complex64)
        for k in (0, 2):
            vis[..., k] = (rng.standard_normal((self.nrow, self.nchan)) +
                           1j * rng.standard_normal((self.nrow, self.nchan)))
        for k, f in SCALE.items():
            vis[..., k] = vis[..., SOURCE[k]] * np.float32(f)
        self.vis = vis
        # models: a few hundred bright pixels over a low-level random field
        real = np.zeros((4, width, width), np.float32)
        imag = np.zeros((4, width, width), np.float32)
        for k in (0, 2):
            real[k] = 1e-3 * rng.standard_normal((width, width))
            idx = rng.integers(0, width, (300, 2))
            real[k][idx[:, 1], idx[:, 0]] += rng.uniform(0.1, 5.0, 300)
            if k in COMPLEX:
                imag[k] = 1e-3 * rng.standard_normal((width, width))
                imag[k][idx[:, 0], idx[:, 1]] += rng.uniform(-1.0, 1.0, 300)
        for k, f in SCALE.items():
            real[k] = real[SOURCE[k]] * np.float32(f)
            imag[k] = imag[SOURCE[k]] * np.float32(f)
        self.model_real, self.model_imag = real, imag
        self._oracle = {}

    def layer_choice(self, cplx, n=3):
        """The n most populated w-layers of an image kind and one sparse one
        (WToLayer, gridding/wstackinggridder.h:182-193)."""
        w = self.uvw_l[:, 2]
        if cplx:
            lay = np.rint((w + self.max_w) * (self.n_w - 1) / (2 * self.max_w))
        else:
            lay = np.rint((np.abs(w) - self.min_w) * (self.n_w - 1) / (self.max_w - self.min_w))
        counts = np.bincount(lay.astype(np.int64), minlength=self.n_w)[:self.n_w]
        order = np.argsort(counts)[::-1]
        sparse = [i for i in order[::-1] if counts[i] > 0][:1]
        return sorted(set(order[:n].tolist() + sparse))

    def oracle(self, oracle, k):
        """Oracle image (re, im|None), the layers of layer_choice and the
        predicted visibilities of image k (0 or 2)."""
        if k in self._oracle:
            return self._oracle[k]
        cplx = k in COMPLEX
        layers = self.layer_choice(cplx)
        W, pix = self.width, self.pixel
        g = oracle.RefGridder(W, W, pix, pix, ORACLE_THREADS, 7, 63, "float")
        g.set_is_complex(cplx)
        g.prepare_w_layers(self.n_w, 1e13, self.min_w, self.max_w)
        g.start_inversion_pass(0)
        g.add_samples(self.uvw_l, self.vis[..., k].reshape(-1))
        lay = {i: g.layer(i) for i in layers}
        g.finish_inversion_pass()
        g.finalize_image(1.0)
        img = g.image(imag=cplx)
        del g
        g = oracle.RefGridder(W, W, pix, pix, ORACLE_THREADS, 7, 63, "float")
        g.set_is_complex(cplx)
        g.prepare_w_layers(self.n_w, 1e13, self.min_w, self.max_w)
        g.initialize_prediction(self.model_real[k], self.model_imag[k] if cplx else None)
        g.start_prediction_pass(0)
        v = g.sample(self.uvw_l)
        del g
        self._oracle[k] = (img, lay, v)
        return self._oracle[k]

    def expected(self, oracle, k):
        """Oracle outputs of image k, scaled for the derived images."""
        src = SOURCE.get(k, k)
        img, _, v = self.oracle(oracle, src)
        f = SCALE.get(k, 1.0)
        if isinstance(img, tuple):
            img = (img[0] * f, img[1] * f)
        else:
            img = img * f
        return img, v * np.complex64(f)


def run_device(case, precision, tile_size=0):
    """Plan + invert + predict of all four images on the device path."""
    import wsclean_amd
    dev = torch.device("cuda:0")
    W = case.width
    g = wsclean_amd.DeviceGridder(W, W, case.pixel, case.pixel, 7, 63, precision=precision,
                                  n_images=4, complex_images=COMPLEX, tile_size=tile_size)
    assert g.prepare_w_layers(case.n_w, case.min_w, case.max_w) == 1
    uvw = torch.from_numpy(case.uvw).to(dev)
    wl = torch.from_numpy(case.wl).to(dev)
    vis = torch.from_numpy(case.vis).to(dev)
    re = torch.zeros((4, W, W), dtype=torch.float32, device=dev)
    im = torch.zeros_like(re)
    g.plan(uvw, wl)
    g.invert(vis, 1.0, re, im)
    out = torch.empty_like(vis)
    g.predict(torch.from_numpy(case.model_real).to(dev),
              torch.from_numpy(case.model_imag).to(dev), out)
    g.synchronize()
    stats = g.stats()
    g.close()
    return re.cpu().numpy(), im.cpu().numpy(), out.cpu().numpy(), stats


def check_against_oracle(oracle, case, precision, re, im, vis_out, label):
    for k in range(4):
        img_o, v_o = case.expected(oracle, k)
        if k in COMPLEX:
            e = _rel(re[k] + 1j * im[k], img_o[0] + 1j * img_o[1])
        else:
            e = _rel(re[k], img_o)
        got = vis_out[..., k].reshape(-1)
        nan_o = np.isnan(v_o)
        if precision == "plain":
            # FP32 sample positions: a sample on the edge of the acceptance
            # window (-W/2 < x <= W/2) may fall the other way than in FP64
            flips = int((np.isnan(got) != nan_o).sum())
            assert flips <= 1e-5 * nan_o.size, (label, IMAGES[k], flips)
            nan_o = nan_o | np.isnan(got)
        else:
            assert np.array_equal(np.isnan(got), nan_o), (label, IMAGES[k], "NaN positions")
        ev = _rel(got[~nan_o], v_o[~nan_o])
        _say(f"{label} {precision} {IMAGES[k]}: image rel {e:.2e}, predict rel {ev:.2e}, "
             f"{int(nan_o.sum())} NaN of {nan_o.size}")
        assert e < TOL[precision], (label, IMAGES[k], e)
        assert ev < TOL[precision], (label, IMAGES[k], ev)


def _same_multiple(a, b, f):
    # negative factors: the FP16 matrix cores of the mixed precision's grid
    # kernel round their sums asymmetrically about zero (for negated operands
    # v_mfma_f32_16x16x32_f16 returns the negated sum but for the last bit of
    # a few outputs), so an image gridded from negated data matches to the
    # sums' rounding rather than bit for bit
    if f < 0:
        return _rel(a, b) < NEG_MULTIPLE_TOL
    return np.array_equal(a, b)


def check_exact_multiples(re, im, vis_out, label):
    for k, f in SCALE.items():
        s = SOURCE[k]
        assert _same_multiple(re[k], re[s] * np.float32(f), f), (label, IMAGES[k], "real")
        if k in COMPLEX:
            assert _same_multiple(im[k], im[s] * np.float32(f), f), (label, IMAGES[k], "imag")
        a, b = vis_out[..., k], vis_out[..., s] * np.complex64(f)
        ok = ~np.isnan(b)
        assert np.array_equal(np.isnan(a), ~ok), (label, IMAGES[k])
        assert np.array_equal(a[ok], b[ok]), (label, IMAGES[k], "visibilities")


@pytest.fixture(scope="module")
def case4096():
    # every 48th row: 70,913 rows x 64 channels = 4.5 M samples, 18 M visibilities
    return Case(4096, 4.0 * ARCSEC, 64, 48, seed=4096)


@pytest.mark.parametrize("precision", ["mixed", "plain"])
def test_config2_4096_device_path_matches_oracle(oracle, case4096, precision):
    c = case4096
    _say(f"4096^2 x {c.n_w} layers, {c.nrow} rows x {c.nchan} ch x 4, {precision}")
    re, im, out, st = run_device(c, precision)
    assert st["samples_in"] == c.nrow * c.nchan
    if precision != "plain":  # the plain form keeps no placement counters
        assert st["samples_gridded"] > 0.9 * st["samples_in"]
    assert np.isfinite(re).all() and np.isfinite(im).all()
    if precision != "plain":
        check_exact_multiples(re, im, out, "4096")
    check_against_oracle(oracle, c, precision, re, im, out, "4096")


def test_config2_4096_gridded_layers_match_oracle(oracle, case4096):
    """Layers of the host protocol (the same bin + matrix-core grid kernels at
    65,536 tiles per layer) against the oracle's layers, XX and YX: the three
    most populated layers of each kind and a sparse one. A dense uv cell here
    sums thousands of FP32 products, so the tolerance is set from the FP32
    oracle's own distance to its FP64 form on the same layer (the noise floor
    of FP32 gridding at this size): three times that, and never below the
    small cases' 2e-6."""
    import wsclean_amd
    c = case4096
    for k in (0, 2):
        cplx = k in COMPLEX
        _, lay_o, _ = c.oracle(oracle, k)
        g = oracle.RefGridder(c.width, c.width, c.pixel, c.pixel, ORACLE_THREADS, 7, 63,
                              "double")
        g.set_is_complex(cplx)
        g.prepare_w_layers(c.n_w, 1e13, c.min_w, c.max_w)
        g.start_inversion_pass(0)
        g.add_samples(c.uvw_l, c.vis[..., k].reshape(-1))
        lay_d = {i: g.layer(i) for i in lay_o}
        del g
        h = wsclean_amd.WStackingGridder(c.width, c.width, c.pixel, c.pixel, 7, 63,
                                         precision="mixed", n_images=1,
                                         complex_images=(0,) if cplx else ())
        h.prepare_w_layers(c.n_w, 0.0, c.min_w, c.max_w)
        h.start_inversion_pass(0)
        h.add_data(c.uvw, c.wl, c.vis[..., k:k + 1])
        for i, ref in lay_o.items():
            floor = _rel(ref, lay_d[i])
            e = _rel(h.get_gridded_uv_layer(i), ref)
            e_d = _rel(h.get_gridded_uv_layer(i), lay_d[i])
            _say(f"4096 {IMAGES[k]} layer {i}: rel {e:.2e} vs oracle FP32 ({e_d:.2e} vs FP64; "
                 f"oracle FP32 vs FP64 {floor:.2e})")
            assert e < max(2e-6, 3.0 * floor), (IMAGES[k], i, e, floor)
        h.close()


def test_config2_2048_device_path_matches_oracle(oracle):
    c = Case(2048, 8.0 * ARCSEC, 32, 48, seed=2048)
    _say(f"2048^2 x {c.n_w} layers, {c.nrow} rows x {c.nchan} ch x 4, mixed")
    re, im, out, _ = run_device(c, "mixed")
    check_exact_multiples(re, im, out, "2048")
    check_against_oracle(oracle, c, "mixed", re, im, out, "2048")


def test_config4_facet_batch_matches_oracle(oracle):
    """The 16-facet set (576-2048 px, runtime-radix sizes included), w-layers
    from the reference's suggestion per facet, the autotuned tile, captured
    into one graph per direction."""
    from wsclean_amd import synth
    from wsclean_amd.autotune import autotune
    from wsclean_amd.facets import FacetBatch
    dev = torch.device("cuda:0")
    pix = 4.0 * ARCSEC
    obs = synth.config2(scale=0.25)
    uvw = np.ascontiguousarray(obs.uvw_m[::16])
    wl = obs.wavelengths
    nrow, nchan = uvw.shape[0], wl.shape[0]
    rng = np.random.default_rng(44)
    vis = np.zeros((nrow, nchan, 4), np.complex64)
    for k in (0, 2):
        vis[..., k] = rng.standard_normal((nrow, nchan)) + 1j * rng.standard_normal((nrow, nchan))
    for k, f in SCALE.items():
        vis[..., k] = vis[..., SOURCE[k]] * np.float32(f)
    facets = synth.config4_facets()
    d_uvw, d_wl = torch.from_numpy(uvw).to(dev), torch.from_numpy(wl).to(dev)
    d_vis = torch.from_numpy(vis).to(dev)
    big = max(facets, key=lambda f: f.width)
    tuned = autotune(big.width, big.height, pix, pix, d_uvw, d_wl, d_vis, n_images=4,
                     complex_images=COMPLEX, l_shift=big.l_shift, m_shift=big.m_shift,
                     chunks=(0, 8192), repeats=1)
    fb = FacetBatch(facets, pix, pix, d_uvw, d_wl, n_images=4, complex_images=COMPLEX,
                    tile_size=tuned["tile_size"], work_chunk=tuned["work_chunk"])
    _say(f"facets: sizes {[f.width for f in facets]}, w-layers {fb.n_w}, tile "
         f"{tuned['tile_size']}, chunk {tuned['work_chunk']}, {nrow} rows x {nchan} ch")
    assert fb.passes == [1] * len(facets)
    re, im = fb.images(), fb.images()
    gi = fb.capture_invert(d_vis, 1.0, re, im)
    models_re, models_im = [], []
    for f in facets:
        m_re = np.zeros((4, f.height, f.width), np.float32)
        m_im = np.zeros((4, f.height, f.width), np.float32)
        for k in (0, 2):
            idx = rng.integers(0, f.width, (50, 2))
            m_re[k][idx[:, 1], idx[:, 0]] = rng.uniform(0.1, 5.0, 50)
            if k in COMPLEX:
                m_im[k][idx[:, 1], idx[:, 0]] = rng.uniform(-1.0, 1.0, 50)
        for k, fct in SCALE.items():
            m_re[k] = m_re[SOURCE[k]] * np.float32(fct)
            m_im[k] = m_im[SOURCE[k]] * np.float32(fct)
        models_re.append(torch.from_numpy(m_re).to(dev))
        models_im.append(torch.from_numpy(m_im).to(dev))
    outs = [torch.empty_like(d_vis) for _ in facets]
    gp = fb.capture_predict(models_re, models_im, outs)
    gi.launch()
    gp.launch()
    torch.cuda.synchronize()
    aw = np.abs(uvw[:, 2])
    min_w, max_w = float(aw.min() / wl.max()), float(aw.max() / wl.min())
    assert fb.w_limits == (min_w, max_w)
    uvw_l = (uvw[:, None, :] / wl[None, :, None]).reshape(-1, 3)
    for i, f in enumerate(facets):
        r, m = re[i].cpu().numpy(), im[i].cpu().numpy()
        o = outs[i].cpu().numpy()
        check_exact_multiples(r, m, o, f"facet {i}")
        ph = 2 * np.pi * (uvw_l[:, 0] * f.l_shift + uvw_l[:, 1] * f.m_shift)
        for k in (0, 2):
            cplx = k in COMPLEX
            g = oracle.RefGridder(f.width, f.height, pix, pix, ORACLE_THREADS, 7, 63, "float")
            g.set_is_complex(cplx)
            g.set_denormal_phase_centre(f.l_shift, f.m_shift)
            g.prepare_w_layers(fb.n_w[i], 1e13, min_w, max_w)
            g.start_inversion_pass(0)
            rot = (vis[..., k].reshape(-1).astype(np.complex128) *
                   np.exp(-1j * ph)).astype(np.complex64)
            g.add_samples(uvw_l, rot)
            g.finish_inversion_pass()
            g.finalize_image(1.0)
            img_o = g.image(imag=cplx)
            del g
            if cplx:
                e = _rel(r[k] + 1j * m[k], img_o[0] + 1j * img_o[1])
            else:
                e = _rel(r[k], img_o)
            g = oracle.RefGridder(f.width, f.height, pix, pix, ORACLE_THREADS, 7, 63, "float")
            g.set_is_complex(cplx)
            g.set_denormal_phase_centre(f.l_shift, f.m_shift)
            g.prepare_w_layers(fb.n_w[i], 1e13, min_w, max_w)
            g.initialize_prediction(models_re[i][k].cpu().numpy(),
                                    models_im[i][k].cpu().numpy() if cplx else None)
            g.start_prediction_pass(0)
            v_o = g.sample(uvw_l).astype(np.complex128) * np.exp(1j * ph)
            del g
            got = o[..., k].reshape(-1)
            nan_o = np.isnan(v_o)
            assert np.array_equal(np.isnan(got), nan_o), (i, IMAGES[k])
            ev = _rel(got[~nan_o], v_o[~nan_o])
            _say(f"facet {i} {f.width}^2 x {fb.n_w[i]} shift ({f.l_shift:+.4f}, "
                 f"{f.m_shift:+.4f}) {IMAGES[k]}: image rel {e:.2e}, predict rel {ev:.2e}")
            assert e < 3e-5, (i, IMAGES[k], e)
            assert ev < 3e-5, (i, IMAGES[k], ev)
    gi.close()
    gp.close()
    fb.close()
