"""The runtime-radix transforms of the blocked layer layout take four rows
(columns) per workgroup where they fit the LDS, else two or one (a four-row
block then spans two or four workgroups). The count is forced with
HWS_FFT_LINES=1, 2 or 4 when a context is created.
- at small sizes, one and two lines against four: the transforms do the same
  operations on every line, so predicted visibilities are equal bit for bit;
  the images agree within 1e-6 of the peak (each context plans its samples
  itself, and the binning scatter may order a bin's records differently,
  which changes the gridded FP32 sums at rounding level);
- sizes whose rows (or columns) take two lines or one by themselves (4800:
  four rows would need 163 KB of LDS; 12288: two would need 209 KB) against
  the oracle, to the parity tolerances."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL_IMG = {"reference": 1e-5, "mixed": 3e-5}
C_LIGHT = 299792458.0


def rel(a, b):
    a = np.asarray(a, np.complex128)
    b = np.asarray(b, np.complex128)
    return np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel())


def _case(seed, width, nrow=300, nchan=3, dl=0.004, frac=0.42):
    rng = np.random.default_rng(seed)
    wl = C_LIGHT / np.linspace(120e6, 150e6, nchan)
    umax = frac / dl * wl.min()
    uvw = np.column_stack([rng.uniform(-umax, umax, nrow), rng.uniform(-umax, umax, nrow),
                           rng.uniform(-60.0, 60.0, nrow)])
    vis = (rng.standard_normal((nrow, nchan, 1)) +
           1j * rng.standard_normal((nrow, nchan, 1))).astype(np.complex64)
    return uvw, wl, vis, dl


def _gpu(w, h, dl, uvw, wl, vis, nw, maxw, precision, model, lines):
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    import wsclean_amd
    old = os.environ.get("HWS_FFT_LINES")
    os.environ["HWS_FFT_LINES"] = str(lines)
    try:
        g = wsclean_amd.WStackingGridder(w, h, dl, dl, 7, 63, precision=precision)
        p = wsclean_amd.WStackingGridder(w, h, dl, dl, 7, 63, precision=precision)
    finally:
        if old is None:
            del os.environ["HWS_FFT_LINES"]
        else:
            os.environ["HWS_FFT_LINES"] = old
    g.prepare_w_layers(nw, 0.0, 0.0, maxw)
    g.start_inversion_pass(0)
    g.add_data(uvw, wl, vis)
    g.finish_inversion_pass()
    re, _ = g.finalize_image(1.0)
    g.close()
    p.prepare_w_layers(nw, 0.0, 0.0, maxw)
    p.initialize_prediction(model)
    p.start_prediction_pass(0)
    v = p.sample_data(uvw, wl).reshape(-1)
    p.close()
    return np.asarray(re), v


@pytest.mark.parametrize("shape,precision", [((144, 112), "mixed"), ((96, 80), "reference"),
                                             ((480, 240), "mixed")])
@pytest.mark.parametrize("lines", [1, 2])
def test_fewer_lines_per_workgroup_equal_four(shape, precision, lines):
    w, h = shape
    uvw, wl, vis, dl = _case(31, min(w, h))
    maxw = float(np.abs(uvw[:, 2]).max() / wl.min())
    model = np.random.default_rng(3).standard_normal((h, w)).astype(np.float32)
    re4, v4 = _gpu(w, h, dl, uvw, wl, vis, 6, maxw, precision, model, 4)
    re2, v2 = _gpu(w, h, dl, uvw, wl, vis, 6, maxw, precision, model, lines)
    assert np.abs(re4).max() > 0
    assert np.abs(re2 - re4).max() <= 1e-6 * np.abs(re4).max()
    assert np.array_equal(np.isnan(v2), np.isnan(v4))
    ok = ~np.isnan(v4)
    assert ok.sum() > 0 and np.array_equal(v2[ok], v4[ok])


@pytest.mark.parametrize("shape,precision", [((4800, 160), "mixed"), ((160, 4800), "reference"),
                                             ((12288, 64), "mixed"), ((64, 12288), "mixed")])
def test_long_line_sizes_match_oracle(oracle, shape, precision):
    # 4800: two lines per workgroup; 12288 (= 3 x 4096): one
    w, h = shape
    # (an image 0.4 rad across its long side, inside the horizon)
    uvw, wl, vis, dl = _case(41, min(w, h), nrow=200, dl=0.4 / max(w, h))
    uvw_l = (uvw[:, None, :] / wl[None, :, None]).reshape(-1, 3)
    maxw = float(np.abs(uvw_l[:, 2]).max())
    nw = 4
    og = oracle.RefGridder(w, h, dl, dl, 2, 7, 63, "float")
    og.prepare_w_layers(nw, 1e12, 0.0, maxw)
    og.start_inversion_pass(0)
    og.add_samples(uvw_l, vis.reshape(-1))
    og.finish_inversion_pass()
    og.finalize_image(1.0)
    img_o = og.image(imag=False)
    model = np.random.default_rng(9).standard_normal((h, w)).astype(np.float32)
    op = oracle.RefGridder(w, h, dl, dl, 2, 7, 63, "float")
    op.prepare_w_layers(nw, 1e12, 0.0, maxw)
    op.initialize_prediction(model, None)
    op.start_prediction_pass(0)
    v_o = op.sample(uvw_l)
    re, v = _gpu(w, h, dl, uvw, wl, vis, nw, maxw, precision, model, 0)
    e_img = rel(re, img_o)
    ok = ~np.isnan(v_o)
    e_v = rel(v[ok], v_o[ok])
    print(f"{w}x{h} {precision}: image rel L2 {e_img:.2e}, predict rel L2 {e_v:.2e}")
    assert e_img < TOL_IMG[precision]
    assert np.array_equal(np.isnan(v), ~ok)
    assert e_v < TOL_IMG[precision]
