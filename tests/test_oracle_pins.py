"""Closed-form pins of the oracle's host-side restatements that the GPU
path's parity rests on: MSGridderBase::calculateWLimits with the wLimit trim
of calculateOverallMetaData (gridding/msgridderbase.cpp:200-286, 397-400),
and the centre-keeping image trim/untrim WSMSGridder applies
(gridding/wsmsgridder.cpp:476-484, 516-520).

Every expected value is fixed by construction: a 64 x 64 grid with pixel
size 1/4096 rad maps u (wavelengths) to x = u / 64 exactly, so the uv window
floor(x) > -32 and ceil(x) < 32 of calculateWLimits admits exactly
-31 <= x <= 31. Each case adds one row whose inclusion or exclusion moves the
minimum or maximum |w| away from the two base rows' values (5 and 7).
"""
import math

import numpy as np
import pytest

W = 64
DL = 1.0 / 4096.0
BASE = [(640.0, 320.0, 5.0), (-1280.0, -640.0, -7.0)]


def _limits(oracle, rows, wavelengths=(1.0,), weights=None, grid_rows=None):
    uvw = np.array(rows, np.float64)
    wl = np.array(wavelengths, np.float64)
    s = oracle.ms_settings(W, W, DL, DL)
    iw = oracle.ImageWeights("natural", W, W, DL, DL)
    g = np.array(grid_rows if grid_rows is not None else rows, np.float64)
    iw.grid(g, wl)
    iw.finish()
    return oracle.calculate_w_limits(s, uvw, wl, iw, weights), iw


@pytest.mark.parametrize("row,in_window", [
    ((1984.0, 0.0, 9.0), True),            # x = 31: ceil(31) < 32
    ((1984.0625, 0.0, 9.0), False),        # x = 31.001: ceil = 32
    ((-1984.0, 0.0, 9.0), True),           # x = -31: floor(-31) > -32
    ((-1984.0625, 0.0, 9.0), False),       # floor = -32
    ((0.0, 1984.0, 9.0), True),
    ((0.0, 1984.0625, 9.0), False),
    ((0.0, -1984.0, -9.0), True),
    ((0.0, -1984.0625, -9.0), False),
])
def test_w_limits_uv_window_edges(oracle, row, in_window):
    (lo, hi, bl), _ = _limits(oracle, BASE + [row])
    assert lo == 5.0
    assert hi == (9.0 if in_window else 7.0)
    b_base = math.sqrt(1280.0 ** 2 + 640.0 ** 2 + 49.0)
    b_row = math.sqrt(row[0] ** 2 + row[1] ** 2 + row[2] ** 2)
    assert bl == (max(b_base, b_row) if in_window else b_base)


def test_w_limits_base_rows(oracle):
    (lo, hi, bl), _ = _limits(oracle, BASE)
    assert (lo, hi) == (5.0, 7.0)
    assert bl == math.sqrt(1280.0 ** 2 + 640.0 ** 2 + 49.0)
    (lo, _, _), _ = _limits(oracle, BASE + [(320.0, 64.0, 0.25)])
    assert lo == 0.25


def test_w_limits_skip_zero_visibility_weight(oracle):
    rows = BASE + [(320.0, 64.0, 0.5), (320.0, 128.0, 50.0)]
    weights = np.array([1, 1, 0, 0], np.float32).reshape(4, 1)
    (lo, hi, _), _ = _limits(oracle, rows, weights=weights)
    assert (lo, hi) == (5.0, 7.0)
    weights[2] = 0.5
    (lo, hi, _), _ = _limits(oracle, rows, weights=weights)
    assert (lo, hi) == (0.5, 7.0)


def test_w_limits_skip_zero_imaging_weight(oracle):
    odd = (320.0, 1280.0, 0.5)     # cell (37, 20) of the weight grid
    rows = BASE + [odd]
    (lo, hi, _), iw = _limits(oracle, rows, grid_rows=BASE)
    assert iw.get_weight(odd[0], odd[1]) == 0.0   # precondition: empty cell
    assert (lo, hi) == (5.0, 7.0)
    (lo, hi, _), _ = _limits(oracle, rows)          # gridded: weight 1
    assert (lo, hi) == (0.5, 7.0)


def test_w_limits_all_flagged_give_zero(oracle):
    weights = np.zeros((2, 1), np.float32)
    (lo, hi, bl), _ = _limits(oracle, BASE, weights=weights)
    assert (lo, hi, bl) == (0.0, 0.0, 0.0)


def test_w_limits_per_channel_wavelength(oracle):
    """Channel 0 (lambda 1 m) of the extra row is outside the window
    (x = 31.25), channel 1 (lambda 2 m) inside (x = 15.625): only w / 2
    counts. Base rows contribute |w| and |w| / 2."""
    rows = BASE + [(2000.0, 0.0, 100.0)]
    (lo, hi, bl), _ = _limits(oracle, rows, wavelengths=(1.0, 2.0))
    assert (lo, hi) == (2.5, 50.0)
    assert bl == max(math.sqrt(1280.0 ** 2 + 640.0 ** 2 + 49.0),
                     math.sqrt(2000.0 ** 2 + 100.0 ** 2) / 2.0)


@pytest.mark.parametrize("w_limit,expected_max", [(0.0, 7.0), (10.0, 7.0 * (1.0 - 0.1)),
                                                  (50.0, 5.0)])
def test_w_limit_trims_max_w(oracle, w_limit, expected_max):
    """calculateOverallMetaData: maxW *= 1 - wLimit/100, not below minW."""
    uvw = np.array(BASE, np.float64)
    wl = np.array([1.0])
    vis = np.ones((2, 1), np.complex64)
    s = oracle.ms_settings(W, W, DL, DL, nw=4, w_limit=w_limit)
    _, _, st = oracle.ms_invert(s, uvw, wl, vis)
    assert st.min_w == 5.0
    assert st.max_w == expected_max


def test_trim_keeps_the_phase_centre_pixel(oracle):
    """Trimming keeps each pixel's (l, m): the FP64 direct transform at the
    trimmed size equals the trimmed direct transform of the full size,
    square and not."""
    rng = np.random.default_rng(7)
    n = 40
    uvw = np.column_stack([rng.uniform(-900, 900, n), rng.uniform(-900, 900, n),
                           rng.uniform(-50, 50, n)])
    vis = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    dl = 2e-4
    for (w, h, tw, th) in [(64, 64, 32, 32), (64, 48, 40, 16), (96, 64, 64, 64)]:
        full = oracle.dft_invert(w, h, dl, dl, uvw, vis, nthreads=1).astype(np.float32)
        small = oracle.dft_invert(tw, th, dl, dl, uvw, vis, nthreads=1).astype(np.float32)
        assert np.array_equal(oracle.trim_image(full, tw, th), small), (w, h, tw, th)
        un = oracle.untrim_image(small, w, h)
        assert np.array_equal(oracle.trim_image(un, tw, th), small)
        sx, sy = (w - tw) // 2, (h - th) // 2
        assert un[h // 2, w // 2] == small[th // 2, tw // 2]
        mask = np.ones((h, w), bool)
        mask[sy:sy + th, sx:sx + tw] = False
        assert not un[mask].any()


# ---- the small-inversion resampler (schaapcommon::fft::Resampler, not
# vendored in the reference tree; its published algorithm restated in
# oracle/msref.cpp::Resample). Pinned by closed forms: a constant stays
# constant, and an image that is a trigonometric polynomial with frequencies
# below both grids' Nyquist frequencies comes out as the same function
# sampled on the new grid. Content on a Nyquist line (where the
# real-to-complex copy keeps one of the two half-weight terms) is not
# pinned by any reference vector or closed form.

def _trig_image(terms, w, h):
    y, x = np.mgrid[0:h, 0:w]
    img = np.zeros((h, w))
    for a, fx, fy, ph in terms:
        img += a * np.cos(2 * np.pi * (fx * x / w + fy * y / h) + ph)
    return img


@pytest.mark.parametrize("in_w,in_h,out_w,out_h", [(32, 32, 64, 64), (64, 48, 40, 32),
                                                   (36, 40, 96, 72), (128, 128, 60, 100)])
def test_resample_constant_and_band_limited(oracle, in_w, in_h, out_w, out_h):
    out = oracle.resample(np.full((in_h, in_w), 2.5), out_w, out_h)
    assert np.max(np.abs(out - 2.5)) < 1e-13
    rng = np.random.default_rng(in_w * out_w + in_h)
    lim_x, lim_y = min(in_w, out_w) // 2, min(in_h, out_h) // 2
    terms = [(rng.uniform(0.2, 2.0), int(rng.integers(-lim_x + 1, lim_x)),
              int(rng.integers(-lim_y + 1, lim_y)), rng.uniform(0, 2 * np.pi)) for _ in range(8)]
    terms.append((1.0, lim_x - 1, -(lim_y - 1), 0.3))  # the highest kept frequencies
    got = oracle.resample(_trig_image(terms, in_w, in_h), out_w, out_h)
    want = _trig_image(terms, out_w, out_h)
    assert np.max(np.abs(got - want)) < 1e-11 * sum(abs(t[0]) for t in terms)
