"""Facet geometry (wsclean_amd.facets), CPU only.

- The grid split is the reference's own test cases
  (tests/structures/tfacetutil.cpp:10-58).
- The facet shift is pinned by the direct Fourier transform: the oracle's
  FP64 DFT of the facets, each at its shift, sums to the DFT of the whole
  image (a sign or offset slip in the shift moves every facet's sources).
- The gridded facet size is checked against brute force.
"""
import numpy as np
import pytest

from wsclean_amd.facets import (create_facet_grid, cut_facet, facet_shift, facets_of_grid,
                                gridded_size, paste_facet)

C_LIGHT = 299792458.0


def test_single_facet_grid():
    # tfacetutil.cpp:10-25
    boxes = create_facet_grid(100, 142, 1, 1)
    assert boxes == [((0, 0, 100, 142), "0, 0")]


def test_multiple_facet_grid():
    # tfacetutil.cpp:27-56: every cell of a 4 x 5 grid over 100 x 100, with
    # the box edges at grid_x * size / grid_width
    size, gw, gh = 100, 4, 5
    boxes = create_facet_grid(size, size, gw, gh)
    assert len(boxes) == gw * gh
    cells = set()
    for (x0, y0, x1, y1), label in boxes:
        cx, cy = (x0 + x1) // 2, (y0 + y1) // 2
        gx, gy = cx // (x1 - x0), cy // (y1 - y0)
        assert 0 <= gx < gw and 0 <= gy < gh
        cells.add((gx, gy))
        assert (x0, y0) == (gx * size // gw, gy * size // gh)
        assert (x1, y1) == ((gx + 1) * size // gw, (gy + 1) * size // gh)
        assert label == f"{gx}, {gy}"
    assert len(cells) == gw * gh


def test_facet_boxes_tile_the_image():
    boxes = create_facet_grid(257, 130, 3, 4)
    cover = np.zeros((130, 257), int)
    for (x0, y0, x1, y1), _ in boxes:
        cover[y0:y1, x0:x1] += 1
    assert (cover == 1).all()


def _brute_size(n, padding, multiple):
    m = int(np.ceil(n * padding - 1e-9))
    while True:
        if m % multiple == 0:
            r = m
            for p in (2, 3, 5, 7):
                while r % p == 0:
                    r //= p
            if r == 1:
                return m
        m += 1


@pytest.mark.parametrize("n,padding,multiple", [(85, 1.0, 32), (100, 1.2, 32), (511, 1.0, 64),
                                                (1000, 1.0, 16), (64, 1.0, 32), (2049, 1.0, 4),
                                                (330, 1.5, 32)])
def test_gridded_size_is_smallest_smooth_multiple(n, padding, multiple):
    assert gridded_size(n, padding, multiple) == _brute_size(n, padding, multiple)


@pytest.mark.parametrize("grid", [(2, 2), (3, 2)])
def test_facet_dfts_sum_to_the_full_dft(oracle, grid):
    rng = np.random.default_rng(41)
    W, H, dl = 96, 64, 0.01
    model = np.zeros((H, W))
    for _ in range(7):
        model[rng.integers(0, H), rng.integers(0, W)] = rng.uniform(0.2, 2.0)
    wl = C_LIGHT / np.array([120e6, 150e6])
    umax = 0.4 / dl * wl.min()
    uvw = np.column_stack([rng.uniform(-umax, umax, 40), rng.uniform(-umax, umax, 40),
                           rng.uniform(-30, 30, 40)])
    uvw_l = (uvw[:, None, :] / wl[None, :, None]).reshape(-1, 3)
    full = oracle.dft_predict(W, H, dl, dl, uvw_l, model, shift=(0.003, -0.002))
    total = np.zeros_like(full)
    for box, _ in create_facet_grid(W, H, *grid):
        x0, y0, x1, y1 = box
        fw, fh = gridded_size(x1 - x0, 1.0, 8), gridded_size(y1 - y0, 1.0, 8)
        ls, ms = facet_shift(box, W, H, dl, dl, 0.003, -0.002)
        sub = cut_facet(model, box, fw, fh)
        assert sub.sum() == model[y0:y1, x0:x1].sum()
        total += oracle.dft_predict(fw, fh, dl, dl, uvw_l, sub, shift=(ls, ms))
    err = np.linalg.norm(total - full) / np.linalg.norm(full)
    assert err < 1e-12, err


def test_facet_dirty_images_paste_into_the_full_dirty_image(oracle):
    # the DFT dirty image of each facet at its shift, trimmed to its box and
    # pasted, is the DFT dirty image of the whole field
    rng = np.random.default_rng(43)
    W, H, dl = 64, 48, 0.01
    wl = C_LIGHT / np.array([130e6])
    umax = 0.4 / dl * wl.min()
    uvw = np.column_stack([rng.uniform(-umax, umax, 30), rng.uniform(-umax, umax, 30),
                           rng.uniform(-30, 30, 30)])
    uvw_l = uvw / wl[0]
    vis = (rng.standard_normal(30) + 1j * rng.standard_normal(30)).astype(np.complex64)
    full = oracle.dft_invert(W, H, dl, dl, uvw_l, vis)
    pasted = np.full((H, W), np.nan)
    for f in facets_of_grid(W, H, 3, 2, dl, dl, multiple=8):
        img = oracle.dft_invert(f.width, f.height, dl, dl, uvw_l, vis, shift=(f.l_shift, f.m_shift))
        paste_facet(pasted, img, f.box)
    assert np.abs(pasted - full).max() < 1e-12 * np.abs(full).max()
