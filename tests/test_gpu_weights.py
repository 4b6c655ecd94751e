"""GPU imaging weights (hws_weights_*, hws_set_weighting) against the oracle's
restatement of ImageWeights (structures/imageweights.cpp:33-49, 75-144) and
of WSMSGridder::Invert with readAndWeightVisibilities
(gridding/msgridderbase.cpp:563-589, 640-659; gridding/wsmsgridder.cpp:436).

- Weight grids (natural, uniform, Briggs at several robustness values,
  weights as taper, super weight, several polarisations): bit-equal with the
  oracle when every partial sum is exact in FP64 (weights on a 1/64 lattice,
  whatever order the GPU adds them in), and to 1e-12 relative for arbitrary
  float weights (summation order only).
- A weighted invert: each image normalised by its own weight sum, the sums
  (total weight, visibility weight sum, maximum weight, gridded count) equal
  to the oracle's, images within the batched precision's tolerance (3e-5).
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

C_LIGHT = 299792458.0


def _rows(seed, nrow=600, nchan=5, frac=0.45, dl=0.003, wmax=80.0):
    rng = np.random.default_rng(seed)
    wl = C_LIGHT / np.linspace(120e6, 160e6, nchan)
    umax = frac / dl * wl.min()
    uvw = np.column_stack([rng.uniform(-umax, umax, nrow), rng.uniform(-umax, umax, nrow),
                           rng.uniform(-wmax, wmax, nrow)])
    # repeat some rows so that cells collect many weights
    uvw[nrow // 2:] = uvw[:nrow - nrow // 2] * rng.choice([1.0, -1.0], (nrow - nrow // 2, 1))
    return uvw, wl, dl


def _lattice_weights(rng, shape):
    """Weights on a 1/64 lattice in [0, 4] with some zeros: all sums exact."""
    w = rng.integers(0, 257, shape).astype(np.float32) / 64.0
    w[rng.uniform(size=shape) < 0.1] = 0.0
    return w


def _device_grid(mode, W, H, dl, uvw, wl, weights, **kw):
    from wsclean_amd.weights import DeviceImageWeights
    dev = torch.device("cuda:0")
    iw = DeviceImageWeights(mode, W, H, dl, dl, **kw)
    half = len(uvw) // 2  # two calls, as for two MeasurementSets
    for sl in (slice(0, half), slice(half, None)):
        iw.grid(torch.from_numpy(np.ascontiguousarray(uvw[sl])).to(dev),
                torch.from_numpy(wl).to(dev),
                torch.from_numpy(np.ascontiguousarray(weights[sl])).to(dev)
                if weights is not None else None)
    summed, total = iw.weight_grid()
    iw.finish()
    grid, _ = iw.weight_grid()
    return iw, summed, total, grid


def _oracle_grid(oracle, mode, W, H, dl, uvw, wl, weights, robust=0.0, taper=False, sw=1.0):
    iw = oracle.ImageWeights(mode, W, H, dl, dl, robust=robust, weights_as_taper=taper,
                             super_weight=sw)
    npol = weights.shape[2] if weights is not None else 1
    half = len(uvw) // 2
    for sl in (slice(0, half), slice(half, None)):
        iw.grid(uvw[sl], wl, weights[sl].reshape(-1) if weights is not None else None,
                npol=npol)
    summed = iw.grid_values().copy()
    iw.finish()
    return iw, summed, iw.grid_values()


@pytest.mark.parametrize("mode,robust,taper,sw,npol", [
    ("natural", 0.0, False, 1.0, 1), ("uniform", 0.0, False, 1.0, 2),
    ("briggs", 0.0, False, 1.0, 4), ("briggs", -1.5, False, 1.0, 1),
    ("briggs", 1.0, True, 1.0, 2), ("uniform", 0.0, False, 2.5, 1),
])
def test_weight_grid_bit_equal_with_exact_sums(oracle, mode, robust, taper, sw, npol):
    uvw, wl, dl = _rows(11)
    W, H = 128, 96
    rng = np.random.default_rng(12)
    weights = _lattice_weights(rng, (len(uvw), len(wl), npol))
    ow, o_summed, o_grid = _oracle_grid(oracle, mode, W, H, dl, uvw, wl, weights, robust, taper,
                                        sw)
    iw, summed, total, grid = _device_grid(mode, W, H, dl, uvw, wl, weights, robust=robust,
                                           weights_as_taper=taper, super_weight=sw)
    assert grid.shape == o_grid.shape
    assert np.array_equal(summed, o_summed), f"summed grid: {np.abs(summed - o_summed).max()}"
    assert np.array_equal(grid, o_grid), f"{mode}: {np.abs(grid - o_grid).max()}"
    assert total == float(o_summed.sum())  # exact sums: every order agrees
    print(f"{mode} R={robust} taper={taper} sw={sw} npol={npol}: {int((grid != 0).sum())} "
          f"weighted cells of {grid.size}, total {total}")
    iw.close()


def test_weight_grid_general_weights_to_rounding(oracle):
    uvw, wl, dl = _rows(13, nrow=2000)
    rng = np.random.default_rng(14)
    weights = rng.uniform(0.01, 3.0, (len(uvw), len(wl), 2)).astype(np.float32)
    for mode in ("uniform", "briggs"):
        _, o_summed, o_grid = _oracle_grid(oracle, mode, 96, 96, dl, uvw, wl, weights)
        iw, summed, _, grid = _device_grid(mode, 96, 96, dl, uvw, wl, weights)
        nz = o_summed != 0
        assert np.array_equal(summed != 0, nz)
        assert np.max(np.abs(summed - o_summed)[nz] / o_summed[nz]) < 1e-12
        assert np.max(np.abs(grid - o_grid)[nz] / np.abs(o_grid[nz])) < 1e-12
        iw.close()


def test_weights_reject_second_finish():
    from wsclean_amd._lib import HwsError
    from wsclean_amd.weights import DeviceImageWeights
    iw = DeviceImageWeights("uniform", 32, 32, 0.01, 0.01)
    iw.finish()
    with pytest.raises(HwsError, match="twice"):
        iw.finish()
    iw.close()


@pytest.mark.parametrize("mode,vis_mode", [("briggs", "normal"), ("uniform", "squared"),
                                           ("natural", "unit")])
def test_weighted_invert_matches_oracle(oracle, mode, vis_mode):
    """Two images (XX real, XY complex) with their own visibility weights,
    gridded with the device weights; each compared with the oracle's
    WSMSGridder invert of that correlation with the same ImageWeights."""
    import wsclean_amd
    from wsclean_amd.weights import DeviceImageWeights
    uvw, wl, dl = _rows(21, nrow=500, nchan=4)
    W = 128
    rng = np.random.default_rng(22)
    vis = (rng.standard_normal((len(uvw), len(wl), 2)) +
           1j * rng.standard_normal((len(uvw), len(wl), 2))).astype(np.complex64)
    vw = _lattice_weights(rng, (len(uvw), len(wl), 2))
    # the oracle grids ImageWeights from the weights of the first correlation
    # (one polarisation per ImageWeights here) and applies them to both
    ow = oracle.ImageWeights(mode, W, W, dl, dl)
    ow.grid(uvw, wl, np.ascontiguousarray(vw[..., 0]).reshape(-1), npol=1)
    ow.finish()
    dev = torch.device("cuda:0")
    d_uvw, d_wl = torch.from_numpy(uvw).to(dev), torch.from_numpy(wl).to(dev)
    iw = DeviceImageWeights(mode, W, W, dl, dl)
    iw.grid(d_uvw, d_wl, torch.from_numpy(np.ascontiguousarray(vw[..., :1])).to(dev))
    iw.finish()
    assert np.array_equal(iw.weight_grid()[0], ow.grid_values())
    vis_mode_code = {"normal": 0, "squared": 1, "unit": 2}[vis_mode]
    eff = vw if vis_mode == "normal" else (vw * vw if vis_mode == "squared"
                                           else (vw != 0).astype(np.float32))
    for k, cplx in ((0, False), (1, True)):
        s = oracle.ms_settings(W, W, dl, dl, is_complex=cplx)
        re_o, im_o, st = oracle.ms_invert(s, uvw, wl, np.ascontiguousarray(vis[..., k]),
                                          np.ascontiguousarray(eff[..., k]), image_weights=ow)
        g = wsclean_amd.DeviceGridder(W, W, dl, dl, 7, 63, precision="mixed", n_images=2,
                                      complex_images=(1,))
        g.prepare_w_layers(st.nw, st.min_w, st.max_w)
        g.set_weighting(iw, torch.from_numpy(vw).to(dev), vis_mode)
        g.plan(d_uvw, d_wl)
        re = torch.zeros((2, W, W), dtype=torch.float32, device=dev)
        im = torch.zeros_like(re)
        g.invert(torch.from_numpy(vis).to(dev), 1.0, re, im)
        sums = g.stats()
        g.close()
        # Briggs and uniform imaging weights are arbitrary doubles: their
        # products with the weights need not sum exactly in FP64
        assert abs(sums["total_weight"][k] - st.total_weight) <= 1e-12 * st.total_weight
        assert sums["visibility_weight_sum"][k] == st.visibility_weight_sum
        assert sums["max_gridded_weight"][k] == st.max_gridded_weight
        assert sums["gridded_visibility_count"][k] == st.gridded_count
        got = re[k].cpu().numpy().astype(np.float64)
        ref = re_o.astype(np.float64)
        if cplx:
            got = got + 1j * im[k].cpu().numpy()
            ref = ref + 1j * im_o
        e = np.linalg.norm(got - ref) / np.linalg.norm(ref)
        print(f"{mode}/{vis_mode} image {k} (complex={cplx}): rel {e:.2e}, sum w "
              f"{st.total_weight}, gridded {st.gridded_count}")
        assert e < 3e-5
    iw.close()
