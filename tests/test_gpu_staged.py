"""The staged host path (hws_stage_rows / hws_invert_staged /
hws_predict_staged, what the MSGridderBase plugin calls) at a size where the
rows and the predicted visibilities cross the bus in several pinned pieces
(307 MB of visibilities: five 64 MB pieces each way, each copied by several
host threads), against the device path on the same samples:
- the predicted visibilities are equal bit for bit (degridding does not
  depend on how the plan orders a bin's records), NaN positions included;
- the images agree within 1e-6 of the peak (each path plans its samples
  itself, and the binning scatter may order a bin's records differently,
  which changes the gridded FP32 sums at rounding level)."""
import ctypes as C

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

C_LIGHT = 299792458.0


def test_staged_path_matches_device_path_across_pinned_pieces():
    import wsclean_amd
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(17)
    nrow, nchan, nimg, w = 150000, 64, 4, 512
    dl = 0.5 / w
    wl = C_LIGHT / np.linspace(120e6, 160e6, nchan)
    umax = 0.4 / dl * wl.min()
    uvw = np.column_stack([rng.uniform(-umax, umax, nrow), rng.uniform(-umax, umax, nrow),
                           rng.uniform(-60, 60, nrow)])
    vis = (rng.standard_normal((nrow, nchan, nimg)) +
           1j * rng.standard_normal((nrow, nchan, nimg))).astype(np.complex64)
    assert vis.nbytes > 4 * (64 << 20)
    maxw = float(np.abs(uvw[:, 2]).max() / wl.min())
    nw, cplx = 8, (2, 3)
    model_re = rng.standard_normal((nimg, w, w)).astype(np.float32)
    model_im = rng.standard_normal((nimg, w, w)).astype(np.float32)

    # staged: rows in batches of 40000 from host memory, results to host memory
    g = wsclean_amd.WStackingGridder(w, w, dl, dl, 7, 63, precision="mixed", n_images=nimg,
                                     complex_images=cplx)
    g.prepare_w_layers(nw, 0.0, 0.0, maxw)
    L, h = g.L, g.h
    pd, pf = C.POINTER(C.c_double), C.POINTER(C.c_float)
    uvw_c, wl_c = np.ascontiguousarray(uvw), np.ascontiguousarray(wl)

    def stage(with_vis):
        g._check(L.hws_stage_begin(h, nrow, nchan, wl_c.ctypes.data_as(pd), int(with_vis)))
        for r0 in range(0, nrow, 40000):
            r1 = min(nrow, r0 + 40000)
            g._check(L.hws_stage_rows(h, r1 - r0, uvw_c[r0:r1].ctypes.data_as(pd),
                                      vis[r0:r1].ctypes.data_as(pf) if with_vis else None))
        g._check(L.hws_stage_end(h))

    re_s = np.zeros((nimg, w, w), np.float32)
    im_s = np.zeros_like(re_s)
    stage(True)
    g._check(L.hws_invert_staged(h, 1.0, w, w, w, w, re_s.ctypes.data_as(pf),
                                 im_s.ctypes.data_as(pf)))
    out_s = np.empty_like(vis)
    stage(False)
    g._check(L.hws_predict_staged(h, model_re.ctypes.data_as(pf), model_im.ctypes.data_as(pf),
                                  w, w, w, w, out_s.ctypes.data_as(pf)))
    g.close()

    # device path on the same samples
    d = wsclean_amd.DeviceGridder(w, w, dl, dl, 7, 63, precision="mixed", n_images=nimg,
                                  complex_images=cplx)
    d.prepare_w_layers(nw, 0.0, maxw)
    d.plan(torch.from_numpy(uvw).to(dev), torch.from_numpy(wl).to(dev))
    re_d = torch.zeros((nimg, w, w), dtype=torch.float32, device=dev)
    im_d = torch.zeros_like(re_d)
    d.invert(torch.from_numpy(vis).to(dev), 1.0, re_d, im_d)
    out_d = torch.empty((nrow, nchan, nimg), dtype=torch.complex64, device=dev)
    d.predict(torch.from_numpy(model_re).to(dev), torch.from_numpy(model_im).to(dev), out_d)
    d.synchronize()
    re_d, im_d, out_d = re_d.cpu().numpy(), im_d.cpu().numpy(), out_d.cpu().numpy()
    d.close()

    peak = np.abs(re_d).max()
    assert peak > 0
    diff = np.abs(re_s - re_d).max() / peak
    diff_i = np.abs(im_s[list(cplx)] - im_d[list(cplx)]).max() / peak
    nan_s, nan_d = np.isnan(out_s), np.isnan(out_d)
    ok = ~nan_d
    print(f"staged vs device: image max diff {diff:.2e}, imag {diff_i:.2e} of the peak; "
          f"{int(ok.sum())} predicted, equal {np.array_equal(out_s[ok], out_d[ok])}")
    assert diff <= 1e-6 and diff_i <= 1e-6
    assert np.array_equal(nan_s, nan_d)
    assert ok.sum() > 0 and np.array_equal(out_s[ok], out_d[ok])
