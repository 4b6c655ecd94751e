"""The mixed precision's FP16 matrix-core grid kernel (gridMfmaF16Kernel:
operands split into FP16 high and low parts, FP32 sums) against the FP32
matrix-core kernel it replaces (selected with HWS_NO_F16_MFMA=1; at OS = 1023,
where the FP16 kernel reads its padded table from device memory, the VALU
cell kernel) and the oracle: the gridded layers of both agree to the FP32 kernel's own accuracy
(relative L2 2e-6, the layer tolerance of the parity tests), also for
visibilities spanning twelve decades and bins split into many work items,
and the image stays within the mixed precision's tolerance of the oracle."""
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

C_LIGHT = 299792458.0


def rel(a, b):
    a = np.asarray(a, np.complex128)
    b = np.asarray(b, np.complex128)
    return np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel())


def _layers(w, dl, uvw, wl, vis, nw, maxw, f16, chunk, complex_images=(1,), os_=63):
    import wsclean_amd
    old = os.environ.get("HWS_NO_F16_MFMA")
    os.environ["HWS_NO_F16_MFMA"] = "0" if f16 else "1"
    try:
        g = wsclean_amd.WStackingGridder(w, w, dl, dl, 7, os_, precision="mixed",
                                         n_images=vis.shape[2], complex_images=complex_images,
                                         work_chunk=chunk)
    finally:
        if old is None:
            del os.environ["HWS_NO_F16_MFMA"]
        else:
            os.environ["HWS_NO_F16_MFMA"] = old
    g.prepare_w_layers(nw, 0.0, 0.0, maxw)
    g.start_inversion_pass(0)
    g.add_data(uvw, wl, vis)
    lay = [[g.get_gridded_uv_layer(i, image=k) for i in range(nw)]
           for k in range(vis.shape[2])]
    g.finish_inversion_pass()
    re, im = g.finalize_image(1.0)
    g.close()
    return lay, re, im


# OS = 1023: the padded table exceeds the LDS, the FP16 kernel reads it from
# device memory, and the kernel it replaces is the VALU cell kernel
@pytest.mark.parametrize("spread,chunk,os_", [(0.0, 0, 63), (12.0, 0, 63), (3.0, 37, 63),
                                              (3.0, 0, 1023), (0.0, 37, 1023)])
def test_f16_grid_kernel_matches_fp32_kernel_and_oracle(oracle, spread, chunk, os_):
    rng = np.random.default_rng(int(spread * 10) + chunk)
    w, nrow, nchan, nw = 128, 1200, 4, 6
    dl = 0.004
    wl = C_LIGHT / np.linspace(120e6, 150e6, nchan)
    umax = 0.3 / dl * wl.min()
    uvw = np.column_stack([rng.uniform(-umax, umax, nrow), rng.uniform(-umax, umax, nrow),
                           rng.uniform(-50, 50, nrow)])
    # visibility magnitudes spread over `spread` decades (weighted data)
    mag = 10.0 ** rng.uniform(-spread / 2, spread / 2, (nrow, nchan, 2))
    vis = ((rng.standard_normal((nrow, nchan, 2)) + 1j * rng.standard_normal((nrow, nchan, 2)))
           * mag).astype(np.complex64)
    maxw = float(np.abs(uvw[:, 2]).max() / wl.min())
    lay16, re16, im16 = _layers(w, dl, uvw, wl, vis, nw, maxw, True, chunk, os_=os_)
    lay32, re32, im32 = _layers(w, dl, uvw, wl, vis, nw, maxw, False, chunk, os_=os_)
    worst = 0.0
    for k in range(2):
        for i in range(nw):
            if np.abs(lay32[k][i]).sum() == 0:
                assert np.abs(lay16[k][i]).sum() == 0
                continue
            worst = max(worst, rel(lay16[k][i], lay32[k][i]))
    uvw_l = (uvw[:, None, :] / wl[None, :, None]).reshape(-1, 3)
    g = oracle.RefGridder(w, w, dl, dl, 2, 7, os_, "float")
    g.set_is_complex(False)
    g.prepare_w_layers(nw, 1e12, 0.0, maxw)
    g.start_inversion_pass(0)
    g.add_samples(uvw_l, vis[:, :, 0].reshape(-1))
    lay_o = [g.layer(i) for i in range(nw)]
    g.finish_inversion_pass()
    g.finalize_image(1.0)
    img_o = g.image()
    worst_o = max(rel(lay16[0][i], lay_o[i]) for i in range(nw) if np.abs(lay_o[i]).sum() > 0)
    e_img = rel(re16[0], img_o)
    print(f"OS {os_}, spread {spread} decades, chunk {chunk}: FP16 vs FP32 kernel layers "
          f"{worst:.2e}; "
          f"FP16 layers vs oracle {worst_o:.2e}, image {e_img:.2e} (FP32 kernel image "
          f"{rel(re32[0], img_o):.2e})")
    assert worst < 2e-6
    assert worst_o < 2e-6
    assert e_img < 3e-5


def test_f16_grid_kernel_keeps_power_of_two_multiples_exact():
    # images that are positive power-of-two multiples of another (2 and 1/2
    # times the first) grid to exact multiples of its layers, also where
    # twelve decades of visibility magnitude push the small contributions into
    # FP16's subnormal range (each image takes its own power-of-two scale);
    # for a negative multiple the matrix cores' sums round asymmetrically
    # about zero, so that one agrees to the sums' rounding
    rng = np.random.default_rng(7)
    w, nrow, nchan, nw = 128, 900, 4, 4
    dl = 0.004
    wl = C_LIGHT / np.linspace(120e6, 150e6, nchan)
    umax = 0.3 / dl * wl.min()
    uvw = np.column_stack([rng.uniform(-umax, umax, nrow), rng.uniform(-umax, umax, nrow),
                           rng.uniform(-50, 50, nrow)])
    mag = 10.0 ** rng.uniform(-6, 6, (nrow, nchan))
    base = ((rng.standard_normal((nrow, nchan)) + 1j * rng.standard_normal((nrow, nchan)))
            * mag).astype(np.complex64)
    vis = np.stack([base, base * np.complex64(2.0), base * np.complex64(0.5),
                    base * np.complex64(-1.0)], axis=2)
    maxw = float(np.abs(uvw[:, 2]).max() / wl.min())
    lay, _, _ = _layers(w, dl, uvw, wl, vis, nw, maxw, True, 0, complex_images=())
    for i in range(nw):
        assert np.array_equal(lay[1][i], lay[0][i] * np.complex64(2.0)), i
        assert np.array_equal(lay[2][i], lay[0][i] * np.complex64(0.5)), i
        if np.abs(lay[0][i]).sum() > 0:
            assert rel(lay[3][i], -lay[0][i]) < 1e-6, i
