"""The fused layer-transform kernels against the kernels they replace (real
and complex images, one pass and several, both batched precisions, every
power-of-two transform size the fused kernels cover and runtime-radix sizes
up to 6144); the replaced kernels are selected with HWS_NO_FUSED_PROJECT=1
when a context is created.
- predict: the forward rows pass over every layer per workgroup from the
  model (fftRowsModelLayersKernel) against four layers per workgroup
  (fftRowsFromModelKernel). Degridding does not depend on the order of a
  bin's records, so the visibilities show the transforms alone: in the
  reference precision they are equal bit for bit; in the mixed precision
  they agree within 1e-6 of the peak (measured: 1.5e-7, last-place
  differences of the FP32 phase evaluation between the two kernels).
- invert: the backward columns pass fused with the w-projection
  (fftColsProjectKernel) against the sparse columns pass and projectKernel.
  The transforms and the projection are the same operations, but each
  context plans its samples itself, and the binning scatter (atomics) may
  order a bin's records differently, which changes the gridded FP32 sums at
  rounding level: equal within 1e-6 of the peak."""
import os

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

C_LIGHT = 299792458.0


def _case(w, nrow, seed, nimg=4, frac=0.4):
    rng = np.random.default_rng(seed)
    dl = 0.5 / w
    wl = C_LIGHT / np.linspace(120e6, 160e6, 5)
    umax = frac / dl * wl.min()
    uvw = np.column_stack([rng.uniform(-umax, umax, nrow), rng.uniform(-umax, umax, nrow),
                           rng.uniform(-80, 80, nrow)])
    vis = (rng.standard_normal((nrow, len(wl), nimg)) +
           1j * rng.standard_normal((nrow, len(wl), nimg))).astype(np.complex64)
    maxw = float(np.abs(uvw[:, 2]).max() / wl.min())
    return uvw, wl, vis, dl, maxw


def _run(w, h, dl, uvw, wl, vis, maxw, precision, nw, max_mem, fused):
    """Invert of vis, then predict from a model derived from the image."""
    import wsclean_amd
    dev = torch.device("cuda:0")
    old = os.environ.get("HWS_NO_FUSED_PROJECT")
    os.environ["HWS_NO_FUSED_PROJECT"] = "0" if fused else "1"
    try:
        g = wsclean_amd.DeviceGridder(w, h, dl, dl, 7, 63, precision=precision,
                                      n_images=vis.shape[2], complex_images=(1, 2))
    finally:
        if old is None:
            del os.environ["HWS_NO_FUSED_PROJECT"]
        else:
            os.environ["HWS_NO_FUSED_PROJECT"] = old
    npass = g.prepare_w_layers(nw, 0.0, maxw, max_mem=max_mem)
    d_uvw, d_wl = torch.from_numpy(uvw).to(dev), torch.from_numpy(wl).to(dev)
    d_vis = torch.from_numpy(vis).to(dev)
    g.plan(d_uvw, d_wl)
    re = torch.zeros((vis.shape[2], h, w), dtype=torch.float32, device=dev)
    im = torch.zeros_like(re)
    g.invert(d_vis, 1.0, re, im)
    g.synchronize()
    # a seeded model, the same for both forms
    gen = torch.Generator(device=dev)
    gen.manual_seed(w * 7 + h)
    mre = torch.randn(re.shape, dtype=torch.float32, device=dev, generator=gen)
    mim = torch.randn(re.shape, dtype=torch.float32, device=dev, generator=gen)
    out = torch.empty_like(d_vis)
    g.predict(mre, mim, out)
    g.synchronize()
    res = re.cpu().numpy(), im.cpu().numpy(), out.cpu().numpy(), npass
    g.close()
    return res


@pytest.mark.parametrize("w,h,precision,nw,passes", [
    (256, 256, "mixed", 8, 1), (512, 256, "reference", 6, 1), (1024, 1024, "mixed", 12, 3),
    (2048, 2048, "mixed", 8, 1), (4096, 4096, "mixed", 6, 2), (1024, 512, "reference", 9, 2),
    # runtime-radix sizes (fftColsProjectGenericKernel, fftRowsModelLayersGenericKernel)
    (480, 240, "mixed", 8, 1), (960, 720, "reference", 6, 2), (1536, 1536, "mixed", 12, 3),
    (512, 6144, "mixed", 4, 1),
])
def test_fused_projection_equals_separate_kernels(w, h, precision, nw, passes):
    uvw, wl, vis, dl, maxw = _case(min(w, h), 400, seed=w + h + nw)
    # a layer budget that splits the layers into the wanted number of passes
    per_layer = w * h * 8 * vis.shape[2]
    max_mem = 0.0 if passes == 1 else per_layer * ((nw + passes - 1) // passes)
    re_f, im_f, v_f, n_f = _run(w, h, dl, uvw, wl, vis, maxw, precision, nw, max_mem, True)
    re_s, im_s, v_s, n_s = _run(w, h, dl, uvw, wl, vis, maxw, precision, nw, max_mem, False)
    assert n_f == n_s == passes
    assert np.isfinite(re_f).all() and np.abs(re_f).max() > 0
    same_re = np.array_equal(re_f, re_s)
    same_im = np.array_equal(im_f[1:3], im_s[1:3])
    d = np.abs(re_f - re_s).max() / np.abs(re_s).max()
    nan_same = np.array_equal(np.isnan(v_f), np.isnan(v_s))
    ok = ~np.isnan(v_s)
    same_v = nan_same and np.array_equal(v_f[ok], v_s[ok])
    dv = np.abs(v_f[ok] - v_s[ok]).max() / np.abs(v_s[ok]).max()
    print(f"{w}x{h} {precision} nw={nw} passes={passes}: image real equal {same_re}, imag "
          f"equal {same_im} (max rel diff {d:.2e}); predict equal {same_v} (NaN positions "
          f"equal {nan_same}, max rel diff {dv:.2e}, {int(ok.sum())} predicted)")
    assert nan_same
    if precision == "reference":
        assert same_v
    assert dv < 1e-6
    assert d < 1e-6
    assert np.abs(im_f[1:3] - im_s[1:3]).max() / np.abs(im_s[1:3]).max() < 1e-6
