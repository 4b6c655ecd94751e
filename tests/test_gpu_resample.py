"""Small inversion and trimming on the GPU (hws_resample_device,
hws_trim_device, hws_untrim_device) against the oracle's restatement of the
resampler (oracle/msref.cpp::Resample, itself pinned by closed forms in
tests/test_oracle_pins.py) and of Image::Trim/Untrim.

The GPU evaluates the resampler as two FP64 resampling matrices (closed-form
Dirichlet sums) applied with matrix products; the oracle runs the FFT, copy
and inverse FFT. Images here carry content at every frequency, Nyquist lines
included, so the two forms agree only if the copy rule is the same. Output is
float32: tolerance 2e-7 of the image's largest value."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("in_w,in_h,out_w,out_h", [
    (64, 64, 128, 128), (128, 128, 64, 64), (48, 64, 96, 80), (200, 120, 64, 96),
    (96, 96, 96, 96), (512, 384, 1024, 768)])
def test_resample_matches_oracle(oracle, in_w, in_h, out_w, out_h):
    from wsclean_amd import imageops
    rng = np.random.default_rng(in_w * 1000 + out_h)
    imgs = rng.standard_normal((2, in_h, in_w)).astype(np.float32)
    got = imageops.resample(torch.from_numpy(imgs).cuda(), out_w, out_h).cpu().numpy()
    for k in range(2):
        ref = oracle.resample(imgs[k].astype(np.float64), out_w, out_h)
        err = np.max(np.abs(got[k] - ref)) / np.max(np.abs(ref))
        print(f"{in_w}x{in_h} -> {out_w}x{out_h} image {k}: max rel {err:.2e}")
        assert err < 2e-7


def test_trim_untrim_match_oracle(oracle):
    from wsclean_amd import imageops
    rng = np.random.default_rng(5)
    imgs = rng.standard_normal((3, 96, 128)).astype(np.float32)
    d = torch.from_numpy(imgs).cuda()
    t = imageops.trim(d, 64, 40).cpu().numpy()
    u = imageops.untrim(torch.from_numpy(t).cuda(), 128, 96).cpu().numpy()
    for k in range(3):
        assert np.array_equal(t[k], oracle.trim_image(imgs[k], 64, 40))
        assert np.array_equal(u[k], oracle.untrim_image(t[k], 128, 96))
