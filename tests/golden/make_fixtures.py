"""Writes the small golden fixtures in tests/golden/.

The reference holds no vectors, known-answer tests or fixtures for the
w-stacking path (its only numeric checks need downloaded MeasurementSets), so
these fixtures come from closed forms evaluated here with numpy/scipy,
independently of the C++ oracle:
  kernel_kb.npz   Kaiser-Bessel x sinc kernel (alpha = 8.6) via scipy.special.i0,
                  the closed form of gridding/wstackinggridder.cpp:425-447;
  dct3.npz        REDFT01 of the half kernel (gridding/wstackinggridder.cpp:784-789)
                  via scipy.fft.dct(type=3), which is the same transform;
  dft_small.npz   brute-force FP64 direct Fourier transform invert
                  (gridding/directmsgridder.cpp:76-114 with the remap at :51-68)
                  and its adjoint (predict) on a 24 x 20 image, 40 samples.
Run: python tests/golden/make_fixtures.py
"""
import os

import numpy as np
import scipy.fft
import scipy.special

HERE = os.path.dirname(os.path.abspath(__file__))


def kb_kernel(os_, k, alpha=8.6, with_sinc=True):
    n = os_ * k
    mid = n // 2
    i = np.arange(mid + 1, dtype=np.float64)
    ratio = 1.0 / os_
    sinc = np.where(i == 0, ratio, np.sin(np.pi * ratio * i) / (np.pi * np.maximum(i, 1)))
    if not with_sinc:
        sinc = np.full_like(i, ratio)
    t = i / mid
    half = sinc * scipy.special.i0(alpha * np.sqrt(1 - t * t)) * os_ / scipy.special.i0(alpha)
    out = np.zeros(n)
    out[mid:] = half
    out[:mid] = out[n - 1:mid:-1]
    return out


def dft_invert(w, h, dl, dm, uvw, vis):
    """I(x, y) = Re sum_k V_k exp(-2 pi i (u l + v m + w n')) on image pixels."""
    x = np.arange(w)
    y = np.arange(h)
    l = (w // 2 - x) * dl
    m = (y - h // 2) * dm
    L, M = np.meshgrid(l, m)
    inside = L * L + M * M < 1
    n1 = np.where(inside, np.sqrt(np.maximum(1 - L * L - M * M, 0)) - 1, 0.0)
    ph = -2 * np.pi * (uvw[:, 0, None, None] * L + uvw[:, 1, None, None] * M +
                       uvw[:, 2, None, None] * n1)
    return (vis[:, None, None] * np.exp(1j * ph)).real.sum(axis=0), L, M, n1


def main():
    rng = np.random.default_rng(20240601)
    k63 = kb_kernel(63, 7)
    k1023 = kb_kernel(1023, 7)
    k63ns = kb_kernel(63, 7, with_sinc=False)
    np.savez_compressed(os.path.join(HERE, "kernel_kb.npz"), kb_63_7=k63, kb_1023_7=k1023,
                        kb_nosinc_63_7=k63ns)

    out = {}
    for width in (64, 96):
        n = width * 63 // 2
        x = np.zeros(n)
        x[: 7 * 63 // 2 + 1] = k63[7 * 63 // 2:]
        out[f"d_{width}"] = scipy.fft.dct(x, type=3)[: width // 2 + 1]
    np.savez_compressed(os.path.join(HERE, "dct3.npz"), **out)

    w, h = 24, 20
    dl, dm = 0.01, 0.012
    n = 40
    uvw = np.column_stack([rng.uniform(-40, 40, n), rng.uniform(-35, 35, n),
                           rng.uniform(-30, 30, n)])
    vis = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    img, L, M, n1 = dft_invert(w, h, dl, dm, uvw, vis.astype(np.complex128))
    model = rng.standard_normal((h, w))
    ph = 2 * np.pi * (uvw[:, 0, None, None] * L + uvw[:, 1, None, None] * M +
                      uvw[:, 2, None, None] * n1)
    pred = (model[None] * np.exp(1j * ph)).sum(axis=(1, 2))
    np.savez_compressed(os.path.join(HERE, "dft_small.npz"), width=w, height=h, dl=dl, dm=dm,
                        uvw=uvw, vis=vis, image=img, model=model, predicted=pred)


if __name__ == "__main__":
    main()
