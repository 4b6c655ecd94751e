"""Autotuner for the batched gridder: uv tile size, work-item size and
w-layer count for one set of samples on one device.

- w-layers: the lower bound is the reference's suggestion
  (WSMSGridder::getSuggestedWGridSize, gridding/wsmsgridder.cpp:77-128):
  nW = ceil(f 2 pi (maxW - cMinW) (1 - sqrt(1 - lmax^2 - mmax^2))) with
  lmax = W dl / 2 + |l_shift| (and likewise m) and cMinW = -maxW for complex
  images, minW otherwise. More layers never run faster,
  so the tuner keeps the larger of that bound and the caller's minimum, and
  reports how many passes the device memory allows.
- with max_image_error, the w-layer count is instead the smallest one from
  that bound up whose dirty image of a random subset of the caller's samples
  is within max_image_error (RMS over the central half, relative to the
  peak) of the FP64 direct transform of the same samples
  (accuracy.w_layers_for_accuracy): the w-stacking approximation's error
  falls as the layers get closer.
- tile size and work chunk (0 = the library's own rule from the sample
  count): every candidate pair is timed (plan + invert +
  predict on the caller's samples, best of `repeats` after one warm-up, host
  clock around device-synchronised runs), and the fastest is chosen.
"""
from __future__ import annotations

import math
import time

from .gridder import DeviceGridder


def suggested_w_layers(width, height, dl, dm, min_w, max_w, l_shift=0.0, m_shift=0.0,
                       factor=1.0):
    """The reference's w-grid size suggestion (without its NCpus floor,
    which has no meaning on one GPU)."""
    lmax = width * dl / 2.0 + abs(l_shift)
    mmax = height * dm / 2.0 + abs(m_shift)
    r2 = lmax * lmax + mmax * mmax
    one_minus = 1.0 - math.sqrt(1.0 - r2) if r2 < 1.0 else 1.0
    return max(1, int(math.ceil(factor * 2.0 * math.pi * (max_w - min_w) * one_minus)))


def valid_tiles(width, height, kernel_size=7, candidates=(16, 32)):
    """Tile edges the library accepts: multiples of 4 dividing both sizes,
    >= K - 1, with T + K - 1 <= 64."""
    return [t for t in candidates
            if t % 4 == 0 and width % t == 0 and height % t == 0 and t >= kernel_size - 1
            and t + kernel_size - 1 <= 64]


def autotune(width, height, dl, dm, uvw_m, wavelengths, vis, *, n_images=1, complex_images=(),
             precision="mixed", kernel_size=7, oversampling=63, min_layers=1, l_shift=0.0,
             m_shift=0.0, tiles=(16, 32), chunks=(4096, 8192, 16384), repeats=2,
             max_image_error=None, accuracy_samples=20_000, max_w_layers=1024):
    """Returns a dict with the chosen tile_size, work_chunk and n_w_layers, the
    number of passes, and the timings (ms per plan + invert + predict) of
    every candidate. uvw_m, wavelengths and vis are device tensors in the
    layout DeviceGridder takes. With max_image_error, n_w_layers is the
    smallest count meeting it (ValueError if none up to max_w_layers does),
    and "accuracy" lists the (n_w_layers, error) pairs tried."""
    import torch
    aw = torch.abs(uvw_m[:, 2])
    min_w = float(aw.min() / wavelengths.max())
    max_w = float(aw.max() / wavelengths.min())
    if not complex_images:
        n_w = suggested_w_layers(width, height, dl, dm, min_w, max_w, l_shift, m_shift)
    else:
        # complex images span [-maxW, maxW]
        n_w = suggested_w_layers(width, height, dl, dm, -max_w, max_w, l_shift, m_shift)
    n_w = max(n_w, int(min_layers))
    accuracy_trials = None
    if max_image_error is not None:
        from .accuracy import w_layers_for_accuracy
        if complex_images:
            raise ValueError("the accuracy search grids a real image: pass the visibilities "
                             "of a real correlation")
        found, accuracy_trials = w_layers_for_accuracy(
            width, height, dl, dm, uvw_m, wavelengths, vis, max_image_error, n_w_start=n_w,
            n_w_max=max_w_layers, oversampling=oversampling, kernel_size=kernel_size,
            precision=precision, n_samples=accuracy_samples)
        if found is None:
            raise ValueError(f"no w-layer count up to {max_w_layers} meets an image error of "
                             f"{max_image_error}: {accuracy_trials}")
        n_w = found
    dev = vis.device
    real = torch.zeros((n_images, height, width), dtype=torch.float32, device=dev)
    imag = torch.zeros_like(real) if complex_images else None
    out = torch.empty_like(vis)
    trials = []
    passes = None
    for tile in valid_tiles(width, height, kernel_size, tiles):
        for chunk in chunks:
            g = DeviceGridder(width, height, dl, dm, kernel_size, oversampling,
                              precision=precision, n_images=n_images,
                              complex_images=complex_images, l_shift=l_shift, m_shift=m_shift,
                              tile_size=tile, work_chunk=chunk)
            passes = g.prepare_w_layers(n_w, min_w, max_w)
            best = float("inf")
            for it in range(repeats + 1):
                g.synchronize()
                t0 = time.perf_counter()
                g.plan(uvw_m, wavelengths)
                g.invert(vis, 1.0, real, imag)
                g.predict(real, imag, out)
                g.synchronize()
                if it > 0:
                    best = min(best, (time.perf_counter() - t0) * 1e3)
            g.close()
            trials.append({"tile_size": tile, "work_chunk": chunk, "ms": best})
    if not trials:
        raise ValueError("no uv tile size fits the image size")
    pick = min(trials, key=lambda t: t["ms"])
    return {"tile_size": pick["tile_size"], "work_chunk": pick["work_chunk"],
            "n_w_layers": n_w, "passes": passes, "ms": pick["ms"], "trials": trials,
            "w_range": (min_w, max_w), "accuracy": accuracy_trials}


__all__ = ["autotune", "suggested_w_layers", "valid_tiles"]
