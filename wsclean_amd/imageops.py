"""Small inversion and trimming on the GPU: WSMSGridder's resampling between
the inversion size and the image size and the centre-keeping trim/untrim
(gridding/wsmsgridder.cpp:452-484, 516-551), and the small-inversion size
rule of MSGridderBase::calculateOverallMetaData (gridding/msgridderbase.cpp:
409-439) with CalculateFFTSize (math/calculatefftsize.h:4-32).

Images are float32 torch tensors [n_images, H, W] (or [H, W]) on a HIP
device; work runs on torch's current stream.
"""
from __future__ import annotations

import ctypes as C
import math

from . import _checks, _lib


def _prep(images, what):
    import torch
    if images.dim() == 2:
        images = images.unsqueeze(0)
    if images.dim() != 3:
        raise ValueError(f"{what} must be [H, W] or [n_images, H, W]")
    dev = images.device.index if images.device.index is not None else 0
    _checks.tensor(images, what, torch.float32, device=dev)
    return images, dev


def _stream(dev):
    import torch
    return C.c_void_p(torch.cuda.current_stream(torch.device("cuda", dev)).cuda_stream)


def resample(images, out_width, out_height):
    """schaapcommon::fft::Resampler::Resample of every image to out_height x
    out_width (FP64 arithmetic, float32 result)."""
    import torch
    squeeze = images.dim() == 2
    x, dev = _prep(images, "images")
    out = torch.empty((x.shape[0], out_height, out_width), dtype=torch.float32, device=x.device)
    L = _lib.load()
    _lib.check(L.hws_resample_device(dev, _stream(dev), C.c_void_p(x.data_ptr()), x.shape[2],
                                     x.shape[1], C.c_void_p(out.data_ptr()), out_width,
                                     out_height, x.shape[0]))
    return out[0] if squeeze else out


def trim(images, trim_width, trim_height):
    """aocommon::Image::Trim: the trim_height x trim_width centre."""
    import torch
    squeeze = images.dim() == 2
    x, dev = _prep(images, "images")
    out = torch.empty((x.shape[0], trim_height, trim_width), dtype=torch.float32,
                      device=x.device)
    L = _lib.load()
    _lib.check(L.hws_trim_device(dev, _stream(dev), C.c_void_p(x.data_ptr()), x.shape[2],
                                 x.shape[1], C.c_void_p(out.data_ptr()), trim_width,
                                 trim_height, x.shape[0]))
    return out[0] if squeeze else out


def untrim(images, width, height):
    """aocommon::Image::Untrim: each image centred in a zeroed height x width."""
    import torch
    squeeze = images.dim() == 2
    x, dev = _prep(images, "images")
    out = torch.empty((x.shape[0], height, width), dtype=torch.float32, device=x.device)
    L = _lib.load()
    _lib.check(L.hws_untrim_device(dev, _stream(dev), C.c_void_p(x.data_ptr()), x.shape[2],
                                   x.shape[1], C.c_void_p(out.data_ptr()), width, height,
                                   x.shape[0]))
    return out[0] if squeeze else out


def _low_primes(n):
    # math/calculatefftsize.h:4-19
    while n > 7:
        if n % 2 == 0:
            n //= 2
        elif n % 3 == 0:
            n //= 3
        elif n % 5 == 0:
            n //= 5
        elif n % 7 == 0:
            n //= 7
        else:
            return False
    return True


def calculate_fft_size(size, pixel_scale, beam_size):
    """CalculateFFTSize (math/calculatefftsize.h:22-32): (minimal, optimal)
    image size that samples a beam of beam_size radians at 2 pixels."""
    total = size * pixel_scale
    minimal = int(math.ceil(total * 2 / beam_size))
    optimal = minimal
    if optimal % 4 != 0:
        optimal += 4 - optimal % 4
    while not _low_primes(optimal) and optimal < size:
        optimal += 4
    return minimal, min(optimal, size)


def small_inversion_size(width, height, pixel_x, pixel_y, beam_size):
    """The inversion size and pixel scales calculateOverallMetaData picks with
    small inversion on (gridding/msgridderbase.cpp:409-439): unchanged unless
    the beam is sampled by more than 2 pixels, else at least 32 pixels."""
    _, opt_w = calculate_fft_size(width, pixel_x, beam_size)
    _, opt_h = calculate_fft_size(height, pixel_y, beam_size)
    if opt_w < width or opt_h < height:
        nw = max(min(opt_w, width), 32)
        nh = max(min(opt_h, height), 32)
        return nw, nh, width * pixel_x / nw, height * pixel_y / nh
    return width, height, pixel_x, pixel_y


__all__ = ["calculate_fft_size", "resample", "small_inversion_size", "trim", "untrim"]
