"""FP64 direct Fourier transform on the GPU (DirectMSGridder<double>,
gridding/directmsgridder.cpp:16-114, 175-196, and its adjoint predict).

Tensors live on a HIP device; work runs on torch's current stream.
uvw_m: float64 [nrow, 3]; wavelengths: float64 [nchan];
vis: complex64 [nrow, nchan] or [nrow, nchan, n_images] (select one image
with `image`); images: [H, W] (float64 out of invert, float32 into predict).
"""
from __future__ import annotations

import ctypes as C

from . import _checks, _lib


def _stream(t):
    import torch
    return C.c_void_p(torch.cuda.current_stream(t.device).cuda_stream)


def _device_of(uvw_m):
    import torch
    if not isinstance(uvw_m, torch.Tensor) or uvw_m.device.type != "cuda":
        raise ValueError("uvw_m must be a tensor on a HIP device (expected cuda)")
    return _checks._dev_index(uvw_m)


def _vis_view(vis, image):
    if vis.dim() == 2:
        return vis, 1, 0
    return vis, vis.shape[2], image


def direct_invert(uvw_m, wavelengths, vis, width, height, dl, dm, l_shift=0.0, m_shift=0.0,
                  multiplication_factor=1.0, image=0, out=None):
    """image[y, x] = mult * sum_k Re(V_k exp(-2 pi i (u l + v m + w n'))), FP64."""
    import torch
    L = _lib.load()
    dev = _device_of(uvw_m)
    nrow, nchan = _checks.plan_args(uvw_m, wavelengths, dev)
    vis, stride, offset = _vis_view(vis, image)
    if vis.dim() not in (2, 3) or tuple(vis.shape[:2]) != (nrow, nchan):
        raise ValueError(f"vis must be [nrow, nchan] or [nrow, nchan, n_images] with "
                         f"nrow={nrow}, nchan={nchan}; got {tuple(vis.shape)}")
    if not 0 <= offset < stride:
        raise ValueError(f"image {image} out of range for {stride} images")
    _checks.tensor(vis, "vis", torch.complex64, device=dev)
    if out is None:
        out = torch.empty((height, width), dtype=torch.float64, device=vis.device)
    _checks.tensor(out, "out", torch.float64, numel=width * height, device=dev)
    vptr = vis.data_ptr() + offset * 8
    _lib.check(L.hws_direct_invert_device(
        dev, _stream(vis), width, height, dl, dm, l_shift, m_shift, nrow,
        nchan, C.c_void_p(uvw_m.data_ptr()), C.c_void_p(wavelengths.data_ptr()),
        C.c_void_p(vptr), stride, multiplication_factor, C.c_void_p(out.data_ptr())))
    return out


def direct_predict(model, uvw_m, wavelengths, dl, dm, l_shift=0.0, m_shift=0.0, out=None,
                   image=0):
    """V_k = sum_(x,y) I(x, y) exp(+2 pi i (u l + v m + w n')), FP64 sums."""
    import torch
    L = _lib.load()
    dev = _device_of(uvw_m)
    nrow, nchan = _checks.plan_args(uvw_m, wavelengths, dev)
    if model is None or model.dim() != 2:
        raise ValueError("model must be a float32 [H, W] tensor")
    _checks.tensor(model, "model", torch.float32, device=dev)
    height, width = model.shape
    if out is None:
        out = torch.empty((nrow, nchan), dtype=torch.complex64, device=model.device)
    out_, stride, offset = _vis_view(out, image)
    if out_.dim() not in (2, 3) or tuple(out_.shape[:2]) != (nrow, nchan):
        raise ValueError(f"out must be [nrow, nchan] or [nrow, nchan, n_images] with "
                         f"nrow={nrow}, nchan={nchan}; got {tuple(out_.shape)}")
    if not 0 <= offset < stride:
        raise ValueError(f"image {image} out of range for {stride} images")
    _checks.tensor(out_, "out", torch.complex64, device=dev)
    _lib.check(L.hws_direct_predict_device(
        dev, _stream(model), width, height, dl, dm, l_shift, m_shift,
        C.c_void_p(model.data_ptr()), nrow, nchan, C.c_void_p(uvw_m.data_ptr()),
        C.c_void_p(wavelengths.data_ptr()), C.c_void_p(out_.data_ptr() + offset * 8), stride))
    return out
