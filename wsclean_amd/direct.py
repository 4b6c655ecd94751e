"""FP64 direct Fourier transform on the GPU (DirectMSGridder<double>,
gridding/directmsgridder.cpp:16-114, 175-196, and its adjoint predict).

Tensors live on a HIP device; work runs on torch's current stream.
uvw_m: float64 [nrow, 3]; wavelengths: float64 [nchan];
vis: complex64 [nrow, nchan] or [nrow, nchan, n_images] (select one image
with `image`); images: [H, W] (float64 out of invert, float32 into predict).
"""
from __future__ import annotations

import ctypes as C

from . import _lib


def _stream(t):
    import torch
    return C.c_void_p(torch.cuda.current_stream(t.device).cuda_stream)


def _vis_view(vis, image):
    if vis.dim() == 2:
        return vis, 1, 0
    return vis, vis.shape[2], image


def direct_invert(uvw_m, wavelengths, vis, width, height, dl, dm, l_shift=0.0, m_shift=0.0,
                  multiplication_factor=1.0, image=0, out=None):
    """image[y, x] = mult * sum_k Re(V_k exp(-2 pi i (u l + v m + w n'))), FP64."""
    import torch
    L = _lib.load()
    vis, stride, offset = _vis_view(vis, image)
    assert vis.dtype == torch.complex64 and vis.is_contiguous()
    assert uvw_m.dtype == torch.float64 and wavelengths.dtype == torch.float64
    nrow, nchan = uvw_m.shape[0], wavelengths.shape[0]
    assert vis.shape[0] == nrow and vis.shape[1] == nchan
    if out is None:
        out = torch.empty((height, width), dtype=torch.float64, device=vis.device)
    assert out.dtype == torch.float64 and out.is_contiguous() and out.numel() == width * height
    vptr = vis.data_ptr() + offset * 8
    _lib.check(L.hws_direct_invert_device(
        vis.device.index or 0, _stream(vis), width, height, dl, dm, l_shift, m_shift, nrow,
        nchan, C.c_void_p(uvw_m.data_ptr()), C.c_void_p(wavelengths.data_ptr()),
        C.c_void_p(vptr), stride, multiplication_factor, C.c_void_p(out.data_ptr())))
    return out


def direct_predict(model, uvw_m, wavelengths, dl, dm, l_shift=0.0, m_shift=0.0, out=None,
                   image=0):
    """V_k = sum_(x,y) I(x, y) exp(+2 pi i (u l + v m + w n')), FP64 sums."""
    import torch
    L = _lib.load()
    assert model.dtype == torch.float32 and model.is_contiguous() and model.dim() == 2
    height, width = model.shape
    nrow, nchan = uvw_m.shape[0], wavelengths.shape[0]
    if out is None:
        out = torch.empty((nrow, nchan), dtype=torch.complex64, device=model.device)
    out_, stride, offset = _vis_view(out, image)
    assert out_.is_contiguous() and out_.dtype == torch.complex64
    _lib.check(L.hws_direct_predict_device(
        model.device.index or 0, _stream(model), width, height, dl, dm, l_shift, m_shift,
        C.c_void_p(model.data_ptr()), nrow, nchan, C.c_void_p(uvw_m.data_ptr()),
        C.c_void_p(wavelengths.data_ptr()), C.c_void_p(out_.data_ptr() + offset * 8), stride))
    return out
