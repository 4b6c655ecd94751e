"""Argument checks for every tensor whose raw device pointer is handed to the
C ABI. The kernels trust the sizes they are given: a tensor of the wrong
dtype, element count, layout or device would make them read or write past
its buffer. Each check raises ValueError before anything is launched.

The checks look only at tensor metadata (dtype, shape, strides, device), so
they run, and are tested, without a GPU.
"""
from __future__ import annotations


def _dev_index(t):
    import torch
    idx = t.device.index
    if idx is None and t.device.type == "cuda":
        idx = torch.cuda.current_device()
    return idx


def tensor(t, name, dtype, *, shape=None, numel=None, device=None, optional=False):
    """Checks one tensor: dtype, then shape or element count, then that it is
    contiguous, then that it lives on cuda:<device>. Returns t."""
    import torch
    if t is None:
        if optional:
            return None
        raise ValueError(f"{name} is required")
    if not isinstance(t, torch.Tensor):
        raise ValueError(f"{name} must be a torch tensor, got {type(t).__name__}")
    if t.dtype != dtype:
        raise ValueError(f"{name} has dtype {t.dtype}, expected {dtype}")
    if shape is not None and tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name} has shape {tuple(t.shape)}, expected {tuple(shape)}")
    if numel is not None and t.numel() != numel:
        raise ValueError(f"{name} has {t.numel()} elements, expected {numel}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if device is not None and (t.device.type != "cuda" or _dev_index(t) != device):
        raise ValueError(f"{name} is on {t.device}, expected cuda:{device}")
    return t


def plan_args(uvw_m, wavelengths, device):
    """uvw float64 [nrow, 3] metres, wavelengths float64 [nchan]; returns
    (nrow, nchan)."""
    import torch
    if uvw_m is None or getattr(uvw_m, "dim", lambda: 0)() != 2:
        raise ValueError("uvw_m must be a float64 tensor of shape [nrow, 3]")
    tensor(uvw_m, "uvw_m", torch.float64, shape=(uvw_m.shape[0], 3))
    if wavelengths is None or getattr(wavelengths, "dim", lambda: 0)() != 1:
        raise ValueError("wavelengths must be a float64 tensor of shape [nchan]")
    tensor(wavelengths, "wavelengths", torch.float64)
    on_device([("uvw_m", uvw_m), ("wavelengths", wavelengths)], device)
    return uvw_m.shape[0], wavelengths.shape[0]


def on_device(tensors, device):
    """Every (name, tensor) pair, tensor None allowed, on cuda:<device>."""
    for name, t in tensors:
        if t is not None and (t.device.type != "cuda" or _dev_index(t) != device):
            raise ValueError(f"{name} is on {t.device}, expected cuda:{device}")


def image_args(real, imag, n_images, height, width, has_complex, device, what="image"):
    """real (and imag) float32 [n_images, H, W]; imag is required iff an
    image is complex, and may be given otherwise."""
    import torch
    shp = (n_images, height, width)
    tensor(real, f"real {what}", torch.float32, shape=shp)
    tensor(imag, f"imaginary {what}", torch.float32, shape=shp, optional=not has_complex)
    on_device([(f"real {what}", real), (f"imaginary {what}", imag)], device)


def vis_args(vis, nrow, nchan, n_images, device, name="vis"):
    """complex64 [nrow, nchan, n_images] (any shape with that element count)."""
    import torch
    tensor(vis, name, torch.complex64, numel=nrow * nchan * n_images)
    on_device([(name, vis)], device)


def transfer_args(vis, real, imag, nrow, nchan, n_images, height, width, has_complex, device,
                  vis_name="vis", what="image", planned=True):
    """The tensors of one invert or predict: structure of all of them first
    (vis_args, image_args without the device), then their device."""
    import torch
    if not planned:
        raise ValueError("no plan: call plan() with the samples first")
    tensor(vis, vis_name, torch.complex64, numel=nrow * nchan * n_images)
    shp = (n_images, height, width)
    tensor(real, f"real {what}", torch.float32, shape=shp)
    tensor(imag, f"imaginary {what}", torch.float32, shape=shp, optional=not has_complex)
    on_device([(vis_name, vis), (f"real {what}", real), (f"imaginary {what}", imag)], device)


__all__ = ["image_args", "on_device", "plan_args", "tensor", "transfer_args",
           "vis_args"]
