"""Degridding as a differentiable torch operator.

`WStackingOperator` holds one planned device gridder for a fixed set of
(row, channel) samples. Its forward map A takes a real model image to
visibilities, the reference's Predict (gridding/wsmsgridder.cpp:488-586):
image -> w-layers (w-correction, forward FFT) -> degridding. Its adjoint A^H
takes visibilities to an image, the reference's Invert without the 1/sum(w)
normalisation (gridding/wsmsgridder.cpp:367-486): gridding -> inverse FFTs ->
w-correction -> kernel correction. The two are an exact adjoint pair of the
discretised operator (same kernel tables, same layers, same corrections), so

  for y = A x with x real and a real loss L(y), dL/dx = Re(A^H dL/dy*)

where dL/dy* is the gradient torch hands to backward for a complex output.
The backward is therefore one invert of the incoming gradient: the w-layers
are rebuilt from the visibilities, never stored between forward and backward.

Samples that the forward map cannot predict (off the uv grid; NaN in the
reference, gridding/wstackinggridder.cpp:626-627) are returned as 0 and get
no gradient: the adjoint drops them the same way the invert does.
"""
from __future__ import annotations

from .gridder import DeviceGridder


class WStackingOperator:
    """A (predict) and A^H (invert, unnormalised) on a fixed sample set.

    uvw_m: float64 [nrow, 3] and wavelengths: float64 [nchan] on the device;
    images are float32 [H, W]; visibilities complex64 [nrow, nchan].
    """

    def __init__(self, width, height, pixel_size_x, pixel_size_y, uvw_m, wavelengths,
                 n_w_layers, *, kernel_size=7, oversampling=63, kernel_mode="kb",
                 precision="mixed", l_shift=0.0, m_shift=0.0, w_limits=None, tile_size=0):
        import torch
        self.width, self.height = width, height
        self.device = uvw_m.device
        self.uvw, self.wl = uvw_m.contiguous(), wavelengths.contiguous()
        self.nrow, self.nchan = self.uvw.shape[0], self.wl.shape[0]
        if w_limits is None:
            aw = torch.abs(self.uvw[:, 2])
            w_limits = (float(aw.min() / self.wl.max()), float(aw.max() / self.wl.min()))
        self.g = DeviceGridder(width, height, pixel_size_x, pixel_size_y, kernel_size,
                               oversampling, kernel_mode=kernel_mode, precision=precision,
                               device=self.device.index or 0, n_images=1, l_shift=l_shift,
                               m_shift=m_shift, tile_size=tile_size,
                               stream=torch.cuda.current_stream(self.device))
        self.n_passes = self.g.prepare_w_layers(n_w_layers, *w_limits)
        self.g.plan(self.uvw, self.wl)

    def predict(self, image):
        import torch
        assert image.shape == (self.height, self.width)
        x = image.detach().to(torch.float32).contiguous()
        out = torch.empty((self.nrow, self.nchan, 1), dtype=torch.complex64, device=self.device)
        self.g.predict(x.reshape(1, self.height, self.width), None, out)
        return torch.nan_to_num(out.reshape(self.nrow, self.nchan), nan=0.0)

    def adjoint(self, vis):
        import torch
        v = vis.detach().to(torch.complex64).contiguous().reshape(self.nrow, self.nchan, 1)
        img = torch.empty((1, self.height, self.width), dtype=torch.float32,
                          device=self.device)
        self.g.invert(v, 1.0, img)
        return img.reshape(self.height, self.width)

    def close(self):
        self.g.close()


def _make_function():
    import torch

    class Degrid(torch.autograd.Function):
        """y = A x; backward: grad_x = Re(A^H grad_y) (one adjoint gridding)."""

        @staticmethod
        def forward(ctx, image, op):
            ctx.op = op
            return op.predict(image)

        @staticmethod
        def backward(ctx, grad_vis):
            return ctx.op.adjoint(grad_vis), None

    return Degrid


_DEGRID = None


def degrid(image, op: WStackingOperator):
    """Differentiable degridding of a float32 [H, W] model image."""
    global _DEGRID
    if _DEGRID is None:
        _DEGRID = _make_function()
    return _DEGRID.apply(image, op)


def chi_squared(vis_model, vis_observed, weights=None):
    """sum w |V_model - V_obs|^2, accumulated in FP64."""
    import torch
    r = (vis_model - vis_observed).to(torch.complex128)
    r2 = r.real * r.real + r.imag * r.imag
    if weights is not None:
        r2 = r2 * weights.to(torch.float64)
    return r2.sum()


def dot_product_test(op: WStackingOperator, seed=0):
    """<A x, y> against <x, A^H y> for random x (real) and y (complex):
    Re<A x, y> = <x, Re A^H y>. Returns (lhs, rhs)."""
    import torch
    gen = torch.Generator(device=op.device)
    gen.manual_seed(seed)
    x = torch.randn((op.height, op.width), dtype=torch.float32, device=op.device,
                    generator=gen)
    y = torch.randn((op.nrow, op.nchan), dtype=torch.complex64, device=op.device,
                    generator=gen)
    ax = op.predict(x).to(torch.complex128)
    ahy = op.adjoint(y).to(torch.float64)
    lhs = float(torch.real(torch.sum(ax * torch.conj(y.to(torch.complex128)))))
    rhs = float(torch.sum(x.to(torch.float64) * ahy))
    return lhs, rhs


__all__ = ["WStackingOperator", "degrid", "chi_squared", "dot_product_test"]
