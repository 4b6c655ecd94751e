"""Python face of the MI355X w-stacking core.

``WStackingGridder`` mirrors the reference class WStackingGridder<T>
(gridding/wstackinggridder.h:92-635) method for method, over host numpy
arrays. ``DeviceGridder`` runs the batched one-call forms on torch tensors
that already live on the GPU (used by the benchmark and the autograd op).
Both call the C ABI of libhipwstacking.so; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _checks, _lib
from ._lib import HwsConfig, HwsStats, KERNEL_MODES, PRECISIONS, check

SPEED_OF_LIGHT = 299792458.0


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t)) if a is not None else None


def _cfg(width, height, dl, dm, kernel_size, oversampling, kernel_mode, precision,
         device, n_images, complex_mask, shift, rotate, tile_size, work_chunk=0):
    mode = KERNEL_MODES[kernel_mode] if isinstance(kernel_mode, str) else int(kernel_mode)
    prec = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
    return HwsConfig(width, height, dl, dm, kernel_size, oversampling, mode, prec, device,
                     n_images, complex_mask, shift[0], shift[1], int(bool(rotate)),
                     tile_size, work_chunk)


class _Context:
    def __init__(self, cfg: HwsConfig):
        self.L = _lib.load()
        h = C.c_void_p()
        st = self.L.hws_create(C.byref(cfg), C.byref(h))
        check(st, None)
        self.h = h
        self.cfg = cfg

    def close(self):
        if getattr(self, "h", None):
            self.L.hws_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _check(self, st):
        check(st, self.h)


def get_kernel(mode="kb", oversampling=63, size=7):
    """WStackingGridder::GetKernel (gridding/wstackinggridder.cpp:387-423)."""
    out = np.zeros(oversampling * size, np.float64)
    L = _lib.load()
    check(L.hws_get_kernel(KERNEL_MODES[mode], oversampling, size, _p(out, C.c_double)))
    return out


class WStackingGridder(_Context):
    """Host-array mirror of WStackingGridder<float>.

    width/height/pixel sizes as in the reference constructor
    (gridding/wstackinggridder.h:120-122). ``n_images`` images sharing uvw can
    be gridded at once; ``complex_images`` lists which of them are complex
    (SetIsComplex). Visibility arrays are complex64 [row][chan][image]
    (trailing image axis dropped when n_images == 1).
    """

    def __init__(self, width, height, pixel_size_x, pixel_size_y, kernel_size=7,
                 oversampling=1023, *, kernel_mode="kb", precision="reference", device=0,
                 n_images=1, complex_images=(), l_shift=0.0, m_shift=0.0,
                 rotate_visibilities=False, tile_size=0, work_chunk=0):
        mask = 0
        for i in complex_images:
            mask |= 1 << i
        super().__init__(_cfg(width, height, pixel_size_x, pixel_size_y, kernel_size,
                              oversampling, kernel_mode, precision, device, n_images, mask,
                              (l_shift, m_shift), rotate_visibilities, tile_size,
                              work_chunk))
        self.width, self.height, self.n_images = width, height, n_images
        self.complex_images = tuple(complex_images)
        self.n_passes = 0

    # ---- w-layers (gridding/wstackinggridder.h:158-267)
    def prepare_w_layers(self, n_w_layers, max_mem, min_w, max_w):
        n = C.c_size_t()
        self._check(self.L.hws_prepare_w_layers(self.h, n_w_layers, max_mem, min_w, max_w,
                                                C.byref(n)))
        self.n_passes = n.value
        return self.n_passes

    def w_to_layer(self, w, is_complex=False):
        r = self.L.hws_w_to_layer(self.h, w, int(is_complex))
        return None if r == 2**64 - 1 else r

    def layer_to_w(self, layer, is_complex=False):
        return self.L.hws_layer_to_w(self.h, layer, int(is_complex))

    def pass_layer_range(self, p):
        a, b = C.c_size_t(), C.c_size_t()
        self._check(self.L.hws_pass_layer_range(self.h, p, C.byref(a), C.byref(b)))
        return a.value, b.value

    # ---- inversion
    def start_inversion_pass(self, p):
        self._check(self.L.hws_start_inversion_pass(self.h, p))

    def _vis(self, vis, nrow, nchan):
        v = np.ascontiguousarray(vis, np.complex64).reshape(nrow, nchan, self.n_images)
        return v.view(np.float32)

    def add_data(self, uvw_m, wavelengths, vis):
        uvw = np.ascontiguousarray(uvw_m, np.float64).reshape(-1, 3)
        wl = np.ascontiguousarray(wavelengths, np.float64).reshape(-1)
        v = self._vis(vis, uvw.shape[0], wl.shape[0])
        self._check(self.L.hws_add_data(self.h, uvw.shape[0], wl.shape[0], _p(uvw, C.c_double),
                                        _p(wl, C.c_double), _p(v, C.c_float)))

    def add_samples(self, uvw_lambda, vis):
        uvw = np.ascontiguousarray(uvw_lambda, np.float64).reshape(-1, 3)
        v = self._vis(vis, uvw.shape[0], 1)
        self._check(self.L.hws_add_samples(self.h, uvw.shape[0], _p(uvw, C.c_double),
                                           _p(v, C.c_float)))

    def finish_inversion_pass(self):
        self._check(self.L.hws_finish_inversion_pass(self.h))

    def finalize_image(self, multiplication_factor):
        """Returns (real [n_images,H,W], imag or None)."""
        n = self.n_images * self.width * self.height
        re = np.zeros(n, np.float32)
        im = np.zeros(n, np.float32) if self.complex_images else None
        self._check(self.L.hws_finalize_image(self.h, multiplication_factor,
                                              _p(re, C.c_float), _p(im, C.c_float)))
        shape = (self.n_images, self.height, self.width)
        return re.reshape(shape), (im.reshape(shape) if im is not None else None)

    def get_gridded_uv_layer(self, layer, image=0):
        out = np.zeros(self.width * self.height * 2, np.float32)
        self._check(self.L.hws_get_gridded_uv_layer(self.h, image, layer, _p(out, C.c_float)))
        return out.view(np.complex64).reshape(self.height, self.width)

    # ---- prediction
    def initialize_prediction(self, real, imag=None):
        re = np.ascontiguousarray(real, np.float32).reshape(-1)
        im = np.ascontiguousarray(imag, np.float32).reshape(-1) if imag is not None else None
        self._check(self.L.hws_initialize_prediction(self.h, _p(re, C.c_float), _p(im, C.c_float)))

    def start_prediction_pass(self, p):
        self._check(self.L.hws_start_prediction_pass(self.h, p))

    def sample_data(self, uvw_m, wavelengths):
        uvw = np.ascontiguousarray(uvw_m, np.float64).reshape(-1, 3)
        wl = np.ascontiguousarray(wavelengths, np.float64).reshape(-1)
        out = np.zeros(uvw.shape[0] * wl.shape[0] * self.n_images * 2, np.float32)
        self._check(self.L.hws_sample_data(self.h, uvw.shape[0], wl.shape[0], _p(uvw, C.c_double),
                                           _p(wl, C.c_double), _p(out, C.c_float)))
        return out.view(np.complex64).reshape(uvw.shape[0], wl.shape[0], self.n_images)

    def sample_samples(self, uvw_lambda):
        uvw = np.ascontiguousarray(uvw_lambda, np.float64).reshape(-1, 3)
        out = np.zeros(uvw.shape[0] * self.n_images * 2, np.float32)
        self._check(self.L.hws_sample_samples(self.h, uvw.shape[0], _p(uvw, C.c_double),
                                              _p(out, C.c_float)))
        return out.view(np.complex64).reshape(uvw.shape[0], self.n_images)


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


class DeviceGridder(_Context):
    """Batched one-call invert/predict on torch CUDA (HIP) tensors, ordered on
    torch's current stream of `device` at construction (or on `stream`).

    uvw_m: float64 [nrow, 3]; wavelengths: float64 [nchan];
    vis: complex64 [nrow, nchan, n_images]; images: float32 [n_images, H, W].
    """

    def __init__(self, width, height, pixel_size_x, pixel_size_y, kernel_size=7,
                 oversampling=63, *, kernel_mode="kb", precision="mixed", device=0,
                 n_images=1, complex_images=(), l_shift=0.0, m_shift=0.0,
                 rotate_visibilities=True, tile_size=0, stream=None, work_chunk=0):
        mask = 0
        for i in complex_images:
            mask |= 1 << i
        super().__init__(_cfg(width, height, pixel_size_x, pixel_size_y, kernel_size,
                              oversampling, kernel_mode, precision, device, n_images, mask,
                              (l_shift, m_shift), rotate_visibilities, tile_size,
                              work_chunk))
        self.width, self.height, self.n_images = width, height, n_images
        self.complex_images = tuple(complex_images)
        self.device = device
        self.n_passes = 0
        self.nrow = self.nchan = 0
        self.planned = False
        # order the gridder's work with torch's on the same stream unless the
        # caller names another one
        if stream is None:
            import torch
            stream = torch.cuda.current_stream(torch.device("cuda", device))
        self.set_stream(stream)

    def set_stream(self, stream):
        self._check(self.L.hws_set_stream(self.h, C.c_void_p(stream.cuda_stream)))

    def prepare_w_layers(self, n_w_layers, min_w, max_w, max_mem=0.0):
        n = C.c_size_t()
        self._check(self.L.hws_prepare_w_layers(self.h, n_w_layers, max_mem, min_w, max_w,
                                                C.byref(n)))
        self.n_passes = n.value
        return self.n_passes

    def plan(self, uvw_m, wavelengths):
        nrow, nchan = _checks.plan_args(uvw_m, wavelengths, self.device)
        w = getattr(self, "_weighting", None)
        if w is not None and w[1] is not None and w[1].numel() != nrow * nchan * self.n_images:
            raise ValueError(f"the attached vis_weights hold {w[1].numel()} elements, the plan "
                             f"needs {nrow * nchan * self.n_images}")
        self._uvw, self._wl = uvw_m, wavelengths  # keep alive
        self.nrow, self.nchan = nrow, nchan
        self.planned = False
        self._check(self.L.hws_plan_device(self.h, self.nrow, self.nchan, _ptr(uvw_m),
                                           _ptr(wavelengths)))
        self.planned = True

    def check_invert_args(self, vis, real_out, imag_out):
        """Raises ValueError unless vis is complex64 with nrow * nchan *
        n_images elements and the images are float32 [n_images, H, W] (imag
        required iff an image is complex), all contiguous on this device."""
        _checks.transfer_args(vis, real_out, imag_out, self.nrow, self.nchan, self.n_images,
                              self.height, self.width, bool(self.complex_images), self.device,
                              planned=self.planned)

    def check_predict_args(self, real, imag, vis_out):
        _checks.transfer_args(vis_out, real, imag, self.nrow, self.nchan, self.n_images,
                              self.height, self.width, bool(self.complex_images), self.device,
                              vis_name="vis_out", what="model", planned=self.planned)

    def invert(self, vis, multiplication_factor, real_out, imag_out=None):
        self.check_invert_args(vis, real_out, imag_out)
        self._check(self.L.hws_invert_planned_device(self.h, _ptr(vis), multiplication_factor,
                                                     _ptr(real_out), _ptr(imag_out)))

    def predict(self, real, imag, vis_out):
        self.check_predict_args(real, imag, vis_out)
        self._check(self.L.hws_predict_planned_device(self.h, _ptr(real), _ptr(imag),
                                                      _ptr(vis_out)))

    def synchronize(self):
        self._check(self.L.hws_synchronize(self.h))

    def set_weighting(self, image_weights, vis_weights=None, mode="normal"):
        """Grid every visibility times visibility weight x imaging weight, as
        readAndWeightVisibilities weights them (gridding/msgridderbase.cpp:
        563-589, 640-659), and normalise each inverted image by its own weight
        sum (gridding/wsmsgridder.cpp:436). image_weights: a finished
        DeviceImageWeights on this device (None detaches); vis_weights:
        float32 [nrow, nchan, n_images] (None: 1); mode: "normal", "squared"
        or "unit". Takes effect from the next plan() (or invert() of the
        current plan); the sums are in stats()."""
        import torch
        from .weights import VISIBILITY_WEIGHTING
        if image_weights is None:
            if vis_weights is not None:
                raise ValueError("visibility weights need imaging weights")
            self._check(self.L.hws_set_weighting(self.h, None, None, 0))
            self._weighting = None
            return
        if image_weights.device != self.device:
            raise ValueError("imaging weights live on another device")
        if mode not in VISIBILITY_WEIGHTING:
            raise ValueError(f"visibility weighting mode must be one of {list(VISIBILITY_WEIGHTING)}")
        if vis_weights is not None:
            _checks.tensor(vis_weights, "vis_weights", torch.float32, device=self.device)
            if self.planned and vis_weights.numel() != self.nrow * self.nchan * self.n_images:
                raise ValueError(f"vis_weights has {vis_weights.numel()} elements, expected "
                                 f"{self.nrow * self.nchan * self.n_images}")
        self._check(self.L.hws_set_weighting(self.h, image_weights.h, _ptr(vis_weights),
                                             VISIBILITY_WEIGHTING[mode]))
        self._weighting = (image_weights, vis_weights)  # keep alive while attached

    def capture_invert(self, vis, multiplication_factor, real_out, imag_out=None):
        """A HIP graph of invert() on the planned samples (single pass);
        replay with .launch(). Valid until the next plan(), prepare_w_layers()
        or set_weighting(), or close(): launch() then raises HwsError."""
        self.check_invert_args(vis, real_out, imag_out)
        gr = C.c_void_p()
        self._check(self.L.hws_graph_capture_invert(self.h, _ptr(vis), multiplication_factor,
                                                    _ptr(real_out), _ptr(imag_out),
                                                    C.byref(gr)))
        return Graph(self, gr, (vis, real_out, imag_out))

    def capture_predict(self, real, imag, vis_out):
        """A HIP graph of predict() on the planned samples (single pass)."""
        self.check_predict_args(real, imag, vis_out)
        gr = C.c_void_p()
        self._check(self.L.hws_graph_capture_predict(self.h, _ptr(real), _ptr(imag),
                                                     _ptr(vis_out), C.byref(gr)))
        return Graph(self, gr, (real, imag, vis_out))

    def stats(self):
        s = HwsStats()
        self._check(self.L.hws_get_stats(self.h, C.byref(s)))
        out = {}
        for k, _ in HwsStats._fields_:
            v = getattr(s, k)
            out[k] = list(v) if hasattr(v, "__len__") else v
        return out


class Graph:
    """A captured invert or predict; launch() replays it on torch's current
    stream (or the given one)."""

    def __init__(self, owner, handle, keep):
        self.owner, self.h, self._keep = owner, handle, keep  # keep tensors alive

    def launch(self, stream=None):
        import torch
        st = stream if stream is not None else torch.cuda.current_stream(
            torch.device("cuda", self.owner.device))
        # graph errors are reported per thread, not on the context
        check(self.owner.L.hws_graph_launch(self.h, C.c_void_p(st.cuda_stream)), None)

    def close(self):
        if self.h:
            self.owner.L.hws_graph_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
