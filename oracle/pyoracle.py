"""ORACLE / TEST INFRASTRUCTURE ONLY.

ctypes wrapper over oracle/_build/liboracle.so (the CPU restatement of the
reference w-stacking gridder, its WSMSGridder host pipeline and the FP64
direct-FT oracle; see oracle/oracle.h). Only tests/, __graft_entry__.smoke()
and bench.py's cpu_baseline leg import this module, and only as the checker.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

KERNEL_MODES = {
    "nn": 0, "kb": 1, "kb_nosinc": 2, "rect": 3, "gaussian": 4,
    "gaussian_nosinc": 5, "bn": 6, "bn_nosinc": 7, "bh": 8,
}


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = _setup(C.CDLL(_LIB_PATH))
    return _lib


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t)) if a is not None else None


class MsSettings(C.Structure):
    _fields_ = [
        ("width", C.c_size_t), ("height", C.c_size_t),
        ("trim_width", C.c_size_t), ("trim_height", C.c_size_t),
        ("dl", C.c_double), ("dm", C.c_double),
        ("kernel_size", C.c_size_t), ("oversampling", C.c_size_t),
        ("grid_mode", C.c_int), ("nw_layers", C.c_size_t),
        ("nw_factor", C.c_double), ("l_shift", C.c_double),
        ("m_shift", C.c_double), ("is_complex", C.c_int),
        ("ncpus", C.c_size_t), ("memory", C.c_double),
        ("w_limit", C.c_double), ("psf_mode", C.c_int),
        ("small_inversion", C.c_int),
    ]


class InvertStats(C.Structure):
    _fields_ = [
        ("total_weight", C.c_double), ("visibility_weight_sum", C.c_double),
        ("max_gridded_weight", C.c_double), ("gridded_count", C.c_size_t),
        ("min_w", C.c_double), ("max_w", C.c_double),
        ("nw", C.c_size_t), ("n_passes", C.c_size_t),
    ]


def _setup(L):
    sz, d, i, vp = C.c_size_t, C.c_double, C.c_int, C.c_void_p
    pd, pf = C.POINTER(C.c_double), C.POINTER(C.c_float)
    sig = {
        "orc_make_kernel": (i, [i, sz, sz, pd]),
        "orc_bessel0": (d, [d, d]),
        "orc_wsg_create": (vp, [i, sz, sz, d, d, sz, sz, sz]),
        "orc_wsg_destroy": (None, [vp]),
        "orc_wsg_set_grid_mode": (i, [vp, i]),
        "orc_wsg_set_is_complex": (None, [vp, i]),
        "orc_wsg_set_denormal_phase_centre": (None, [vp, d, d]),
        "orc_wsg_prepare_w_layers": (None, [vp, sz, d, d, d]),
        "orc_wsg_n_passes": (sz, [vp]),
        "orc_wsg_n_fft_threads": (sz, [vp]),
        "orc_wsg_w_to_layer": (sz, [vp, d]),
        "orc_wsg_layer_to_w": (d, [vp, sz]),
        "orc_wsg_is_in_layer_range": (i, [vp, d]),
        "orc_wsg_is_in_layer_range2": (i, [vp, d, d]),
        "orc_wsg_start_inversion_pass": (None, [vp, sz]),
        "orc_wsg_add_data_sample": (None, [vp, C.c_float, C.c_float, d, d, d]),
        "orc_wsg_add_samples": (None, [vp, sz, pd, pf]),
        "orc_wsg_finish_inversion_pass": (None, [vp]),
        "orc_wsg_finalize_image": (None, [vp, d]),
        "orc_wsg_get_image": (None, [vp, pd, pd]),
        "orc_wsg_get_layer": (None, [vp, sz, pd]),
        "orc_wsg_initialize_prediction": (None, [vp, pf, pf]),
        "orc_wsg_start_prediction_pass": (None, [vp, sz]),
        "orc_wsg_sample_data_sample": (None, [vp, d, d, d, pf]),
        "orc_wsg_sample_samples": (None, [vp, sz, pd, pf]),
        "orc_wsg_kernel_correction": (None, [vp, sz, pd]),
        "orc_wsg_sqrt_lm_table": (None, [vp, pd]),
        "orc_iw_create": (vp, [i, d, sz, sz, d, d, i, d]),
        "orc_iw_destroy": (None, [vp]),
        "orc_iw_grid": (None, [vp, sz, sz, sz, pd, pd, pf]),
        "orc_iw_finish": (None, [vp]),
        "orc_iw_get_weight": (d, [vp, d, d]),
        "orc_iw_lookup": (None, [vp, sz, sz, pd, pd, pf]),
        "orc_iw_width": (sz, [vp]),
        "orc_iw_height": (sz, [vp]),
        "orc_iw_grid_values": (None, [vp, pd]),
        "orc_ms_invert": (i, [i, C.POINTER(MsSettings), sz, sz, pd, pd, pf, pf,
                              vp, pf, pf, C.POINTER(InvertStats)]),
        "orc_ms_predict": (i, [i, C.POINTER(MsSettings), sz, sz, pd, pd, pf,
                               vp, pf, pf, pf, C.POINTER(InvertStats)]),
        "orc_suggested_w_grid_size": (sz, [C.POINTER(MsSettings), d, d]),
        "orc_calculate_fft_size": (None, [sz, d, d, C.POINTER(sz),
                                          C.POINTER(sz)]),
        "orc_dft_invert": (None, [sz, sz, d, d, d, d, sz, pd, pf, d, pd, sz]),
        "orc_dft_predict": (None, [sz, sz, d, d, d, d, sz, pd, pd, pd, sz]),
        "orc_fft2d": (None, [sz, sz, i, pd]),
        "orc_calculate_w_limits": (None, [C.POINTER(MsSettings), sz, sz, pd, pd, pf, vp, pd,
                                          pd, pd]),
        "orc_trim_image": (None, [pf, sz, sz, sz, sz, pf]),
        "orc_resample": (None, [pd, sz, sz, pd, sz, sz]),
        "orc_untrim_image": (None, [pf, sz, sz, sz, sz, pf]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    return L


def _c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


def make_kernel(mode="kb", oversampling=63, size=7):
    out = np.zeros(oversampling * size, np.float64)
    rc = lib().orc_make_kernel(KERNEL_MODES[mode], oversampling, size, _p(out, C.c_double))
    if rc != 0:
        raise RuntimeError("kernel mode not supported by the w-stacking gridder")
    return out


class RefGridder:
    """Restatement of WStackingGridder<float|double> (prec 'float'|'double')."""

    def __init__(self, width, height, dl, dm, nthreads=1, kernel_size=7,
                 oversampling=63, prec="float"):
        self.L = lib()
        self.width, self.height = width, height
        self.h = self.L.orc_wsg_create(0 if prec == "float" else 1, width, height,
                                       dl, dm, nthreads, kernel_size, oversampling)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_wsg_destroy(self.h)
            self.h = None

    def set_grid_mode(self, mode):
        if self.L.orc_wsg_set_grid_mode(self.h, KERNEL_MODES[mode]) != 0:
            raise RuntimeError("Blackman-Harris kernel not supported by W-Stacking gridder")

    def set_is_complex(self, c):
        self.L.orc_wsg_set_is_complex(self.h, int(bool(c)))

    def set_denormal_phase_centre(self, dl, dm):
        self.L.orc_wsg_set_denormal_phase_centre(self.h, dl, dm)

    def prepare_w_layers(self, nw, max_mem, min_w, max_w):
        self.L.orc_wsg_prepare_w_layers(self.h, nw, max_mem, min_w, max_w)

    @property
    def n_passes(self):
        return self.L.orc_wsg_n_passes(self.h)

    def w_to_layer(self, w):
        return self.L.orc_wsg_w_to_layer(self.h, w)

    def layer_to_w(self, layer):
        return self.L.orc_wsg_layer_to_w(self.h, layer)

    def is_in_layer_range(self, w, w2=None):
        if w2 is None:
            return bool(self.L.orc_wsg_is_in_layer_range(self.h, w))
        return bool(self.L.orc_wsg_is_in_layer_range2(self.h, w, w2))

    def start_inversion_pass(self, p):
        self.L.orc_wsg_start_inversion_pass(self.h, p)

    def add_samples(self, uvw, vis):
        uvw = _c(uvw, np.float64).reshape(-1, 3)
        v = np.ascontiguousarray(vis, np.complex64).view(np.float32)
        self.L.orc_wsg_add_samples(self.h, uvw.shape[0], _p(uvw, C.c_double), _p(v, C.c_float))

    def finish_inversion_pass(self):
        self.L.orc_wsg_finish_inversion_pass(self.h)

    def finalize_image(self, mult):
        self.L.orc_wsg_finalize_image(self.h, mult)

    def image(self, imag=False):
        re = np.zeros(self.width * self.height, np.float64)
        im = np.zeros_like(re) if imag else None
        self.L.orc_wsg_get_image(self.h, _p(re, C.c_double), _p(im, C.c_double))
        re = re.reshape(self.height, self.width)
        return (re, im.reshape(self.height, self.width)) if imag else re

    def layer(self, i):
        out = np.zeros(self.width * self.height * 2, np.float64)
        self.L.orc_wsg_get_layer(self.h, i, _p(out, C.c_double))
        return out.view(np.complex128).reshape(self.height, self.width)

    def initialize_prediction(self, real, imag=None):
        re = _c(real, np.float32)
        im = _c(imag, np.float32) if imag is not None else None
        self.L.orc_wsg_initialize_prediction(self.h, _p(re, C.c_float), _p(im, C.c_float))

    def start_prediction_pass(self, p):
        self.L.orc_wsg_start_prediction_pass(self.h, p)

    def sample(self, uvw):
        uvw = _c(uvw, np.float64).reshape(-1, 3)
        out = np.zeros(uvw.shape[0] * 2, np.float32)
        self.L.orc_wsg_sample_samples(self.h, uvw.shape[0], _p(uvw, C.c_double), _p(out, C.c_float))
        return out.view(np.complex64)

    def kernel_correction(self, axis):
        out = np.zeros(axis // 2 + 1, np.float64)
        self.L.orc_wsg_kernel_correction(self.h, axis, _p(out, C.c_double))
        return out

    def sqrt_lm_table(self):
        out = np.zeros(self.width * self.height, np.float64)
        self.L.orc_wsg_sqrt_lm_table(self.h, _p(out, C.c_double))
        return out.reshape(self.height, self.width)


class ImageWeights:
    """Restatement of ImageWeights (mode 'natural' | 'uniform' | 'briggs')."""

    MODES = {"natural": 0, "uniform": 1, "briggs": 2}

    def __init__(self, mode, width, height, dl, dm, robust=0.0,
                 weights_as_taper=False, super_weight=1.0):
        self.L = lib()
        self.h = self.L.orc_iw_create(self.MODES[mode], robust, width, height, dl, dm,
                                      int(weights_as_taper), super_weight)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_iw_destroy(self.h)
            self.h = None

    def grid(self, uvw_m, wavelengths, weights=None, npol=1):
        uvw = _c(uvw_m, np.float64).reshape(-1, 3)
        wl = _c(wavelengths, np.float64)
        w = _c(weights, np.float32) if weights is not None else None
        self.L.orc_iw_grid(self.h, uvw.shape[0], wl.shape[0], npol, _p(uvw, C.c_double),
                           _p(wl, C.c_double), _p(w, C.c_float))

    def finish(self):
        self.L.orc_iw_finish(self.h)

    def get_weight(self, u, v):
        return self.L.orc_iw_get_weight(self.h, u, v)

    def lookup(self, uvw_m, wavelengths):
        uvw = _c(uvw_m, np.float64).reshape(-1, 3)
        wl = _c(wavelengths, np.float64)
        out = np.zeros((uvw.shape[0], wl.shape[0]), np.float32)
        self.L.orc_iw_lookup(self.h, uvw.shape[0], wl.shape[0], _p(uvw, C.c_double),
                             _p(wl, C.c_double), _p(out, C.c_float))
        return out

    def grid_values(self):
        w, h = self.L.orc_iw_width(self.h), self.L.orc_iw_height(self.h)
        out = np.zeros(w * h // 2, np.float64)
        self.L.orc_iw_grid_values(self.h, _p(out, C.c_double))
        return out.reshape(h // 2, w)


def ms_settings(width, height, dl, dm, *, trim=None, kernel_size=7,
                oversampling=63, grid_mode="kb", nw=0, nw_factor=1.0,
                shift=(0.0, 0.0), is_complex=False, ncpus=1, memory=64e9,
                w_limit=0.0, psf=False, small_inversion=False):
    tw, th = trim if trim else (0, 0)
    return MsSettings(width, height, tw, th, dl, dm, kernel_size, oversampling,
                      KERNEL_MODES[grid_mode], nw, nw_factor, shift[0], shift[1],
                      int(is_complex), ncpus, memory, w_limit, int(psf),
                      int(small_inversion))


def ms_invert(settings, uvw_m, wavelengths, data, weights=None, image_weights=None,
              prec="float"):
    """WSMSGridder::Invert restatement. Returns (real, imag|None, stats)."""
    uvw = _c(uvw_m, np.float64).reshape(-1, 3)
    wl = _c(wavelengths, np.float64)
    d = np.ascontiguousarray(data, np.complex64).view(np.float32)
    w = _c(weights, np.float32) if weights is not None else None
    tw = settings.trim_width or settings.width
    th = settings.trim_height or settings.height
    re = np.zeros(tw * th, np.float32)
    im = np.zeros(tw * th, np.float32) if settings.is_complex else None
    st = InvertStats()
    rc = lib().orc_ms_invert(0 if prec == "float" else 1, C.byref(settings), uvw.shape[0],
                             wl.shape[0], _p(uvw, C.c_double), _p(wl, C.c_double),
                             _p(d, C.c_float), _p(w, C.c_float),
                             image_weights.h if image_weights else None,
                             _p(re, C.c_float), _p(im, C.c_float), C.byref(st))
    if rc != 0:
        raise RuntimeError("oracle invert failed")
    return re.reshape(th, tw), (im.reshape(th, tw) if im is not None else None), st


def ms_predict(settings, uvw_m, wavelengths, real, imag=None, weights=None,
               image_weights=None, prec="float", initial=np.nan):
    uvw = _c(uvw_m, np.float64).reshape(-1, 3)
    wl = _c(wavelengths, np.float64)
    w = _c(weights, np.float32) if weights is not None else None
    re = _c(real, np.float32)
    im = _c(imag, np.float32) if imag is not None else None
    out = np.full(uvw.shape[0] * wl.shape[0] * 2, initial, np.float32)
    st = InvertStats()
    rc = lib().orc_ms_predict(0 if prec == "float" else 1, C.byref(settings), uvw.shape[0],
                              wl.shape[0], _p(uvw, C.c_double), _p(wl, C.c_double),
                              _p(w, C.c_float), image_weights.h if image_weights else None,
                              _p(re, C.c_float), _p(im, C.c_float), _p(out, C.c_float),
                              C.byref(st))
    if rc != 0:
        raise RuntimeError("oracle predict failed")
    return out.view(np.complex64).reshape(uvw.shape[0], wl.shape[0]), st


def suggested_w_grid_size(settings, min_w, max_w):
    return lib().orc_suggested_w_grid_size(C.byref(settings), min_w, max_w)


def calculate_fft_size(size, pixel_scale, beam_size):
    a, b = C.c_size_t(), C.c_size_t()
    lib().orc_calculate_fft_size(size, pixel_scale, beam_size, C.byref(a), C.byref(b))
    return a.value, b.value


def dft_invert(width, height, dl, dm, uvw_lambda, vis, factor=1.0, shift=(0.0, 0.0),
               nthreads=None):
    uvw = _c(uvw_lambda, np.float64).reshape(-1, 3)
    v = np.ascontiguousarray(vis, np.complex64).view(np.float32)
    out = np.zeros(width * height, np.float64)
    lib().orc_dft_invert(width, height, dl, dm, shift[0], shift[1], uvw.shape[0],
                         _p(uvw, C.c_double), _p(v, C.c_float), factor,
                         _p(out, C.c_double), nthreads or os.cpu_count() or 1)
    return out.reshape(height, width)


def dft_predict(width, height, dl, dm, uvw_lambda, image, shift=(0.0, 0.0), nthreads=None):
    uvw = _c(uvw_lambda, np.float64).reshape(-1, 3)
    img = _c(image, np.float64)
    out = np.zeros(uvw.shape[0] * 2, np.float64)
    lib().orc_dft_predict(width, height, dl, dm, shift[0], shift[1], uvw.shape[0],
                          _p(uvw, C.c_double), _p(img, C.c_double), _p(out, C.c_double),
                          nthreads or os.cpu_count() or 1)
    return out.view(np.complex128)


def fft2d(data, sign):
    a = np.ascontiguousarray(data, np.complex128).copy()
    h, w = a.shape
    lib().orc_fft2d(w, h, sign, a.view(np.float64).ctypes.data_as(C.POINTER(C.c_double)))
    return a


def calculate_w_limits(settings, uvw_m, wavelengths, image_weights, weights=None):
    """MSGridderBase::calculateWLimits restatement: (min_w, max_w,
    max_baseline_uvw) in wavelengths."""
    uvw = _c(uvw_m, np.float64).reshape(-1, 3)
    wl = _c(wavelengths, np.float64)
    w = _c(weights, np.float32) if weights is not None else None
    lo, hi, bl = C.c_double(), C.c_double(), C.c_double()
    lib().orc_calculate_w_limits(C.byref(settings), uvw.shape[0], wl.shape[0],
                                 _p(uvw, C.c_double), _p(wl, C.c_double), _p(w, C.c_float),
                                 image_weights.h, C.byref(lo), C.byref(hi), C.byref(bl))
    return lo.value, hi.value, bl.value


def trim_image(image, trim_width, trim_height):
    img = _c(image, np.float32)
    h, w = img.shape
    out = np.zeros((trim_height, trim_width), np.float32)
    lib().orc_trim_image(_p(img, C.c_float), w, h, trim_width, trim_height, _p(out, C.c_float))
    return out


def untrim_image(image, width, height):
    img = _c(image, np.float32)
    th, tw = img.shape
    out = np.zeros((height, width), np.float32)
    lib().orc_untrim_image(_p(img, C.c_float), tw, th, width, height, _p(out, C.c_float))
    return out


def resample(image, out_width, out_height):
    """FFT resampling of a real [H, W] image to [out_height, out_width] (the
    small-inversion resampler), FP64."""
    img = _c(image, np.float64)
    h, w = img.shape
    out = np.zeros((out_height, out_width), np.float64)
    lib().orc_resample(_p(img, C.c_double), w, h, _p(out, C.c_double), out_width, out_height)
    return out
