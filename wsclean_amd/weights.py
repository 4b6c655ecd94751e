"""Imaging weights on the GPU: ImageWeights (structures/imageweights.h:19-129,
structures/imageweights.cpp:33-49, 75-144) gridded from device-resident rows,
and the weighting a gridder context applies when it grids
(MSGridderBase::readAndWeightVisibilities, gridding/msgridderbase.cpp:563-589,
640-659).

``DeviceImageWeights`` holds the weight grid (natural, uniform or Briggs) on
one device: ``grid()`` adds the weights of a set of rows (one call per
MeasurementSet, as ImageWeights::Grid is called), ``finish()`` turns the
summed grid into weights (FinishGridding), and ``weight_grid()`` copies it to
host memory. ``DeviceGridder.set_weighting`` attaches it to a gridder: the
plan then looks up every sample's imaging weight, the grid kernels multiply
each visibility by visibility weight x imaging weight, and the invert
normalises each image by its own weight sum, as WSMSGridder does
(gridding/wsmsgridder.cpp:436).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _checks, _lib
from ._lib import HwsWeightsConfig, check

WEIGHT_MODES = {"natural": 0, "uniform": 1, "briggs": 2}
VISIBILITY_WEIGHTING = {"normal": 0, "squared": 1, "unit": 2}


class DeviceImageWeights:
    """ImageWeights(mode, width, height, pixel_scale_x, pixel_scale_y,
    weights_as_taper, super_weight) with Briggs robustness `robust`, on
    `device`. Work is ordered on torch's current stream of that device."""

    def __init__(self, mode, width, height, pixel_scale_x, pixel_scale_y, *, robust=0.0,
                 weights_as_taper=False, super_weight=1.0, device=0):
        self.L = _lib.load()
        self.mode = mode
        self.device = device
        cfg = HwsWeightsConfig(WEIGHT_MODES[mode], float(robust), width, height,
                               float(pixel_scale_x), float(pixel_scale_y),
                               int(bool(weights_as_taper)), float(super_weight), device)
        h = C.c_void_p()
        check(self.L.hws_weights_create(C.byref(cfg), C.byref(h)), None)
        self.h = h
        w, hh = C.c_size_t(), C.c_size_t()
        check(self.L.hws_weights_shape(self.h, C.byref(w), C.byref(hh)), None)
        self.width, self.height = w.value, hh.value

    def _stream(self):
        import torch
        return C.c_void_p(torch.cuda.current_stream(torch.device("cuda", self.device)).cuda_stream)

    def grid(self, uvw_m, wavelengths, weights=None):
        """Adds the rows' weights: uvw_m float64 [nrow, 3], wavelengths float64
        [nchan], weights float32 [nrow, nchan, npol] (None: all 1)."""
        import torch
        nrow, nchan = _checks.plan_args(uvw_m, wavelengths, self.device)
        npol = 1
        if weights is not None:
            if weights.dim() != 3 or tuple(weights.shape[:2]) != (nrow, nchan):
                raise ValueError(f"weights must be [nrow, nchan, npol] with nrow={nrow}, "
                                 f"nchan={nchan}; got {tuple(weights.shape)}")
            npol = weights.shape[2]
            _checks.tensor(weights, "weights", torch.float32, device=self.device)
        self._keep = (uvw_m, wavelengths, weights)
        check(self.L.hws_weights_grid_device(
            self.h, self._stream(), nrow, nchan, npol, C.c_void_p(uvw_m.data_ptr()),
            C.c_void_p(wavelengths.data_ptr()),
            C.c_void_p(weights.data_ptr()) if weights is not None else None), None)

    def finish(self):
        """FinishGridding: summed weights -> imaging weights (synchronises)."""
        check(self.L.hws_weights_finish(self.h, self._stream()), None)

    def weight_grid(self):
        """(grid [height/2, width] float64, total of the gridded weights)."""
        out = np.zeros((self.height // 2, self.width), np.float64)
        total = C.c_double()
        check(self.L.hws_weights_get_grid(self.h, out.ctypes.data_as(C.POINTER(C.c_double)),
                                          C.byref(total)), None)
        return out, total.value

    def close(self):
        if getattr(self, "h", None):
            self.L.hws_weights_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


__all__ = ["DeviceImageWeights", "VISIBILITY_WEIGHTING", "WEIGHT_MODES"]
