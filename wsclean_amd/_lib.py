"""ctypes binding of libhipwstacking.so (C ABI: include/hip_wstacking.h).

The library is built in-tree for gfx950 by ``wsclean_amd.build.build()``.
Loading fails loudly when it is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhipwstacking.so")

HWS_OK = 0
STATUS_NAMES = {
    0: "HWS_OK", 1: "HWS_ERR_INVALID_ARGUMENT", 2: "HWS_ERR_HIP",
    3: "HWS_ERR_OUT_OF_MEMORY", 4: "HWS_ERR_UNSUPPORTED", 5: "HWS_ERR_BAD_STATE",
}

KERNEL_MODES = {
    "nn": 0, "kb": 1, "kb_nosinc": 2, "rect": 3, "gaussian": 4,
    "gaussian_nosinc": 5, "bn": 6, "bn_nosinc": 7, "bh": 8,
}
PRECISIONS = {"plain": 0, "reference": 1, "mixed": 2}


class HwsConfig(C.Structure):
    _fields_ = [
        ("width", C.c_size_t), ("height", C.c_size_t),
        ("pixel_size_x", C.c_double), ("pixel_size_y", C.c_double),
        ("kernel_size", C.c_size_t), ("oversampling", C.c_size_t),
        ("kernel_mode", C.c_int), ("precision", C.c_int), ("device", C.c_int),
        ("n_images", C.c_size_t), ("complex_mask", C.c_uint32),
        ("l_shift", C.c_double), ("m_shift", C.c_double),
        ("rotate_visibilities", C.c_int), ("tile_size", C.c_size_t),
        ("work_chunk", C.c_size_t),
    ]


class HwsStats(C.Structure):
    _fields_ = [
        ("samples_in", C.c_uint64), ("samples_gridded", C.c_uint64),
        ("samples_off_grid", C.c_uint64), ("samples_off_layer", C.c_uint64),
        ("bins_used", C.c_uint64), ("chunks", C.c_uint64),
        ("ms_bin", C.c_double), ("ms_grid", C.c_double), ("ms_fft", C.c_double),
        ("ms_project", C.c_double), ("ms_total", C.c_double),
        ("tiles_total", C.c_uint64), ("tiles_occupied", C.c_uint64 * 2),
        ("tiles_needed", C.c_uint64 * 2), ("tile_rows_total", C.c_uint64),
        ("tile_rows_live", C.c_uint64 * 2), ("tile_cols_total", C.c_uint64),
        ("tile_cols_needed", C.c_uint64 * 2),
        ("total_weight", C.c_double * 4), ("visibility_weight_sum", C.c_double * 4),
        ("max_gridded_weight", C.c_double * 4), ("gridded_visibility_count", C.c_uint64 * 4),
    ]


class HwsWeightsConfig(C.Structure):
    _fields_ = [
        ("mode", C.c_int), ("briggs_robustness", C.c_double),
        ("width", C.c_size_t), ("height", C.c_size_t),
        ("pixel_scale_x", C.c_double), ("pixel_scale_y", C.c_double),
        ("weights_as_taper", C.c_int), ("super_weight", C.c_double), ("device", C.c_int),
    ]


class HwsError(RuntimeError):
    def __init__(self, status, message):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


_lib = None

# every entry point declared in include/hip_wstacking.h: (restype, argtypes)
_vp, _sz, _d, _i = C.c_void_p, C.c_size_t, C.c_double, C.c_int
_pd, _pf = C.POINTER(C.c_double), C.POINTER(C.c_float)
SIGNATURES = {
    "hws_create": (_i, [C.POINTER(HwsConfig), C.POINTER(_vp)]),
    "hws_destroy": (None, [_vp]),
    "hws_last_error": (C.c_char_p, [_vp]),
    "hws_status_string": (C.c_char_p, [_i]),
    "hws_prepare_w_layers": (_i, [_vp, _sz, _d, _d, _d, C.POINTER(_sz)]),
    "hws_w_to_layer": (_sz, [_vp, _d, _i]),
    "hws_layer_to_w": (_d, [_vp, _sz, _i]),
    "hws_pass_layer_range": (_i, [_vp, _sz, C.POINTER(_sz), C.POINTER(_sz)]),
    "hws_start_inversion_pass": (_i, [_vp, _sz]),
    "hws_add_data": (_i, [_vp, _sz, _sz, _pd, _pd, _pf]),
    "hws_add_samples": (_i, [_vp, _sz, _pd, _pf]),
    "hws_finish_inversion_pass": (_i, [_vp]),
    "hws_finalize_image": (_i, [_vp, _d, _pf, _pf]),
    "hws_initialize_prediction": (_i, [_vp, _pf, _pf]),
    "hws_start_prediction_pass": (_i, [_vp, _sz]),
    "hws_sample_data": (_i, [_vp, _sz, _sz, _pd, _pd, _pf]),
    "hws_sample_samples": (_i, [_vp, _sz, _pd, _pf]),
    "hws_get_gridded_uv_layer": (_i, [_vp, _sz, _sz, _pf]),
    "hws_get_kernel": (_i, [_i, _sz, _sz, _pd]),
    "hws_set_stream": (_i, [_vp, _vp]),
    "hws_plan_device": (_i, [_vp, _sz, _sz, _vp, _vp]),
    "hws_invert_planned_device": (_i, [_vp, _vp, _d, _vp, _vp]),
    "hws_predict_planned_device": (_i, [_vp, _vp, _vp, _vp]),
    "hws_invert_device": (_i, [_vp, _sz, _sz, _vp, _vp, _vp, _d, _vp, _vp]),
    "hws_predict_device": (_i, [_vp, _sz, _sz, _vp, _vp, _vp, _vp, _vp]),
    "hws_get_stats": (_i, [_vp, C.POINTER(HwsStats)]),
    "hws_synchronize": (_i, [_vp]),
    "hws_graph_capture_invert": (_i, [_vp, _vp, _d, _vp, _vp, C.POINTER(_vp)]),
    "hws_graph_capture_predict": (_i, [_vp, _vp, _vp, _vp, C.POINTER(_vp)]),
    "hws_graph_capture_invert_batch": (_i, [C.POINTER(_vp), _sz, C.POINTER(_vp), _pd,
                                            C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp)]),
    "hws_graph_capture_predict_batch": (_i, [C.POINTER(_vp), _sz, C.POINTER(_vp),
                                             C.POINTER(_vp), C.POINTER(_vp), C.POINTER(_vp)]),
    "hws_weights_create": (_i, [C.POINTER(HwsWeightsConfig), C.POINTER(_vp)]),
    "hws_weights_destroy": (None, [_vp]),
    "hws_weights_grid_device": (_i, [_vp, _vp, _sz, _sz, _sz, _vp, _vp, _vp]),
    "hws_weights_finish": (_i, [_vp, _vp]),
    "hws_weights_shape": (_i, [_vp, C.POINTER(_sz), C.POINTER(_sz)]),
    "hws_weights_get_grid": (_i, [_vp, _pd, _pd]),
    "hws_set_weighting": (_i, [_vp, _vp, _vp, _i]),
    "hws_stage_begin": (_i, [_vp, _sz, _sz, _pd, _i]),
    "hws_stage_rows": (_i, [_vp, _sz, _pd, _pf]),
    "hws_stage_end": (_i, [_vp]),
    "hws_invert_staged": (_i, [_vp, _d, _sz, _sz, _sz, _sz, _pf, _pf]),
    "hws_predict_staged": (_i, [_vp, _pf, _pf, _sz, _sz, _sz, _sz, _pf]),
    "hws_finalize_image_resampled": (_i, [_vp, _d, _sz, _sz, _sz, _sz, _pf, _pf]),
    "hws_initialize_prediction_resampled": (_i, [_vp, _pf, _pf, _sz, _sz, _sz, _sz]),
    "hws_graph_launch": (_i, [_vp, _vp]),
    "hws_graph_destroy": (None, [_vp]),
    "hws_direct_invert_device": (_i, [_i, _vp, _sz, _sz, _d, _d, _d, _d, _sz, _sz, _vp, _vp,
                                      _vp, _sz, _d, _vp]),
    "hws_direct_predict_device": (_i, [_i, _vp, _sz, _sz, _d, _d, _d, _d, _vp, _sz, _sz, _vp,
                                       _vp, _vp, _sz]),
    "hws_direct_invert": (_i, [_i, _sz, _sz, _d, _d, _d, _d, _sz, _sz, _pd, _pd, _pf, _d, _pd]),
    "hws_direct_predict": (_i, [_i, _sz, _sz, _d, _d, _d, _d, _pf, _sz, _sz, _pd, _pd, _pf]),
    "hws_resample_device": (_i, [_i, _vp, _vp, _sz, _sz, _vp, _sz, _sz, _sz]),
    "hws_trim_device": (_i, [_i, _vp, _vp, _sz, _sz, _vp, _sz, _sz, _sz]),
    "hws_untrim_device": (_i, [_i, _vp, _vp, _sz, _sz, _vp, _sz, _sz, _sz]),
    "hws_direct_predict_device_noalloc": (_i, [_i, _vp, _sz, _sz, _d, _d, _d, _d, _vp, _sz, _sz,
                                               _vp, _vp, _vp, _sz]),
}


def load():
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with "
                "`python -c 'import __graft_entry__ as e; e.build()'` "
                "(there is no CPU fallback)")
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def check(status, handle=None):
    if status != HWS_OK:
        msg = load().hws_last_error(handle)
        raise HwsError(status, msg.decode() if msg else "")
