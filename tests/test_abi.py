"""The C-ABI library loads and exports every entry point of
include/hip_wstacking.h; host-only entry points behave (no GPU needed)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "hip_wstacking.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hws_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from wsclean_amd import _lib
    lib = _lib.load()
    names = header_functions()
    assert len(names) >= 27
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_lib.SIGNATURES), "binding and header disagree"


def test_get_kernel_matches_oracle(oracle):
    from wsclean_amd import get_kernel
    for mode in ("kb", "kb_nosinc", "rect", "gaussian", "gaussian_nosinc", "bn", "bn_nosinc",
                 "nn"):
        np.testing.assert_allclose(get_kernel(mode, 63, 7), oracle.make_kernel(mode, 63, 7),
                                   rtol=1e-14, atol=1e-15)
    np.testing.assert_allclose(get_kernel("kb", 1023, 7), oracle.make_kernel("kb", 1023, 7),
                               rtol=1e-14, atol=1e-15)


def test_blackman_harris_rejected():
    from wsclean_amd import get_kernel
    from wsclean_amd._lib import HwsError
    with pytest.raises(HwsError, match="Blackman-Harris"):
        get_kernel("bh", 63, 7)


def test_invalid_configs_rejected_before_device_use():
    from wsclean_amd import _lib
    L = _lib.load()
    base = dict(width=64, height=64, pixel_size_x=0.01, pixel_size_y=0.01, kernel_size=7,
                oversampling=63, kernel_mode=1, precision=1, device=0, n_images=1,
                complex_mask=0, l_shift=0.0, m_shift=0.0, rotate_visibilities=0, tile_size=0)
    bad = [dict(width=0), dict(width=66 * 11), dict(height=63), dict(kernel_size=8),
           dict(oversampling=0), dict(kernel_mode=8), dict(precision=7), dict(n_images=5),
           dict(tile_size=100), dict(pixel_size_x=-1.0)]
    for b in bad:
        cfg = _lib.HwsConfig(**{**base, **b})
        h = C.c_void_p()
        st = L.hws_create(C.byref(cfg), C.byref(h))
        assert st in (1, 4), b
        assert h.value is None
        assert L.hws_last_error(None)
    assert L.hws_status_string(0) == b"ok"
    assert L.hws_status_string(5) == b"call sequence violated"


def test_null_context_calls_fail_cleanly():
    from wsclean_amd import _lib
    L = _lib.load()
    assert L.hws_start_inversion_pass(None, 0) == 1
    assert L.hws_finish_inversion_pass(None) == 1
    assert L.hws_synchronize(None) == 1


def test_direct_transform_rejects_bad_arguments_before_device_use():
    from wsclean_amd import _lib
    L = _lib.load()
    st = L.hws_direct_invert_device(0, None, 0, 16, 0.01, 0.01, 0.0, 0.0, 0, 1, None, None,
                                    None, 1, 1.0, None)
    assert st == 1 and b"image size" in L.hws_last_error(None)
    st = L.hws_direct_predict_device(0, None, 16, 16, -0.01, 0.01, 0.0, 0.0, None, 0, 1, None,
                                     None, None, 1)
    assert st == 1 and b"pixel sizes" in L.hws_last_error(None)
    st = L.hws_direct_invert_device(0, None, 16, 16, 0.01, 0.01, 0.0, 0.0, 0, 1, None, None,
                                    None, 1, 1.0, None)
    assert st == 1 and b"null" in L.hws_last_error(None)
    st = L.hws_direct_invert_device(0, None, 33, 16, 0.01, 0.01, 0.0, 0.0, 0, 1, None, None,
                                    None, 1, 1.0, None)
    assert st == 1 and b"even" in L.hws_last_error(None)
    fake = C.c_void_p(1 << 20)  # never dereferenced: rejected by the size check
    st = L.hws_direct_predict_device(0, None, 16, 16, 0.01, 0.01, 0.0, 0.0, fake, 1 << 28, 16,
                                     fake, fake, fake, 1)
    assert st == 1 and b"2^31" in L.hws_last_error(None)
