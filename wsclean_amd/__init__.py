"""MI355X-native w-stacking gridder/degridder (a drop-in for the w-stacking
path of WSClean: WStackingGridder behind WSMSGridder / MSGridderBase).

The compute path is libhipwstacking.so (hand-written HIP for gfx950) behind
the C ABI in include/hip_wstacking.h.
"""
from .gridder import DeviceGridder, WStackingGridder, get_kernel  # noqa: F401

__all__ = ["WStackingGridder", "DeviceGridder", "get_kernel"]
