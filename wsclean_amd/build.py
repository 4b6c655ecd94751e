"""In-tree builds: the HIP library (libhipwstacking.so, gfx950) and the
host plugin layer over its C ABI (libwsclean_amd_host.so)."""
from __future__ import annotations

import os
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))


def _make(directory: str, force: bool, *extra: str) -> None:
    args = ["make", "-s", "-j8", "-C", directory, *extra]
    if force:
        args.insert(1, "-B")
    subprocess.run(args, check=True)


def build(arch: str = "gfx950", force: bool = True) -> str:
    """Compile wsclean_amd/csrc into wsclean_amd/libhipwstacking.so."""
    _make(os.path.join(_HERE, "csrc"), force, f"ARCH={arch}")
    return os.path.join(_HERE, "libhipwstacking.so")


def build_host(force: bool = True) -> str:
    """Compile wsclean_amd/host into wsclean_amd/libwsclean_amd_host.so
    (needs libhipwstacking.so)."""
    _make(os.path.join(_HERE, "host"), force)
    return os.path.join(_HERE, "libwsclean_amd_host.so")
