"""In-tree build of the HIP library (libhipwstacking.so, gfx950)."""
from __future__ import annotations

import os
import subprocess

_HERE = os.path.dirname(os.path.abspath(__file__))


def build(arch: str = "gfx950", force: bool = True) -> str:
    """Compile wsclean_amd/csrc into wsclean_amd/libhipwstacking.so."""
    csrc = os.path.join(_HERE, "csrc")
    args = ["make", "-s", "-C", csrc, f"ARCH={arch}"]
    if force:
        args.insert(1, "-B")
    subprocess.run(args, check=True)
    return os.path.join(_HERE, "libhipwstacking.so")
