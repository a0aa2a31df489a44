"""meep_amd — MI355X-native FDTD time-stepping engine with Meep's fields/structure API.

The compute path is the HIP library meep_amd/_lib/libmeep_amd.so (kernels for gfx950).
There is no CPU fallback: :func:`load_library` fails when the library is missing.
"""

import ctypes
import os

from . import capi
from .capi import *  # noqa: F401,F403  (reference enums and API classes)

_LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib", "libmeep_amd.so")
_lib = None


def library_path():
    return _LIB_PATH


def load_library():
    """Load and bind the HIP engine library (built by meep_amd/build.py)."""
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(
                f"meep_amd: HIP engine library not built ({_LIB_PATH}); run "
                "python -m meep_amd.build")
        _lib = capi.bind(ctypes.CDLL(_LIB_PATH))
    return _lib
