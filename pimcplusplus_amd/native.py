"""Loads the in-tree HIP library (pimcplusplus_amd/_lib/libpimc_gpu.so).

There is no CPU fallback: a missing or unloadable library raises.
"""

from __future__ import annotations

import ctypes
import os

from . import abi

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib", "libpimc_gpu.so")

_lib = None


def library() -> ctypes.CDLL:
    """The loaded HIP library (symbols bound with their C signatures)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"HIP action library not built: {LIB_PATH} is missing "
                               "(run python -m pimcplusplus_amd.build)")
        _lib = abi.bind(ctypes.CDLL(LIB_PATH))
    return _lib


def open_context(device: int = 0) -> abi.PimcLib:
    """A PimcLib on the HIP library, bound to GPU `device`."""
    return abi.PimcLib(library(), device)
