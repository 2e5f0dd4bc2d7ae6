"""sphcode_amd -- MI355X-native (gfx950) per-step SPH particle pipeline.

The compute path is the HIP library ``sphcode_amd/lib/libsphgpu.so`` behind
the C ABI of ``include/sphgpu.h``; this package only loads it.  There is no
CPU fallback: ``library()`` raises if the library is missing or no HIP
device is present.
"""
from __future__ import annotations

import ctypes as C
import pathlib

from . import _abi
from ._abi import (CUBIC_SPLINE, DISPH, FP64, GSPH, MIXED, SSPH, WENDLAND,  # noqa: F401
                   Params, default_params)
from .engine import Engine, SPHError  # noqa: F401

LIB_PATH = pathlib.Path(__file__).resolve().parent / "lib" / "libsphgpu.so"
_lib = None
_fns = None


def library():
    """Load libsphgpu.so once and return (CDLL, bound functions)."""
    global _lib, _fns
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"HIP library not built: {LIB_PATH} (run python -m sphcode_amd.build)")
        lib = C.CDLL(str(LIB_PATH))
        _fns = _abi.bind(lib, "sphgpu_", _abi.GPU_ONLY)
        _lib = lib
    return _lib, _fns


class GPUEngine(Engine):
    """One sphgpu context on the current HIP device."""

    def __init__(self, params: Params):
        _, fns = library()
        super().__init__(fns, params, name="sphgpu")

    def timers(self) -> dict:
        t = _abi.Timers()
        self._call("get_timers", C.byref(t))
        return t.as_dict()

    def enable_timers(self, on: bool = True) -> None:
        self._call("enable_timers", C.c_int(1 if on else 0))

    def synchronize(self) -> None:
        self._call("synchronize")
