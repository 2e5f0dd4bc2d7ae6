"""sphcode_amd -- MI355X-native (gfx950) per-step SPH particle pipeline.

The compute path is the HIP library ``sphcode_amd/lib/libsphgpu.so`` behind
the C ABI of ``include/sphgpu.h``; this package only loads it.  There is no
CPU fallback: ``library()`` raises if the library is missing or no HIP
device is present.
"""
from __future__ import annotations

import ctypes as C
import pathlib

import numpy as np

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
    """One sphgpu context on the current HIP device.  ``stats`` turns on the
    diagnostic counters that :meth:`stats` reports (interactions, Newton
    iterations, node visits); the library's default is off."""

    def __init__(self, params: Params, stats: bool = True):
        _, fns = library()
        super().__init__(fns, params, name="sphgpu")
        self.enable_stats(stats)

    def enable_stats(self, on: bool = True) -> None:
        self._call("enable_stats", C.c_int(1 if on else 0))

    def timers(self) -> dict:
        t = _abi.Timers()
        self._call("get_timers", C.byref(t))
        return t.as_dict()

    def enable_timers(self, on: bool = True) -> None:
        self._call("enable_timers", C.c_int(1 if on else 0))

    def synchronize(self) -> None:
        self._call("synchronize")

    def restart(self, time: float, dt: float, h_per_v_sig: float) -> None:
        """Resume from an uploaded state saved after a full step."""
        self._call("restart", C.c_double(time), C.c_double(dt), C.c_double(h_per_v_sig))

    # ---------------------------------------------------- differentiable rollout
    def rollout(self, steps: int, dt: float) -> None:
        """Solver::initialize, then ``steps`` steps of fixed ``dt``, with
        per-step checkpoints for :meth:`rollout_adjoint`."""
        self._call("rollout", C.c_int(int(steps)), C.c_double(float(dt)))

    def rollout_adjoint(self, dl_dx=None, dl_dv=None, dl_du=None):
        """Gradient of a loss with respect to the initial (pos, vel, ene),
        given its gradient with respect to the final ones (caller order;
        None means zero)."""
        n, d = self.n, self.dim

        def arg(a, shape):
            if a is None:
                return None, None
            a = np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(shape))
            return a, a.ctypes.data_as(_abi._DP)

        sx, px = arg(dl_dx, (n, d))
        sv, pv = arg(dl_dv, (n, d))
        su, pu = arg(dl_du, (n,))
        gx, gv, gu = np.zeros((n, d)), np.zeros((n, d)), np.zeros(n)
        self._call("rollout_adjoint", px, pv, pu, gx.ctypes.data_as(_abi._DP), gv.ctypes.data_as(_abi._DP),
                   gu.ctypes.data_as(_abi._DP))
        return gx, gv, gu
