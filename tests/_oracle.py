"""Loader for the CPU oracle (oracle/lib/libsphref<D>.so).

Test infrastructure: only tests/, the smoke check and bench.py's CPU
baseline use it.
"""
from __future__ import annotations

import ctypes as C
import pathlib
import subprocess

from sphcode_amd import _abi
from sphcode_amd.engine import Engine

ROOT = pathlib.Path(__file__).resolve().parents[1]
LIBDIR = ROOT / "oracle" / "lib"

_dp = C.POINTER(C.c_double)
_EXTRA = {
    "set_threads": (C.c_int, [C.c_void_p, C.c_int]),
    "set_exhaustive": (C.c_int, [C.c_void_p, C.c_int]),
    "set_tree_pool": (C.c_int, [C.c_void_p, C.c_int64]),
    "stage_seconds": (C.c_int, [C.c_void_p, _dp]),
    "kernel_eval": (C.c_int, [C.c_int, C.c_double, C.c_double, _dp, _dp, _dp]),
    "softening": (C.c_int, [C.c_double, C.c_double, _dp, _dp]),
    "hll": (C.c_int, [_dp, _dp, _dp, _dp]),
    "van_leer": (C.c_double, [C.c_double, C.c_double]),
    "min_image": (C.c_int, [C.c_void_p, _dp, _dp, _dp]),
}


def kernel_eval(dim: int, kernel: int, r: float, h: float):
    """(W, c, dW/dh) of the oracle's kernel, with grad W = r_ij * c."""
    _, fns = library(dim)
    w, c, dh = C.c_double(), C.c_double(), C.c_double()
    fns["kernel_eval"](kernel, r, h, C.byref(w), C.byref(c), C.byref(dh))
    return w.value, c.value, dh.value


def softening(r: float, h: float):
    _, fns = library(3)
    f, g = C.c_double(), C.c_double()
    fns["softening"](r, h, C.byref(f), C.byref(g))
    return f.value, g.value


def hll(left, right):
    _, fns = library(1)
    L = (C.c_double * 4)(*left)
    R = (C.c_double * 4)(*right)
    p, v = C.c_double(), C.c_double()
    fns["hll"](L, R, C.byref(p), C.byref(v))
    return p.value, v.value


def van_leer(a: float, b: float) -> float:
    _, fns = library(1)
    return fns["van_leer"](a, b)
_cache: dict = {}


def build() -> None:
    subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "-j4"], check=True)


def library(dim: int, fast: bool = False):
    key = (dim, fast)
    if key not in _cache:
        path = LIBDIR / f"libsphref{dim}{'_fast' if fast else ''}.so"
        if not path.exists():
            build()
        lib = C.CDLL(str(path))
        _cache[key] = (lib, _abi.bind(lib, "sphref_", _EXTRA))
    return _cache[key]


class OracleEngine(Engine):
    def __init__(self, params: _abi.Params, exhaustive: bool = False, fast: bool = False, threads: int = 0):
        lib, fns = library(int(params.dim), fast)
        super().__init__(fns, params, name="sphref")
        if exhaustive:
            self._call("set_exhaustive", C.c_int(1))
        if threads:
            fns["set_threads"](self._ctx, threads)

    def set_tree_pool(self, nodes: int) -> None:
        self._call("set_tree_pool", C.c_int64(nodes))

    def stage_seconds(self):
        import numpy as np
        out = np.zeros(7)
        self._call("stage_seconds", out.ctypes.data_as(C.POINTER(C.c_double)))
        return out
