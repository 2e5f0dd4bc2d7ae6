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

_EXTRA = {
    "set_threads": (C.c_int, [C.c_void_p, C.c_int]),
    "set_exhaustive": (C.c_int, [C.c_void_p, C.c_int]),
    "set_tree_pool": (C.c_int, [C.c_void_p, C.c_int64]),
    "stage_seconds": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)]),
}
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
