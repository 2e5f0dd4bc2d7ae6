"""Python handle over a library that exports the sphgpu C ABI.

``Engine`` owns one context created through ``<prefix>create`` and exposes
the stage calls of one step in the order of Solver::integrate
(src/solver.cpp:417-429).  A non-zero status raises ``SPHError``, the
counterpart of the reference's SPHException (include/exception.hpp:28-56).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi


class SPHError(RuntimeError):
    """A stage returned a non-zero status."""

    def __init__(self, code: int, message: str, call: str):
        self.code = code
        self.call = call
        name = _abi.ERROR_NAMES.get(code, str(code))
        super().__init__(f"{call}: {name}: {message}")


class Engine:
    def __init__(self, fns: dict, params: _abi.Params, name: str = "sphgpu"):
        self._fns = fns
        self.name = name
        self.params = params.copy()
        self.dim = int(params.dim)
        self.n = 0
        self._ctx = C.c_void_p()
        self._check(fns["create"](C.byref(self.params), C.byref(self._ctx)), "create")

    # ------------------------------------------------------------ helpers
    def _check(self, rc: int, call: str) -> None:
        if rc != 0:
            msg = self._fns["last_error"]()
            raise SPHError(rc, msg.decode() if msg else "", f"{self.name}_{call}")

    def _call(self, name: str, *args) -> None:
        self._check(self._fns[name](self._ctx, *args), name)

    def close(self) -> None:
        if self._ctx:
            self._fns["destroy"](self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):  # pragma: no cover - interpreter shutdown order
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # ---------------------------------------------------------- particles
    def upload(self, fields: dict) -> None:
        n = int(np.asarray(fields["pos"]).shape[0])
        present = tuple(k for k in _abi.ALL_FIELDS if k in fields and fields[k] is not None)
        arrays = _abi.ParticleArrays(n, self.dim, present, fields)
        s = arrays.struct
        for name in _abi.ALL_FIELDS:
            if name not in present:
                setattr(s, name, None)
        self._call("upload", C.byref(s))
        self.n = n

    def download(self, fields=_abi.ALL_FIELDS) -> dict:
        arrays = _abi.ParticleArrays(self.n, self.dim, tuple(fields))
        s = arrays.struct
        for name in _abi.ALL_FIELDS:
            if name not in fields:
                setattr(s, name, None)
        self._call("download", C.byref(s))
        return arrays.arrays

    # -------------------------------------------------------------- stages
    def initialize(self) -> None:
        self._call("initialize")

    def timestep(self) -> float:
        dt = C.c_double()
        self._call("timestep", C.byref(dt))
        return dt.value

    def predict(self) -> None:
        self._call("predict")

    def make_tree(self) -> None:
        self._call("make_tree")

    def pre_interaction(self) -> None:
        self._call("pre_interaction")

    def fluid_force(self) -> None:
        self._call("fluid_force")

    def gravity_force(self) -> None:
        self._call("gravity_force")

    def correct(self) -> None:
        self._call("correct")

    def integrate(self, fixed_dt: float = 0.0) -> None:
        self._call("integrate", C.c_double(fixed_dt))

    # ------------------------------------------------------------- scalars
    def time(self) -> tuple[float, float, float]:
        t, dt, hv = C.c_double(), C.c_double(), C.c_double()
        self._call("get_time", C.byref(t), C.byref(dt), C.byref(hv))
        return t.value, dt.value, hv.value

    def set_time(self, time: float, dt: float) -> None:
        self._call("set_time", C.c_double(time), C.c_double(dt))

    def energy(self) -> np.ndarray:
        out = np.zeros(8)
        self._call("energy", out.ctypes.data_as(C.POINTER(C.c_double)))
        return out[: 4 + self.dim]

    def stats(self) -> dict:
        st = _abi.Stats()
        self._call("get_stats", C.byref(st))
        return st.as_dict()

    def reset_stats(self) -> None:
        self._call("reset_stats")

    def set_record_sets(self, on: bool) -> None:
        self._call("set_record_sets", C.c_int(1 if on else 0))

    def neighbor_sets_packed(self, kind: int, cap: int | None = None) -> tuple[np.ndarray, np.ndarray]:
        """(offsets, members): the set of particle i is members[offsets[i]:offsets[i+1]]."""
        cap = cap or self.n * max(64, int(self.params.neighbor_number) * 8)
        counts = np.zeros(self.n, np.int32)
        lists = np.zeros(cap, np.int32)
        self._call("neighbor_sets", C.c_int(kind), counts.ctypes.data_as(C.POINTER(C.c_int32)),
                   lists.ctypes.data_as(C.POINTER(C.c_int32)), C.c_int64(cap))
        offs = np.concatenate([[0], np.cumsum(counts.astype(np.int64))])
        return offs, lists

    def neighbor_sets(self, kind: int, cap: int | None = None) -> list[np.ndarray]:
        offs, lists = self.neighbor_sets_packed(kind, cap)
        return [lists[offs[i]:offs[i + 1]].copy() for i in range(self.n)]

    def tree_info(self, cap: int | None = None) -> dict:
        cap = cap or 6 * self.n + 16
        d = self.dim
        nn = C.c_int64()
        level = np.zeros(cap, np.int32)
        count = np.zeros(cap, np.int32)
        edge = np.zeros(cap)
        center = np.zeros(cap * d)
        mass = np.zeros(cap)
        com = np.zeros(cap * d)
        ip = C.POINTER(C.c_int32)
        dp = C.POINTER(C.c_double)
        self._call("tree_info", C.byref(nn), level.ctypes.data_as(ip), count.ctypes.data_as(ip),
                   edge.ctypes.data_as(dp), center.ctypes.data_as(dp), mass.ctypes.data_as(dp),
                   com.ctypes.data_as(dp), C.c_int64(cap))
        k = nn.value
        return {"num_nodes": k, "level": level[:k], "count": count[:k], "edge": edge[:k],
                "center": center[: k * d].reshape(k, d), "mass": mass[:k], "com": com[: k * d].reshape(k, d)}
