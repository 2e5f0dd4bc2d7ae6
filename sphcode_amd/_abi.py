"""ctypes mirror of the C ABI in include/sphgpu.h.

The structures below must match the C declarations field for field; the
test ``tests/test_abi.py`` checks their sizes and offsets against a small C
program compiled from the header.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

SSPH, DISPH, GSPH = 0, 1, 2
CUBIC_SPLINE, WENDLAND = 0, 1
FP64, MIXED = 0, 1

E_OK = 0
E_INVALID = -1
E_HIP = -2
E_NO_DEVICE = -3
E_LIST_OVERFLOW = -4
E_TREE = -5
E_STATE = -6

ERROR_NAMES = {
    E_INVALID: "SPHGPU_E_INVALID",
    E_HIP: "SPHGPU_E_HIP",
    E_NO_DEVICE: "SPHGPU_E_NO_DEVICE",
    E_LIST_OVERFLOW: "SPHGPU_E_LIST_OVERFLOW",
    E_TREE: "SPHGPU_E_TREE",
    E_STATE: "SPHGPU_E_STATE",
}

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class Params(C.Structure):
    """sphgpu_params (mirrors SPHParameters, include/parameters.hpp:20-79)."""

    _fields_ = [
        ("dim", C.c_int32),
        ("sph_type", C.c_int32),
        ("kernel", C.c_int32),
        ("precision", C.c_int32),
        ("time_start", C.c_double),
        ("time_end", C.c_double),
        ("time_output", C.c_double),
        ("time_energy", C.c_double),
        ("cfl_sound", C.c_double),
        ("cfl_force", C.c_double),
        ("av_alpha", C.c_double),
        ("av_alpha_max", C.c_double),
        ("av_alpha_min", C.c_double),
        ("av_epsilon", C.c_double),
        ("av_use_balsara_switch", C.c_int32),
        ("av_use_time_dependent", C.c_int32),
        ("ac_alpha", C.c_double),
        ("ac_is_valid", C.c_int32),
        ("tree_max_level", C.c_int32),
        ("tree_leaf_particle_num", C.c_int32),
        ("neighbor_number", C.c_int32),
        ("gamma", C.c_double),
        ("iterative_sml", C.c_int32),
        ("periodic_is_valid", C.c_int32),
        ("range_max", C.c_double * 3),
        ("range_min", C.c_double * 3),
        ("gravity_is_valid", C.c_int32),
        ("gravity_constant", C.c_double),
        ("gravity_theta", C.c_double),
        ("gsph_is_2nd_order", C.c_int32),
        ("neighbor_capacity", C.c_int32),
        ("reserved0", C.c_int32),
    ]

    def copy(self) -> "Params":
        out = Params()
        C.memmove(C.byref(out), C.byref(self), C.sizeof(Params))
        return out


class Particles(C.Structure):
    """sphgpu_particles: structure-of-arrays view of SPHParticle."""

    _fields_ = [
        ("n", C.c_int64),
        ("pos", _dp), ("vel", _dp), ("vel_p", _dp), ("acc", _dp),
        ("mass", _dp), ("dens", _dp), ("pres", _dp), ("ene", _dp),
        ("ene_p", _dp), ("dene", _dp), ("sml", _dp), ("sound", _dp),
        ("balsara", _dp), ("alpha", _dp), ("gradh", _dp), ("phi", _dp),
        ("id", _ip), ("neighbor", _ip),
        ("grad_density", _dp), ("grad_pressure", _dp), ("grad_velocity", _dp),
    ]


class Stats(C.Structure):
    _fields_ = [(name, C.c_int64) for name in (
        "steps", "gather_pairs", "force_pairs", "gravity_pp", "gravity_pn",
        "newton_failures", "newton_iterations", "search_truncations",
        "list_overflows", "max_list_length", "tree_nodes", "tree_max_level", "gravity_opened",
        "search_nodes", "search_candidates", "gravity_wave_steps", "gravity_rewalks",
        "gravity_pp_slots", "gravity_max_records")]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class Timers(C.Structure):
    _fields_ = [(name, C.c_double) for name in (
        "timestep", "predict", "tree", "pre_interaction", "fluid_force", "gravity", "correct",
        "search", "pre_pairs", "force_pairs", "gravity_walk", "gravity_leaf")]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


VECTOR_FIELDS = ("pos", "vel", "vel_p", "acc")
SCALAR_FIELDS = ("mass", "dens", "pres", "ene", "ene_p", "dene", "sml", "sound",
                 "balsara", "alpha", "gradh", "phi")
INT_FIELDS = ("id", "neighbor")
GRAD_FIELDS = ("grad_density", "grad_pressure", "grad_velocity")
ALL_FIELDS = VECTOR_FIELDS + SCALAR_FIELDS + INT_FIELDS + GRAD_FIELDS


def field_shape(name: str, n: int, dim: int) -> tuple:
    if name in VECTOR_FIELDS or name in ("grad_density", "grad_pressure"):
        return (n, dim)
    if name == "grad_velocity":
        return (n, dim, dim)
    return (n,)


def field_dtype(name: str):
    return np.int32 if name in INT_FIELDS else np.float64


class ParticleArrays:
    """numpy storage behind a Particles struct (keeps the arrays alive)."""

    def __init__(self, n: int, dim: int, fields=ALL_FIELDS, data: dict | None = None):
        self.n = int(n)
        self.dim = int(dim)
        self.arrays: dict[str, np.ndarray] = {}
        data = data or {}
        for name in fields:
            if name in data and data[name] is not None:
                a = np.ascontiguousarray(np.asarray(data[name], dtype=field_dtype(name)))
                a = a.reshape(field_shape(name, self.n, self.dim))
            else:
                a = np.zeros(field_shape(name, self.n, self.dim), dtype=field_dtype(name))
            self.arrays[name] = a
        self.struct = Particles()
        self.struct.n = self.n
        for name, a in self.arrays.items():
            ptype = _ip if name in INT_FIELDS else _dp
            setattr(self.struct, name, a.ctypes.data_as(ptype))

    def __getitem__(self, name: str) -> np.ndarray:
        return self.arrays[name]

    def to_dict(self) -> dict:
        return {k: v.copy() for k, v in self.arrays.items()}


_PREFIXED = {
    # name: (restype, argtypes)
    "last_error": (C.c_char_p, []),
    "create": (C.c_int, [C.POINTER(Params), C.POINTER(C.c_void_p)]),
    "destroy": (C.c_int, [C.c_void_p]),
    "upload": (C.c_int, [C.c_void_p, C.POINTER(Particles)]),
    "download": (C.c_int, [C.c_void_p, C.POINTER(Particles)]),
    "initialize": (C.c_int, [C.c_void_p]),
    "timestep": (C.c_int, [C.c_void_p, _dp]),
    "predict": (C.c_int, [C.c_void_p]),
    "make_tree": (C.c_int, [C.c_void_p]),
    "pre_interaction": (C.c_int, [C.c_void_p]),
    "fluid_force": (C.c_int, [C.c_void_p]),
    "gravity_force": (C.c_int, [C.c_void_p]),
    "correct": (C.c_int, [C.c_void_p]),
    "integrate": (C.c_int, [C.c_void_p, C.c_double]),
    "get_time": (C.c_int, [C.c_void_p, _dp, _dp, _dp]),
    "set_time": (C.c_int, [C.c_void_p, C.c_double, C.c_double]),
    "energy": (C.c_int, [C.c_void_p, _dp]),
    "get_stats": (C.c_int, [C.c_void_p, C.POINTER(Stats)]),
    "reset_stats": (C.c_int, [C.c_void_p]),
    "set_record_sets": (C.c_int, [C.c_void_p, C.c_int]),
    "neighbor_sets": (C.c_int, [C.c_void_p, C.c_int, _ip, _ip, C.c_int64]),
    "tree_info": (C.c_int, [C.c_void_p, C.POINTER(C.c_int64), _ip, _ip, _dp, _dp, _dp, _dp, C.c_int64]),
}

_DP = C.POINTER(C.c_double)

GPU_ONLY = {
    "version": (C.c_int, []),
    "params_default": (C.c_int, [C.POINTER(Params)]),
    "get_timers": (C.c_int, [C.c_void_p, C.POINTER(Timers)]),
    "enable_timers": (C.c_int, [C.c_void_p, C.c_int]),
    "enable_stats": (C.c_int, [C.c_void_p, C.c_int]),
    "synchronize": (C.c_int, [C.c_void_p]),
    "rollout": (C.c_int, [C.c_void_p, C.c_int, C.c_double]),
    "restart": (C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_double]),
    "rollout_adjoint": (C.c_int, [C.c_void_p, _DP, _DP, _DP, _DP, _DP, _DP]),
}


def bind(lib: C.CDLL, prefix: str, extra: dict | None = None) -> dict:
    """Attach argtypes/restype to every prefixed entry point; return name->fn."""
    table = dict(_PREFIXED)
    if extra:
        table.update(extra)
    out = {}
    for name, (res, args) in table.items():
        fn = getattr(lib, prefix + name)
        fn.restype = res
        fn.argtypes = args
        out[name] = fn
    return out


def default_params(dim: int) -> Params:
    """The reference's JSON defaults (src/solver.cpp:155-299).

    gamma and endTime are required keys in the reference and are left 0 here.
    """
    p = Params()
    p.dim = dim
    p.sph_type = SSPH
    p.kernel = CUBIC_SPLINE
    p.precision = FP64
    p.time_start = 0.0
    p.time_end = 0.0
    p.time_output = 0.0
    p.time_energy = 0.0
    p.cfl_sound = 0.3
    p.cfl_force = 0.125
    p.av_alpha = 1.0
    p.av_use_balsara_switch = 1
    p.av_use_time_dependent = 0
    p.av_alpha_max = 2.0
    p.av_alpha_min = 0.1
    p.av_epsilon = 0.2
    p.ac_is_valid = 0
    p.ac_alpha = 1.0
    p.tree_max_level = 20
    p.tree_leaf_particle_num = 1
    p.neighbor_number = 32
    p.gamma = 0.0
    p.iterative_sml = 1
    p.periodic_is_valid = 0
    p.gravity_is_valid = 0
    p.gravity_constant = 1.0
    p.gravity_theta = 0.5
    p.gsph_is_2nd_order = 1
    p.neighbor_capacity = 0
    return p
