"""Full-precision binary snapshots and JSON parameter files for sphgpu_run.

The snapshot layout is the one the C++ host layer reads and writes
(little endian): ``b"SPHSNAP2"``, int32 dim, int32 0, int64 n, float64 time,
float64 dt, float64 h_per_v_sig (absent in ``b"SPHSNAP1"`` files), then pos,
vel, vel_p, acc (n*dim float64 each), mass, dens, pres,
ene, ene_p, dene, sml, sound, balsara, alpha, gradh, phi (n float64 each), id,
neighbor (n int32 each).  It replaces the reference's lossy text snapshots
(src/output.cpp:14-64).

``params_to_json`` writes a parameter structure under the reference's JSON
keys (src/solver.cpp:155-299) plus this build's "DIM", "precision" and
"neighborCapacity".
"""
from __future__ import annotations

import json

import numpy as np

from . import _abi

MAGIC = b"SPHSNAP2"
MAGIC1 = b"SPHSNAP1"
VEC_FIELDS = ("pos", "vel", "vel_p", "acc")
SCALAR_FIELDS = ("mass", "dens", "pres", "ene", "ene_p", "dene", "sml", "sound",
                 "balsara", "alpha", "gradh", "phi")
INT_FIELDS = ("id", "neighbor")


def write(path, fields: dict, dim: int, time: float = 0.0, dt: float = 0.0, h_per_v_sig: float = 0.0) -> None:
    """Write a snapshot; absent fields are written as zeros (id as 0..n-1)."""
    n = int(np.asarray(fields["pos"]).shape[0])
    with open(path, "wb") as f:
        f.write(MAGIC)
        f.write(np.array([dim, 0], "<i4").tobytes())
        f.write(np.array([n], "<i8").tobytes())
        f.write(np.array([time, dt, h_per_v_sig], "<f8").tobytes())
        for k in VEC_FIELDS:
            a = fields.get(k)
            a = np.zeros((n, dim)) if a is None else np.asarray(a, dtype=np.float64).reshape(n, dim)
            f.write(np.ascontiguousarray(a, dtype="<f8").tobytes())
        for k in SCALAR_FIELDS:
            a = fields.get(k)
            a = np.zeros(n) if a is None else np.asarray(a, dtype=np.float64).reshape(n)
            f.write(np.ascontiguousarray(a, dtype="<f8").tobytes())
        for k in INT_FIELDS:
            a = fields.get(k)
            if a is None:
                a = np.arange(n) if k == "id" else np.zeros(n)
            f.write(np.ascontiguousarray(np.asarray(a).reshape(n), dtype="<i4").tobytes())


def read(path, with_hvs: bool = False):
    """Read a snapshot: (fields, dim, time, dt), plus h_per_v_sig last if
    ``with_hvs``."""
    with open(path, "rb") as f:
        buf = f.read()
    if buf[:8] not in (MAGIC, MAGIC1):
        raise ValueError(f"{path}: not a snapshot")
    dim, _ = np.frombuffer(buf, "<i4", 2, 8)
    (n,) = np.frombuffer(buf, "<i8", 1, 16)
    time, dt = np.frombuffer(buf, "<f8", 2, 24)
    v2 = buf[:8] == MAGIC
    hvs = float(np.frombuffer(buf, "<f8", 1, 40)[0]) if v2 else 0.0
    dim, n = int(dim), int(n)
    off = 48 if v2 else 40
    out = {}
    for k in VEC_FIELDS:
        out[k] = np.frombuffer(buf, "<f8", n * dim, off).reshape(n, dim).copy()
        off += 8 * n * dim
    for k in SCALAR_FIELDS:
        out[k] = np.frombuffer(buf, "<f8", n, off).copy()
        off += 8 * n
    for k in INT_FIELDS:
        out[k] = np.frombuffer(buf, "<i4", n, off).copy()
        off += 4 * n
    if off != len(buf):
        raise ValueError(f"{path}: {len(buf) - off} trailing bytes")
    if with_hvs:
        return out, dim, float(time), float(dt), hvs
    return out, dim, float(time), float(dt)


def params_to_json(p: _abi.Params, output_directory: str = "results") -> dict:
    """JSON document (as a dict) that sphgpu_run parses back into ``p``."""
    d = {
        "outputDirectory": output_directory,
        "DIM": int(p.dim),
        "precision": "mixed" if p.precision == _abi.MIXED else "fp64",
        "neighborCapacity": int(p.neighbor_capacity),
        "startTime": p.time_start,
        "endTime": p.time_end,
        "outputTime": p.time_output,
        "energyTime": p.time_energy,
        "SPHType": {_abi.SSPH: "ssph", _abi.DISPH: "disph", _abi.GSPH: "gsph"}[p.sph_type],
        "cflSound": p.cfl_sound,
        "cflForce": p.cfl_force,
        "avAlpha": p.av_alpha,
        "useBalsaraSwitch": bool(p.av_use_balsara_switch),
        "useTimeDependentAV": bool(p.av_use_time_dependent),
        "alphaMax": p.av_alpha_max,
        "alphaMin": p.av_alpha_min,
        "epsilonAV": p.av_epsilon,
        "useArtificialConductivity": bool(p.ac_is_valid),
        "alphaAC": p.ac_alpha,
        "maxTreeLevel": int(p.tree_max_level),
        "leafParticleNumber": int(p.tree_leaf_particle_num),
        "neighborNumber": int(p.neighbor_number),
        "gamma": p.gamma,
        "kernel": "wendland" if p.kernel == _abi.WENDLAND else "cubic_spline",
        "iterativeSmoothingLength": bool(p.iterative_sml),
        "periodic": bool(p.periodic_is_valid),
        "useGravity": bool(p.gravity_is_valid),
        "G": p.gravity_constant,
        "theta": p.gravity_theta,
        "use2ndOrderGSPH": bool(p.gsph_is_2nd_order),
    }
    if p.periodic_is_valid:
        d["rangeMax"] = [p.range_max[k] for k in range(p.dim)]
        d["rangeMin"] = [p.range_min[k] for k in range(p.dim)]
    return d


def write_params(path, p: _abi.Params, output_directory: str = "results") -> None:
    with open(path, "w") as f:
        # repr of a float round-trips exactly through strtod
        json.dump(params_to_json(p, output_directory), f, indent=1)
