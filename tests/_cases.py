"""Small seeded workloads shared by the oracle and GPU parity tests.

Each case returns (params, fields) from sphcode_amd.ics at a size the CPU
oracle finishes in about a second.  Every (DIM, SPHType, kernel) pair the
benchmark configurations use appears at least once, plus the second-order
and first-order GSPH variants and artificial conductivity.
"""
from __future__ import annotations

from sphcode_amd import _abi, ics


def sod_ssph():
    return ics.sod(400, 50)


def sod_disph():
    p, f = ics.sod(400, 50)
    p.sph_type = _abi.DISPH
    return p, f


def sod_gsph2():
    p, f = ics.sod(400, 50)
    p.sph_type = _abi.GSPH
    p.gsph_is_2nd_order = 1
    return p, f


def sod_gsph1():
    p, f = ics.sod(400, 50)
    p.sph_type = _abi.GSPH
    p.gsph_is_2nd_order = 0
    return p, f


def sod_ac_tdav():
    p, f = ics.sod(400, 50)
    p.ac_is_valid = 1
    p.ac_alpha = 1.0
    p.av_use_time_dependent = 1
    return p, f


def khi_disph():
    return ics.khi(n_low=32, n_high=45, jitter=0.05)


def khi_ssph():
    p, f = ics.khi(n_low=32, n_high=45, jitter=0.05, sph_type=_abi.SSPH)
    p.kernel = _abi.CUBIC_SPLINE
    p.neighbor_number = 32
    return p, f


def evrard_disph():
    return ics.evrard(4096, leaf=32)


def evrard_ssph():
    p, f = ics.evrard(2048, leaf=8)
    p.sph_type = _abi.SSPH
    p.kernel = _abi.CUBIC_SPLINE
    p.neighbor_number = 50
    p.av_use_time_dependent = 0
    return p, f


def evrard_ac():
    # artificial conductivity with gravity on: its signal speed is then
    # |v_ij . r_ij| / r (src/fluid_force.cpp:111-112); an infall velocity
    # field with noise makes it nonzero from the first force evaluation
    import numpy as np
    p, f = ics.evrard(4096, leaf=32)
    p.ac_is_valid = 1
    p.ac_alpha = 1.0
    x = f["pos"]
    r = np.sqrt((x * x).sum(1))[:, None]
    rng = np.random.default_rng(5)
    f = dict(f)
    f["vel"] = -0.2 * x / np.maximum(r, 1e-3) + rng.normal(0.0, 0.05, x.shape)
    return p, f


def rollout_ssph():
    return ics.rollout(n_side=12)


def rollout_dense():
    # 200 neighbours in 3D: a search tile then holds far more (lane,
    # candidate) hits than the batched append takes at once, so the
    # one-lane-at-a-time append runs beside the batched one
    p, f = ics.rollout(n_side=10)
    p.neighbor_number = 200
    return p, f


def sedov_gsph():
    # A milder contrast than the 256^3 benchmark: with u_bg = 1e-5 and only
    # 1728 particles the reference scheme itself drives u negative within
    # one step (MUSCL reconstruction of an extreme jump).
    return ics.sedov(n_side=12, n_hot=32, u_background=1e-2)


CASES = {
    "sod_ssph": sod_ssph,
    "sod_disph": sod_disph,
    "sod_gsph2": sod_gsph2,
    "sod_gsph1": sod_gsph1,
    "sod_ac_tdav": sod_ac_tdav,
    "khi_disph": khi_disph,
    "khi_ssph": khi_ssph,
    "evrard_disph": evrard_disph,
    "evrard_ssph": evrard_ssph,
    "evrard_ac": evrard_ac,
    "rollout_ssph": rollout_ssph,
    "rollout_dense": rollout_dense,
    "sedov_gsph": sedov_gsph,
}
