"""Seeded initial conditions for the benchmark and parity workloads.

Each generator returns ``(params, fields)``: an ``_abi.Params`` filled with
the workload's settings on top of the reference's JSON defaults
(src/solver.cpp:155-299) and a dict of per-particle numpy arrays (pos, vel,
mass, dens, pres, ene) in the layout of ``sphgpu_particles``.  As in the
reference's samples (src/sample/*.cpp), ``dens`` holds the initial density
guess that seeds the first smoothing-length estimate
(src/pre_interaction.cpp:171-215) and ``pres`` the initial pressure.

Workloads (sizes are the defaults; all accept smaller sizes for tests):
  sod       1D shock tube, SSPH, cubic spline, 1600 + 200 particles
  khi       2D Kelvin-Helmholtz, DISPH, Wendland C4, 512^2/2 + 724^2/2
  evrard    3D Evrard collapse with self-gravity, DISPH, Wendland C4, 2^20
  rollout   3D periodic jittered lattice, SSPH, cubic spline, 64^3
  sedov     3D periodic blast, GSPH 2nd order, Wendland C4, 256^3
"""
from __future__ import annotations

import math

import numpy as np

from . import _abi


def _base(dim: int, **kw) -> _abi.Params:
    p = _abi.default_params(dim)
    for k, v in kw.items():
        if k in ("range_max", "range_min"):
            arr = getattr(p, k)
            for i, x in enumerate(v):
                arr[i] = x
        else:
            setattr(p, k, v)
    return p


def sod(n_left: int = 1600, n_right: int = 200, precision: int = _abi.FP64):
    """1D Sod tube on [-0.5, 0.5], non-periodic, equal-mass particles.

    Left half: rho=1, P=1; right half: rho=0.125, P=0.1; v=0; gamma=1.4.
    Spacing follows from equal masses m = 0.5/n_left.
    """
    gamma = 1.4
    dxl = 0.5 / n_left
    dxr = 0.5 / n_right
    xl = -0.5 + (np.arange(n_left) + 0.5) * dxl
    xr = (np.arange(n_right) + 0.5) * dxr
    x = np.concatenate([xl, xr])
    n = x.size
    mass = np.full(n, 0.5 / n_left)
    dens = np.where(x < 0, 1.0, dxl / dxr)
    pres = np.where(x < 0, 1.0, 0.1)
    ene = pres / ((gamma - 1.0) * dens)
    p = _base(1, sph_type=_abi.SSPH, kernel=_abi.CUBIC_SPLINE, neighbor_number=4, gamma=gamma,
              time_end=0.2, precision=precision)
    fields = dict(pos=x.reshape(n, 1), vel=np.zeros((n, 1)), mass=mass, dens=dens, pres=pres, ene=ene)
    return p, fields


def khi(n_low: int = 512, n_high: int = 724, jitter: float = 1e-3, seed: int = 1,
        precision: int = _abi.FP64, sph_type: int = _abi.DISPH):
    """2D Kelvin-Helmholtz in the periodic unit square.

    An n_low^2 lattice is kept where |y-0.5| >= 0.25 (rho=1) and an
    n_high^2 lattice where |y-0.5| < 0.25 (rho ~ 2); equal masses; Gaussian
    position jitter of ``jitter`` lattice spacings; P=2.5; vx = +0.5 inside
    the band and -0.5 outside; vy = 0.1 sin(4 pi x) times Gaussians of width
    0.05 at both interfaces (the perturbation of src/sample/khi.cpp:33-36).
    """
    gamma = 5.0 / 3.0
    rng = np.random.default_rng(seed)

    def lattice(m):
        g = (np.arange(m) + 0.5) / m
        X, Y = np.meshgrid(g, g, indexing="xy")
        return np.stack([X.ravel(), Y.ravel()], axis=1)

    lo = lattice(n_low)
    lo = lo[np.abs(lo[:, 1] - 0.5) >= 0.25]
    hi = lattice(n_high)
    hi = hi[np.abs(hi[:, 1] - 0.5) < 0.25]
    lo += rng.normal(0.0, jitter / n_low, lo.shape)
    hi += rng.normal(0.0, jitter / n_high, hi.shape)
    pos = np.mod(np.concatenate([lo, hi]), 1.0)
    n = pos.shape[0]
    inside = np.concatenate([np.zeros(len(lo), bool), np.ones(len(hi), bool)])
    rho_hi = (n_high / n_low) ** 2
    mass = np.full(n, (0.5 + 0.5 * rho_hi) / n)
    dens = np.where(inside, rho_hi, 1.0)
    pres = np.full(n, 2.5)
    ene = pres / ((gamma - 1.0) * dens)
    vel = np.zeros((n, 2))
    vel[:, 0] = np.where(inside, 0.5, -0.5)
    s2 = 0.05 ** 2
    x, y = pos[:, 0], pos[:, 1]
    vel[:, 1] = 0.1 * np.sin(4 * math.pi * x) * (np.exp(-(y - 0.25) ** 2 / (2 * s2)) + np.exp(-(y - 0.75) ** 2 / (2 * s2)))
    p = _base(2, sph_type=sph_type, kernel=_abi.WENDLAND, neighbor_number=64, gamma=gamma,
              periodic_is_valid=1, range_max=(1.0, 1.0, 0.0), range_min=(0.0, 0.0, 0.0),
              av_use_time_dependent=1, tree_leaf_particle_num=16, time_end=3.0, precision=precision)
    fields = dict(pos=pos, vel=vel, mass=mass, dens=dens, pres=pres, ene=ene)
    return p, fields


def evrard(n: int = 1 << 20, seed: int = 1, jitter: float = 0.05, leaf: int = 32,
           precision: int = _abi.FP64):
    """3D Evrard collapse: total mass 1, radius 1, rho ~ 1/r, u = 0.05 G.

    A cubic lattice with Gaussian jitter (``jitter`` spacings) is cut to the
    ``n`` points nearest the centre, rescaled to unit radius and remapped
    r -> r^(3/2) (the remap of src/sample/evrard.cpp:36-39), which turns a
    uniform sphere into rho = 1/(2 pi r).  The density guess is 1/(2 pi r)
    (src/sample/evrard.cpp:58).  DISPH, Balsara and time-dependent AV,
    Wendland C4, 64 neighbours, G = 1, theta = 0.5, leaf size ``leaf``.
    """
    gamma = 5.0 / 3.0
    G = 1.0
    rng = np.random.default_rng(seed)
    side = int(math.ceil((6.0 / math.pi * n) ** (1.0 / 3.0))) + 2
    while True:
        dx = 2.0 / side
        g = (np.arange(side) + 0.5) * dx - 1.0
        X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
        pts = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
        pts += rng.normal(0.0, jitter * dx, pts.shape)
        r = np.sqrt((pts * pts).sum(1))
        if np.count_nonzero(r <= 1.0) >= n:
            break
        side += 1
    order = np.argsort(r, kind="stable")[:n]
    pts = pts[order]
    r = r[order]
    pts /= r.max()
    r = np.sqrt((pts * pts).sum(1))
    scale = np.where(r > 0, np.sqrt(r), 1.0)  # r^(3/2) / r
    pos = pts * scale[:, None]
    rr = np.sqrt((pos * pos).sum(1))
    mass = np.full(n, 1.0 / n)
    dens = 1.0 / (2.0 * math.pi * rr)
    u = 0.05 * G
    ene = np.full(n, u)
    pres = (gamma - 1.0) * dens * u
    p = _base(3, sph_type=_abi.DISPH, kernel=_abi.WENDLAND, neighbor_number=64, gamma=gamma,
              av_use_balsara_switch=1, av_use_time_dependent=1, gravity_is_valid=1,
              gravity_constant=G, gravity_theta=0.5, tree_leaf_particle_num=leaf, time_end=3.0,
              precision=precision)
    fields = dict(pos=pos, vel=np.zeros((n, 3)), mass=mass, dens=dens, pres=pres, ene=ene)
    return p, fields


def rollout(n_side: int = 64, seed: int = 1, precision: int = _abi.FP64, dim: int = 3):
    """Periodic unit box for the differentiable rollout (3D by default).

    n_side^dim lattice with Gaussian jitter of 0.05 spacings, rho = 1,
    u = 1 (1 + 0.1 (2U - 1)), velocities N(0, 0.1^2) per axis, gamma = 5/3,
    SSPH with the cubic spline and 50 neighbours in 3D (24 in 2D, 6 in 1D).
    """
    gamma = 5.0 / 3.0
    rng = np.random.default_rng(seed)
    dx = 1.0 / n_side
    g = (np.arange(n_side) + 0.5) * dx
    grids = np.meshgrid(*([g] * dim), indexing="ij")
    pos = np.stack([a.ravel() for a in grids], axis=1)
    pos = np.mod(pos + rng.normal(0.0, 0.05 * dx, pos.shape), 1.0)
    n = pos.shape[0]
    mass = np.full(n, 1.0 / n)
    dens = np.ones(n)
    ene = 1.0 + 0.1 * (2.0 * rng.random(n) - 1.0)
    pres = (gamma - 1.0) * dens * ene
    vel = rng.normal(0.0, 0.1, (n, dim))
    nngb = {1: 6, 2: 24, 3: 50}[dim]
    p = _base(dim, sph_type=_abi.SSPH, kernel=_abi.CUBIC_SPLINE, neighbor_number=nngb, gamma=gamma,
              periodic_is_valid=1, range_max=(1.0,) * dim, range_min=(0.0,) * dim,
              time_end=1.0, precision=precision)
    fields = dict(pos=pos, vel=vel, mass=mass, dens=dens, pres=pres, ene=ene)
    return p, fields


def sedov(n_side: int = 256, seed: int = 1, n_hot: int = 64, energy: float = 1.0,
          precision: int = _abi.FP64, u_background: float = 1e-5):
    """3D Sedov-type blast in the periodic unit cube, GSPH second order.

    n_side^3 lattice with Gaussian jitter of 0.05 spacings, rho = 1,
    background u = 1e-5; ``energy`` is deposited into the ``n_hot`` particles
    nearest the centre with Wendland-C4 weights of support equal to the
    distance of the farthest of them.  Wendland C4, 64 neighbours,
    gamma = 5/3.
    """
    gamma = 5.0 / 3.0
    rng = np.random.default_rng(seed)
    dx = 1.0 / n_side
    g = (np.arange(n_side) + 0.5) * dx
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    pos = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    pos = np.mod(pos + rng.normal(0.0, 0.05 * dx, pos.shape), 1.0)
    n = pos.shape[0]
    mass = np.full(n, 1.0 / n)
    dens = np.ones(n)
    ene = np.full(n, u_background)
    d = pos - 0.5
    r = np.sqrt((d * d).sum(1))
    hot = np.argpartition(r, n_hot)[:n_hot]
    h = r[hot].max() * 1.0000001
    q = r[hot] / h
    w = np.maximum(1.0 - q, 0.0) ** 6 * (1.0 + 6.0 * q + 35.0 / 3.0 * q * q)
    ene[hot] += energy * w / (w.sum() * mass[hot])
    pres = (gamma - 1.0) * dens * ene
    p = _base(3, sph_type=_abi.GSPH, kernel=_abi.WENDLAND, neighbor_number=64, gamma=gamma,
              periodic_is_valid=1, range_max=(1.0, 1.0, 1.0), range_min=(0.0, 0.0, 0.0),
              gsph_is_2nd_order=1, time_end=0.1, precision=precision)
    fields = dict(pos=pos, vel=np.zeros((n, 3)), mass=mass, dens=dens, pres=pres, ene=ene)
    return p, fields


WORKLOADS = {"sod": sod, "khi": khi, "evrard": evrard, "rollout": rollout, "sedov": sedov}
