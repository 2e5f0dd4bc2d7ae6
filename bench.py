#!/usr/bin/env python3
"""Evrard-collapse throughput of the per-step SPH pipeline on one MI355X.

Workload: 3D Evrard collapse, N = 1,048,576 equal-mass particles, DISPH
with Balsara switch and time-dependent viscosity, Wendland C4, 64
neighbours, self-gravity with theta = 0.5, leafParticleNumber 32; seeded
synthetic initial condition (sphcode_amd.ics.evrard).  A step is one full
Solver::integrate (timestep, predict, tree, pre-interaction, fluid force,
gravity, correct) with the state resident in HBM.

Prints one JSON line: particle-steps/s plus the roofline of the dominant
kernel (measured live from HIP events on the context stream) and a CPU
baseline (the oracle built with the reference's compiler flags, timed on
this host's cores).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import pathlib
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "particle-steps/s Evrard N=1M, mixed vs FP64 at stated err; adjoint/fwd ratio"


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(params, fields, n: int, steps: int) -> dict:
    """Time the oracle (reference flags: -O3 -funroll-loops -ffast-math
    -fopenmp) on this host: initialize untimed, then `steps` timed steps."""
    from sphcode_amd import _abi
    from sphcode_amd.engine import Engine
    lib_path = ROOT / "oracle" / "lib" / "libsphref3_fast.so"
    lib = C.CDLL(str(lib_path))
    extra = {"set_threads": (C.c_int, [C.c_void_p, C.c_int])}
    fns = _abi.bind(lib, "sphref_", extra)
    e = Engine(fns, params, name="sphref")
    threads = fns["set_threads"](None, 0)
    e.upload(fields)
    t0 = time.perf_counter()
    e.initialize()
    t_init = time.perf_counter() - t0
    log(f"[cpu] initialize {t_init:.2f} s with {threads} threads")
    t0 = time.perf_counter()
    for _ in range(steps):
        e.integrate()
    dt = time.perf_counter() - t0
    e.close()
    rate = n * steps / dt
    log(f"[cpu] {steps} step(s) {dt:.2f} s -> {rate:.4g} particle-steps/s")
    return {"value": rate, "unit": "particle-steps/s", "cores": int(threads), "kind": "port",
            "sample": f"oracle (reference flags), Evrard N={n}: {steps} full step(s) after initialize, "
                      f"{threads} OpenMP threads, {dt:.1f} s timed"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--precision", choices=("fp64", "mixed"), default="fp64")
    ap.add_argument("--cpu-steps", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is described in DESIGN.md, "
                                   "not implemented", "n_gpus": args.gpus}))
        return 2

    from sphcode_amd import FP64, MIXED, GPUEngine, costs, ics

    prec = MIXED if args.precision == "mixed" else FP64
    t0 = time.perf_counter()
    params, fields = ics.evrard(args.n, leaf=32, precision=prec)
    log(f"[bench] initial condition N={args.n} in {time.perf_counter() - t0:.1f} s")
    g = GPUEngine(params)
    g.upload(fields)
    t0 = time.perf_counter()
    g.initialize()
    log(f"[bench] initialize {time.perf_counter() - t0:.2f} s")
    for k in range(args.warmup):
        g.integrate()
        log(f"[bench] warmup step {k + 1}/{args.warmup} t={g.time()[0]:.6g}")
    g.synchronize()
    g.reset_stats()
    g.enable_timers(True)
    g.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        g.integrate()
    g.synchronize()
    elapsed = time.perf_counter() - t0
    n = args.n
    value = n * args.steps / elapsed
    stats = g.stats()
    timers = g.timers()
    log(f"[bench] {args.steps} steps in {elapsed:.3f} s -> {value:.4g} particle-steps/s")
    log("[bench] stage ms/step: " + ", ".join(f"{k}={v / args.steps:.3f}" for k, v in timers.items()))
    log(f"[bench] stats: {stats}")

    # Dominant kernel: the gravity walk (one launch per step).
    grav_s = timers["gravity"] / args.steps * 1e-3
    flops = costs.gravity_flops(stats) / args.steps
    achieved = flops / grav_s / 1e12 if grav_s > 0 else 0.0
    peak = costs.PEAK_FP64_TFLOPS if prec == FP64 else costs.PEAK_FP32_TFLOPS
    traffic = None
    tfile = ROOT / "profiles" / f"traffic_gravity_evrard_{args.precision}.json"
    if tfile.exists():
        try:
            traffic = json.loads(tfile.read_text()).get("hbm_bytes_per_launch")
        except Exception:
            traffic = None
    roofline = {"bound": "valu", "kernel": "gravity_kernel<3>", "achieved": achieved, "peak": peak,
                "unit": "TFLOP/s", "frac": achieved / peak, "traffic": traffic,
                "flops_per_launch": flops, "launch_ms": grav_s * 1e3}
    g.close()

    cpu = None
    if not args.no_cpu:
        cparams, cfields = ics.evrard(n, leaf=32, precision=FP64)
        cpu = cpu_baseline(cparams, cfields, n, args.cpu_steps)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == FP64 else "f32+f64",
        "data": "synthetic",
        "config": {"workload": "evrard_collapse", "n_particles": n, "sph_type": "disph", "kernel": "wendland_c4",
                   "neighbor_number": 64, "gravity_theta": 0.5, "leaf_particle_num": 32,
                   "precision": args.precision, "balsara": True, "time_dependent_av": True},
        "stage_ms_per_step": {k: v / args.steps for k, v in timers.items()},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
