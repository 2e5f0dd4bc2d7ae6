#!/usr/bin/env python3
"""Evrard-collapse throughput of the per-step SPH pipeline on one MI355X.

Workload: 3D Evrard collapse, N = 1,048,576 equal-mass particles, DISPH
with Balsara switch and time-dependent viscosity, Wendland C4, 64
neighbours, self-gravity with theta = 0.5, leafParticleNumber 32; seeded
synthetic initial condition (sphcode_amd.ics.evrard).  A step is one full
Solver::integrate (timestep, predict, tree, pre-interaction, fluid force,
gravity, correct) with the state resident in HBM.

`value` is the mixed-precision throughput (FP32 pair arithmetic, FP64
kick-drift, Newton test and reductions); the FP64 throughput and the
mixed-vs-FP64 errors on the same initial state are reported beside it.
Also reported: the roofline of the dominant kernel (gravity), measured live
from HIP events on the context stream, and a CPU baseline (the oracle
built with the reference's compiler flags, timed on this host's cores).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import pathlib
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "particle-steps/s Evrard N=1M, mixed vs FP64 at stated err; adjoint/fwd ratio"


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def host_cpu() -> dict:
    """CPU model, sockets, cores per socket and threads per core (lscpu)."""
    import subprocess
    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=20).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            info[k.strip()] = v.strip()
    except Exception:
        pass
    return {"model": info.get("Model name"), "sockets": info.get("Socket(s)"),
            "cores_per_socket": info.get("Core(s) per socket"), "threads_per_core": info.get("Thread(s) per core"),
            "logical_cpus": info.get("CPU(s)")}


def cpu_baseline(params, fields, n: int, steps: int, warmup: int = 1) -> dict:
    """Time the oracle (reference flags: -O3 -funroll-loops -ffast-math
    -fopenmp) on this host's cores: initialize and `warmup` steps untimed,
    then `steps` timed steps, with the oracle's per-stage seconds."""
    import numpy as np
    from sphcode_amd import _abi
    from sphcode_amd.engine import Engine
    lib = C.CDLL(str(ROOT / "oracle" / "lib" / "libsphref3_fast.so"))
    fns = _abi.bind(lib, "sphref_", {"set_threads": (C.c_int, [C.c_void_p, C.c_int]),
                                      "stage_seconds": (C.c_int, [C.c_void_p, C.POINTER(C.c_double)])})
    e = Engine(fns, params, name="sphref")
    threads = fns["set_threads"](None, 0)
    e.upload(fields)
    t0 = time.perf_counter()
    e.initialize()
    log(f"[cpu] initialize {time.perf_counter() - t0:.2f} s with {threads} threads")
    for _ in range(warmup):
        e.integrate()
    e.reset_stats()
    t0 = time.perf_counter()
    for _ in range(steps):
        e.integrate()
    dt = time.perf_counter() - t0
    st = np.zeros(7)
    e._call("stage_seconds", st.ctypes.data_as(C.POINTER(C.c_double)))
    e.close()
    rate = n * steps / dt
    stages = dict(zip(("timestep", "predict", "tree", "pre_interaction", "fluid_force", "gravity", "correct"),
                      (float(x) / steps for x in st)))
    log(f"[cpu] {steps} step(s) {dt:.2f} s -> {rate:.4g} particle-steps/s; per-stage s/step {stages}")
    return {"value": rate, "unit": "particle-steps/s", "cores": int(threads), "kind": "port",
            "sample": f"oracle built with the reference's flags, Evrard N={n}, FP64: initialize and {warmup} "
                      f"warm-up step(s) untimed, then {steps} timed steps on {threads} OpenMP threads "
                      f"({dt:.1f} s timed)",
            "host": host_cpu(), "stage_s_per_step": stages}


def _steps(g, fields, steps: int, warmup: int, timers: bool) -> tuple[float, float]:
    """initialize, `warmup` untimed steps, then `steps` steps: the time of
    the upload and initialize, and of the steps."""
    t0 = time.perf_counter()
    g.upload(fields)
    g.initialize()
    g.synchronize()
    t_init = time.perf_counter() - t0
    for _ in range(warmup):
        g.integrate()
    g.synchronize()
    g.reset_stats()
    g.enable_timers(timers)
    g.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        g.integrate()
    g.synchronize()
    return t_init, time.perf_counter() - t0


def timed_run(params, fields, steps: int, warmup: int, label: str, gravity_alone: int = 0):
    """Times `steps` sphgpu_integrate calls after `warmup` untimed ones, with
    the stage and kernel-span timers (HIP events) on and the diagnostic
    counters off (their default in the library).  The interaction counts of
    those steps come from a second run of the same steps with the counters
    on: every step is deterministic, so its counts are the timed run's.
    With gravity_alone > 0 (and gravity on), the gravity stage is then also
    timed on its own: that many GravityForce stage calls (walk, pair pass
    and combine in order on one stream, nothing beside them) on the state
    the timed steps left, counted the same way; each engine is discarded
    afterwards (each call adds gravity to acc again)."""
    from sphcode_amd import GPUEngine, costs
    grav = gravity_alone > 0 and params.gravity_is_valid

    def alone_calls(g):
        g.gravity_force()  # untimed: packs the records once
        g.synchronize()
        s0 = g.stats()
        t0 = time.perf_counter()
        for _ in range(gravity_alone):
            g.gravity_force()
        g.synchronize()
        dt = time.perf_counter() - t0
        s1 = g.stats()
        return dt, {k: s1[k] - s0[k] for k in ("gravity_pp", "gravity_pn", "gravity_opened")}

    g = GPUEngine(params, stats=False)
    t_init, elapsed = _steps(g, fields, steps, warmup, True)
    log(f"[bench] {label}: upload and initialize {t_init:.2f} s")
    timers = g.timers()
    dt_alone = alone_calls(g)[0] if grav else None
    g.close()
    # the same steps with the counters on, for the interaction counts
    g = GPUEngine(params, stats=True)
    _steps(g, fields, steps, warmup, False)
    stats = g.stats()
    d = alone_calls(g)[1] if grav else None
    g.close()
    n = len(fields["mass"])
    log(f"[bench] {label}: {steps} steps in {elapsed:.3f} s -> {n * steps / elapsed:.4g} particle-steps/s")
    log(f"[bench] {label}: stage ms/step: " + ", ".join(f"{k}={v / steps:.3f}" for k, v in timers.items()))
    log(f"[bench] {label}: stats: {stats}")
    if grav:
        alone = {"calls": gravity_alone, "launch_ms": dt_alone / gravity_alone * 1e3,
                 "flops_per_launch": costs.gravity_flops(d) / gravity_alone}
        log(f"[bench] {label}: gravity stage alone {alone['launch_ms']:.3f} ms")
        timers = dict(timers)
        timers["_gravity_alone"] = alone
    return elapsed, stats, timers


def adjoint_run(n_side: int, steps: int, dt: float, precision: int) -> dict:
    """Differentiable rollout on the 3D periodic box (n_side^3 particles,
    SSPH cubic spline, 50 neighbours): time of the plain forward
    (Solver::initialize + `steps` fixed-dt steps), of the checkpointed
    forward, and of the adjoint of L = sum m |v_K|^2."""
    import numpy as np
    from sphcode_amd import FP64, GPUEngine, ics
    p, f = ics.rollout(n_side=n_side, precision=precision)
    n = n_side ** 3
    g = GPUEngine(p, stats=False)

    def plain():
        g.upload(f)
        g.initialize()
        for _ in range(steps):
            g.integrate(dt)
        g.synchronize()

    plain()  # warm-up (allocation, code load)
    t0 = time.perf_counter()
    plain()
    t_plain = time.perf_counter() - t0
    g.upload(f)
    g.rollout(steps, dt)  # warm-up of the checkpoint buffers
    g.upload(f)
    g.synchronize()
    t0 = time.perf_counter()
    g.rollout(steps, dt)
    g.synchronize()
    t_fwd = time.perf_counter() - t0
    out = g.download(("mass", "vel"))
    seed = 2.0 * out["mass"][:, None] * out["vel"]
    g.rollout_adjoint(dl_dv=seed)  # warm-up
    g.upload(f)
    g.rollout(steps, dt)
    g.synchronize()
    t0 = time.perf_counter()
    gx, gv, gu = g.rollout_adjoint(dl_dv=seed)
    t_adj = time.perf_counter() - t0
    g.close()
    label = "fp64" if precision == FP64 else "mixed"
    log(f"[bench] adjoint {label}: N={n} K={steps} forward {t_plain * 1e3:.1f} ms, checkpointed forward "
        f"{t_fwd * 1e3:.1f} ms, adjoint {t_adj * 1e3:.1f} ms")
    return {"workload": "periodic_box_rollout", "n_particles": n, "steps": steps, "dt": dt, "precision": label,
            "loss": "sum m |v_K|^2", "forward_ms": t_plain * 1e3, "checkpointed_forward_ms": t_fwd * 1e3,
            "adjoint_ms": t_adj * 1e3, "adjoint_over_forward": t_adj / t_plain}, (gx, gv, gu)


def workload_rate(name: str, params, fields, steps: int, warmup: int) -> dict:
    """Throughput of one of the other workloads (mixed precision)."""
    n = int(len(fields["mass"]))
    el, stats, timers = timed_run(params, fields, steps, warmup, name)
    rate = n * steps / el
    log(f"[bench] {name}: N={n} {rate:.4g} particle-steps/s")
    return {"n_particles": n, "steps": steps, "particle_steps_per_s": rate, "ms_per_step": el / steps * 1e3,
            "stage_ms_per_step": {k: v / steps for k, v in timers.items()},
            "force_pairs_per_s": stats["force_pairs"] / el}


def first_step_rate(name: str, params, fields, repeats: int) -> dict:
    """Time of the first integrate() after Solver::initialize: one untimed
    warm-up repetition, then `repeats` timed ones, each from a fresh upload
    of the same initial condition; the median is reported."""
    import statistics

    from sphcode_amd import GPUEngine
    n = int(len(fields["mass"]))
    g = GPUEngine(params, stats=False)
    times, timers = [], None
    for r in range(repeats + 1):
        g.upload(fields)
        g.initialize()
        g.synchronize()
        g.reset_stats()
        g.enable_timers(True)
        g.synchronize()
        t0 = time.perf_counter()
        g.integrate()
        g.synchronize()
        el = time.perf_counter() - t0
        if r > 0:
            times.append(el)
            timers = g.timers()
    # the same step with the counters on, for its pair count
    g.enable_stats(True)
    g.upload(fields)
    g.initialize()
    g.reset_stats()
    g.integrate()
    stats = g.stats()
    g.close()
    el = statistics.median(times)
    log(f"[bench] {name}: first step {[round(t * 1e3, 2) for t in times]} ms -> {n / el:.4g} particle-steps/s")
    return {"n_particles": n, "steps": 1, "repeats": repeats, "particle_steps_per_s": n / el,
            "ms_per_step": el * 1e3, "stage_ms_per_step": timers, "force_pairs_per_s": stats["force_pairs"] / el}


def kernel_rooflines(n: int, stats: dict, timers: dict, steps: int, peak_valu: float) -> dict:
    """Per-kernel figures from the kernel-span timers: the neighbour search
    and the two hydrodynamic pair passes against the HBM peak (algorithmic
    bytes: positions and radii in, the neighbour set out; the stage bytes of
    the pre-interaction and force), and the pair passes and gravity kernels
    against the vector peak (algorithmic FLOPs from the pair counters)."""
    from sphcode_amd import costs
    pre_f, force_f = costs.pair_flops(stats, n, steps, "disph")
    out = {}

    def entry(name, ms, nbytes=None, flops=None):
        ms = ms / steps
        e = {"launch_ms": ms}
        if nbytes is not None:
            gbs = nbytes / (ms * 1e-3) / 1e9
            e.update({"bound": "hbm", "bytes_per_launch": nbytes, "achieved_GBs": gbs,
                      "frac_hbm": gbs / costs.PEAK_HBM_GBS})
        if flops is not None:
            tf = flops / (ms * 1e-3) / 1e12
            e.update({"flops_per_launch": flops, "achieved_TFLOPs": tf, "frac_valu": tf / peak_valu})
        out[name] = e

    entry("search_kernel", timers["search"], nbytes=costs.search_bytes(n, stats, steps))
    entry("pre_kernel", timers["pre_pairs"], nbytes=costs.BYTES_PER_PARTICLE["pre_interaction"] * n, flops=pre_f)
    entry("force_kernel", timers["force_pairs"], nbytes=costs.BYTES_PER_PARTICLE["fluid_force"] * n, flops=force_f)
    walk_f = (costs.GRAV_FLOP_PN * stats["gravity_pn"] +
              costs.GRAV_FLOP_OPEN * (stats["gravity_pn"] + stats["gravity_opened"])) / steps
    entry("gravity_walk_kernel", timers["gravity_walk"], flops=walk_f)
    entry("gravity_leaf_kernel", timers["gravity_leaf"], flops=costs.GRAV_FLOP_PP * stats["gravity_pp"] / steps)
    log("[bench] kernels: " + ", ".join(f"{k} {v['launch_ms']:.3f} ms" for k, v in out.items()))
    return out


def mixed_errors(params_fp64, params_mixed, fields) -> dict:
    """Per-field error of mixed against FP64 after Solver::initialize on the
    benchmark's initial state."""
    import numpy as np

    from sphcode_amd import GPUEngine
    out = {}
    res = []
    for p in (params_fp64, params_mixed):
        g = GPUEngine(p)
        g.upload(fields)
        g.initialize()
        res.append(g.download(("dens", "sml", "pres", "acc", "phi")))
        g.close()
    a, b = res
    rel = lambda x, y: float(np.max(np.abs(x - y) / np.maximum(np.abs(y), 1e-300)))
    rms = lambda x, y: float(np.sqrt(np.mean((x - y) ** 2)) / max(np.sqrt(np.mean(y ** 2)), 1e-300))
    out["dens_max_rel"] = rel(b["dens"], a["dens"])
    out["sml_max_rel"] = rel(b["sml"], a["sml"])
    out["pres_max_rel"] = rel(b["pres"], a["pres"])
    if params_fp64.gravity_is_valid:
        out["phi_max_rel"] = rel(b["phi"], a["phi"])
    out["acc_rms_rel"] = rms(b["acc"], a["acc"])
    out["bars"] = {"dens_max_rel": 1e-5, "acc_rms_rel": 1e-4}
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--precision", choices=("mixed", "fp64"), default="mixed")
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-compare", action="store_true", help="skip the FP64 run and the error check")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the differentiable-rollout timing")
    ap.add_argument("--adjoint-side", type=int, default=64)
    ap.add_argument("--adjoint-steps", type=int, default=16)
    ap.add_argument("--adjoint-only", action="store_true", help="time only the differentiable rollout (profiling)")
    ap.add_argument("--no-workloads", action="store_true", help="skip the KHI and Sedov throughput lines")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="first run one device-to-device copy of this many MiB (calibrates HBM byte counters)")
    args = ap.parse_args()
    if args.copy_mib > 0:
        import torch
        src = torch.ones(args.copy_mib * (1 << 20) // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        log(f"[bench] copy calibration: {args.copy_mib} MiB copied")
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; sharding over several GPUs (Morton-key ranges "
                                   "with halo exchange) is designed but not implemented", "n_gpus": args.gpus}))
        return 2

    from sphcode_amd import FP64, MIXED, costs, ics

    if args.adjoint_only:
        adj, _ = adjoint_run(args.adjoint_side, args.adjoint_steps, 1e-3,
                             MIXED if args.precision == "mixed" else FP64)
        print(json.dumps({"adjoint": adj}), flush=True)
        return 0

    n = args.n
    t0 = time.perf_counter()
    p64, fields = ics.evrard(n, leaf=32, precision=FP64)
    pmx = p64.copy()
    pmx.precision = MIXED
    log(f"[bench] initial condition N={n} in {time.perf_counter() - t0:.1f} s")
    main_p = pmx if args.precision == "mixed" else p64

    errors = None
    fp64_rate = None
    if not args.no_compare:
        errors = mixed_errors(p64, pmx, fields)
        log(f"[bench] mixed vs FP64 after initialize: {errors}")
        if args.precision == "mixed":
            el64, _, _ = timed_run(p64, fields, args.steps, args.warmup, "fp64")
            fp64_rate = n * args.steps / el64

    elapsed, stats, timers = timed_run(main_p, fields, args.steps, args.warmup, args.precision, gravity_alone=5)
    grav_alone = timers.pop("_gravity_alone", None)
    value = n * args.steps / elapsed
    if args.precision == "fp64":
        fp64_rate = value

    # dominant kernel: the gravity pass (a tree walk and a leaf-pair pass per
    # step, timed on the stream they run on)
    grav_s = timers["gravity"] / args.steps * 1e-3
    flops = costs.gravity_flops(stats) / args.steps
    achieved = flops / grav_s / 1e12 if grav_s > 0 else 0.0
    peak = costs.PEAK_FP32_TFLOPS if args.precision == "mixed" else costs.PEAK_FP64_TFLOPS
    roofline = {"bound": "valu", "kernel": "gravity_walk_kernel<3> + gravity_leaf_kernel<3>", "achieved": achieved,
                "peak": peak,
                "unit": "TFLOP/s", "frac": achieved / peak,
                # HBM bytes come from rocprofv3 counter runs, which this run
                # does not make (the profile summaries hold them)
                "traffic": None,
                "flops_per_launch": flops, "launch_ms": grav_s * 1e3,
                "interactions_per_s": (stats["gravity_pp"] + stats["gravity_pn"]) / args.steps / grav_s,
                # node visits of the lanes over 64 x the waves' walk iterations
                "walk_lane_utilisation": (stats["gravity_opened"] + stats["gravity_pn"]) /
                                         max(64 * stats["gravity_wave_steps"], 1),
                # particle pairs over the lane slots the pair pass evaluates
                "pair_lane_utilisation": stats["gravity_pp"] / max(stats["gravity_pp_slots"], 1)}
    if grav_alone is not None:
        # the same two kernels (plus the combine) with nothing beside them:
        # GravityForce stage calls after the timed steps
        a_tf = grav_alone["flops_per_launch"] / (grav_alone["launch_ms"] * 1e-3) / 1e12
        roofline["alone"] = {"launch_ms": grav_alone["launch_ms"], "flops_per_launch": grav_alone["flops_per_launch"],
                             "achieved": a_tf, "frac": a_tf / peak}
    kernels = kernel_rooflines(n, stats, timers, args.steps, peak)

    adjoint = None
    if not args.no_adjoint:
        import numpy as np
        adj64, g64 = adjoint_run(args.adjoint_side, args.adjoint_steps, 1e-3, FP64)
        adjoint, gmx = adjoint_run(args.adjoint_side, args.adjoint_steps, 1e-3, MIXED)
        adjoint["fp64"] = adj64
        adjoint["grad_rel_err_vs_fp64"] = {
            name: float(np.linalg.norm(a - b) / np.linalg.norm(b)) for name, a, b in zip(("pos", "vel", "ene"), gmx, g64)}
        log(f"[bench] adjoint mixed vs FP64 gradient: {adjoint['grad_rel_err_vs_fp64']}")

    workloads = None
    if not args.no_workloads:
        workloads = {}
        try:
            pk, fk = ics.khi(precision=MIXED)
            workloads["khi_disph_2d"] = workload_rate("khi", pk, fk, 5, 1)
            pk64 = pk.copy()
            pk64.precision = FP64
            workloads["khi_disph_2d"]["mixed_vs_fp64"] = mixed_errors(pk64, pk, fk)
            del fk
        except Exception as e:  # a side measurement must not lose the main line
            workloads["khi_disph_2d"] = {"error": str(e)}
        try:
            # Only the first step from the initial condition is timed: with
            # u_bg = 1e-5 the reference scheme itself leaves non-finite states
            # after it (predict drives u of cold particles beside the blast
            # below zero, src/solver.cpp:449-452).  It is timed from fresh
            # uploads after one untimed warm-up repetition.
            ps, fs = ics.sedov(n_side=256, precision=MIXED)
            ps.neighbor_capacity = 256
            workloads["sedov_gsph_3d"] = first_step_rate("sedov", ps, fs, repeats=3)
            del fs
            ps2, fs2 = ics.sedov(n_side=128, precision=FP64)
            ps2.neighbor_capacity = 256
            pm2 = ps2.copy()
            pm2.precision = MIXED
            workloads["sedov_gsph_3d"]["mixed_vs_fp64_at_128^3"] = mixed_errors(ps2, pm2, fs2)
            del fs2
        except Exception as e:
            workloads["sedov_gsph_3d"] = {"error": str(e)}

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(p64, fields, n, args.cpu_steps)

    ms = elapsed / args.steps * 1e3
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if args.precision == "mixed" else "f64",
        "data": "synthetic",
        "config": {"workload": "evrard_collapse", "n_particles": n, "sph_type": "disph", "kernel": "wendland_c4",
                   "neighbor_number": 64, "gravity_theta": 0.5, "leaf_particle_num": 32,
                   "precision": args.precision, "balsara": True, "time_dependent_av": True},
        "fp64_value": fp64_rate,
        "mixed_vs_fp64": errors,
        "pairs_per_s": {"gather": stats["gather_pairs"] / args.steps / (ms * 1e-3),
                        "force": stats["force_pairs"] / args.steps / (ms * 1e-3)},
        "stage_ms_per_step": {k: v / args.steps for k, v in timers.items()},
        "counters": "timed steps run with the diagnostic counters off (the library default); interaction "
                    "counts come from the same steps rerun with them on",
        "roofline": roofline,
        "kernel_rooflines": kernels,
        "adjoint": adjoint,
        "other_workloads": workloads,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
