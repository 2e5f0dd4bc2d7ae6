"""The C++ Module layer on the GPU: ``sphgpu_run`` (Solver::initialize, then
TimeStep / predict / make_tree / PreInteraction / FluidForce / GravityForce /
correct as separate Module calls, src/solver.cpp:392-429) must give exactly
the state of the fused C-ABI path (``sphgpu_initialize`` + ``sphgpu_integrate``)
on the same input: both run the same kernels in the same order, so the bar
is bit equality of every field, the time and the number of steps."""
from __future__ import annotations

import pathlib
import subprocess

import numpy as np
import pytest

from _cases import CASES
from sphcode_amd import snapshot

pytestmark = pytest.mark.gpu

ROOT = pathlib.Path(__file__).resolve().parents[1]
RUN = ROOT / "sphcode_amd" / "bin" / "sphgpu_run"

# fields the reference leaves undefined for a scheme are still computed
# identically by both paths, so every field is compared
COMPARED = snapshot.VEC_FIELDS + snapshot.SCALAR_FIELDS + snapshot.INT_FIELDS


def _run_cli(tmp_path, p, f, steps):
    pj, sin, sout = tmp_path / "p.json", tmp_path / "in.snap", tmp_path / "out.snap"
    snapshot.write_params(pj, p)
    snapshot.write(sin, f, int(p.dim))
    cmd = [str(RUN), str(pj), "--particles", str(sin), "--output", str(sout)]
    if steps is not None:
        cmd += ["--steps", str(steps)]
    print("sphgpu_run", " ".join(cmd[1:]), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    nsteps = int(r.stdout.split("steps=")[1].split()[0])
    return snapshot.read(sout), nsteps


def _run_engine(p, f, steps=None):
    from sphcode_amd import GPUEngine
    g = GPUEngine(p)
    g.upload(f)
    g.initialize()
    n = 0
    if steps is None:
        while g.time()[0] < p.time_end:
            g.integrate()
            n += 1
    else:
        for _ in range(steps):
            g.integrate()
            n += 1
    out = g.download()
    t, dt, _ = g.time()
    g.close()
    return out, t, dt, n


@pytest.mark.parametrize("case", ["sod_ssph", "sod_gsph2", "khi_disph", "evrard_disph", "sedov_gsph"])
def test_module_sequence_matches_fused_step(gpu, tmp_path, case):
    p, f = CASES[case]()
    p.time_end = 1.0
    (got, dim, t_cli, dt_cli), n_cli = _run_cli(tmp_path, p, f, steps=3)
    want, t, dt, n = _run_engine(p, f, steps=3)
    assert dim == p.dim and n_cli == n == 3
    assert t_cli == t and dt_cli == dt
    for k in COMPARED:
        a = np.asarray(got[k]).reshape(np.asarray(want[k]).shape)
        np.testing.assert_array_equal(a, want[k], err_msg=f"{case}: {k}")


def test_run_to_end_time(gpu, tmp_path):
    """Without --steps the driver loops while time < endTime
    (src/solver.cpp:318-325): same step count and final state."""
    p, f = CASES["sod_ssph"]()
    p.time_end = 2e-3
    (got, _, t_cli, _), n_cli = _run_cli(tmp_path, p, f, steps=None)
    want, t, _, n = _run_engine(p, f)
    assert n_cli == n and n > 1 and t_cli == t and t >= p.time_end
    for k in ("pos", "vel", "ene", "sml", "dens"):
        np.testing.assert_array_equal(got[k].reshape(want[k].shape), want[k], err_msg=k)


def test_cli_reports_stage_errors(gpu, tmp_path):
    """An error status from a stage ends the driver with status 1 and the
    C ABI's message, as SPHException does in the reference."""
    p, f = CASES["evrard_disph"]()
    p.time_end = 1.0
    p.neighbor_capacity = 8     # far below the ~90 candidates per particle
    pj, sin = tmp_path / "p.json", tmp_path / "in.snap"
    snapshot.write_params(pj, p)
    snapshot.write(sin, f, int(p.dim))
    r = subprocess.run([str(RUN), str(pj), "--particles", str(sin), "--steps", "1"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 1 and "error:" in r.stderr and "capacity" in r.stderr.lower(), r.stderr


def _cli(tmp_path, pj, sin, sout, steps, restart):
    cmd = [str(RUN), str(pj), "--particles", str(sin), "--output", str(sout), "--steps", str(steps)]
    if restart:
        cmd.append("--restart")
    print("sphgpu_run", " ".join(cmd[1:]), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return snapshot.read(sout, with_hvs=True)


@pytest.mark.parametrize("case", ["sod_ssph", "sod_gsph2", "khi_disph", "evrard_disph"])
def test_restart_continues_run(gpu, tmp_path, case):
    """A run of 3 steps written with --output and continued with --restart
    for 3 more reaches the state of 6 uninterrupted steps: the snapshot keeps
    every FP64 bit plus time, dt and min h/v_sig, and the restart skips the
    initial smoothing.  The continued run sorts the particles from caller
    order rather than from the last step's order, so particles that share a
    Morton key may be summed in another order: the bar is round-off."""
    p, f = CASES[case]()
    p.time_end = 1.0
    pj, sin = tmp_path / "p.json", tmp_path / "in.snap"
    snapshot.write_params(pj, p)
    snapshot.write(sin, f, int(p.dim))
    want, _, t6, dt6, hv6 = _cli(tmp_path, pj, sin, tmp_path / "six.snap", 6, False)
    mid, _, t3, _, hv3 = _cli(tmp_path, pj, sin, tmp_path / "three.snap", 3, False)
    assert hv3 > 0 and 0 < t3 < t6
    got, _, t, dt, hv = _cli(tmp_path, pj, tmp_path / "three.snap", tmp_path / "cont.snap", 3, True)
    np.testing.assert_allclose(t, t6, rtol=1e-12)
    np.testing.assert_allclose(dt, dt6, rtol=1e-10)
    np.testing.assert_allclose(hv, hv6, rtol=1e-10)
    np.testing.assert_array_equal(got["id"], want["id"])
    for k in ("pos", "vel", "ene", "sml", "dens", "pres", "alpha", "acc", "dene"):
        scale = np.abs(want[k]).max() + 1e-300
        err = np.abs(got[k] - want[k]).max() / scale
        assert err < 1e-10, f"{case}: {k} max rel err {err:.3e}"


@pytest.mark.parametrize("case", ["sod_gsph2", "khi_disph", "evrard_disph"])
def test_cli_restart_matches_oracle(gpu, tmp_path, case):
    """The driver's path end to end against the oracle: sphgpu_run for 3
    steps with --output, then 3 more with --restart from that snapshot
    (src/solver.cpp:318-345 loop, src/output.cpp snapshot), against the
    oracle's initialize + 6 uninterrupted steps, with the bars the fused
    path's own 3-step oracle comparison uses."""
    from _oracle import OracleEngine
    p, f = CASES[case]()
    p.time_end = 1.0
    pj, sin = tmp_path / "p.json", tmp_path / "in.snap"
    snapshot.write_params(pj, p)
    snapshot.write(sin, f, int(p.dim))
    _cli(tmp_path, pj, sin, tmp_path / "three.snap", 3, False)
    got, _, t, dt, _ = _cli(tmp_path, pj, tmp_path / "three.snap", tmp_path / "cont.snap", 3, True)
    ref = OracleEngine(p)
    ref.upload(f)
    ref.initialize()
    for _ in range(6):
        ref.integrate()
    t_ref, dt_ref = ref.time()[:2]
    want = ref.download()
    ref.close()
    assert abs(t - t_ref) <= 1e-9 * t_ref and abs(dt - dt_ref) <= 1e-9 * dt_ref, f"{case}: t/dt {t} {dt} vs {t_ref} {dt_ref}"
    # both keep the caller's particle order; align by id regardless
    g, w = np.argsort(got["id"], kind="stable"), np.argsort(want["id"], kind="stable")
    np.testing.assert_array_equal(got["id"][g], want["id"][w])
    n, d = len(want["id"]), int(p.dim)
    for k in ("sml", "dens", "pres"):
        a, b = np.asarray(got[k]).reshape(n)[g], want[k].reshape(n)[w]
        err = np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-300))
        assert err <= 1e-6, f"{case}: {k} rel err {err:.3e}"
    for k, bar in (("pos", 1e-6), ("vel", 1e-6), ("acc", 1e-5)):
        a, b = np.asarray(got[k]).reshape(n, d)[g], want[k].reshape(n, d)[w]
        err = np.max(np.abs(a - b)) / (np.sqrt(np.mean(b * b)) + 1e-300)
        assert err <= bar, f"{case}: {k} err {err:.3e}"


def test_restart_rejects_fresh_snapshot(gpu, tmp_path):
    """--restart needs the min h/v_sig of a finished step; an initial
    condition written without it is refused."""
    p, f = CASES["sod_ssph"]()
    pj, sin = tmp_path / "p.json", tmp_path / "in.snap"
    snapshot.write_params(pj, p)
    snapshot.write(sin, f, int(p.dim))
    r = subprocess.run([str(RUN), str(pj), "--particles", str(sin), "--steps", "1", "--restart"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 1 and "--restart" in r.stderr, r.stderr


@pytest.mark.parametrize("case", ["sod_gsph2", "evrard_disph", "khi_disph"])
def test_reference_shaped_solver_runs(gpu, tmp_path, case):
    """tests/host_solver_mirror.cpp drives the GPU modules the way the
    reference's Solver does (src/solver.cpp:353-474): predict and correct
    as host loops over Simulation::get_particles(), the tree through
    get_tree()->resize / make and make_tree.  Each host change is uploaded
    before the next stage (continuing the run, as a restart does), so three
    such steps reach sphgpu_run's three fused steps to round-off (particles
    sharing a Morton key may be summed in another order after an upload)."""
    from test_host import build_mirror
    exe = build_mirror(tmp_path)
    p, f = CASES[case]()
    p.time_end = 1.0
    pj, sin = tmp_path / "p.json", tmp_path / "in.snap"
    snapshot.write_params(pj, p)
    snapshot.write(sin, f, int(p.dim))
    r = subprocess.run([str(exe), str(pj), str(sin), "3", str(tmp_path / "mirror.snap")], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    got, _, t, dt, hv = snapshot.read(tmp_path / "mirror.snap", with_hvs=True)
    want, _, t3, dt3, hv3 = _cli(tmp_path, pj, sin, tmp_path / "three.snap", 3, False)
    np.testing.assert_allclose(t, t3, rtol=1e-12)
    np.testing.assert_allclose(dt, dt3, rtol=1e-10)
    np.testing.assert_allclose(hv, hv3, rtol=1e-10)
    for k in ("pos", "vel", "ene", "sml", "dens", "pres", "alpha", "acc", "dene"):
        scale = np.abs(want[k]).max() + 1e-300
        err = np.abs(got[k] - want[k]).max() / scale
        print(f"[mirror] {case} {k}: max rel err {err:.2e}", flush=True)
        assert err < 1e-10, f"{case}: {k} max rel err {err:.3e}"
