"""Whole runs of the small workloads on the GPU against the oracle.

  * The 1D Sod tube at its stated size (1,600 + 200 particles, SSPH, cubic
    spline, 4 neighbours, gamma = 1.4) integrated to endTime = 0.2 with the
    CFL time step: the same number of steps and final time as the oracle,
    density within 1e-6 of the oracle's RMS, and the density against the
    exact Riemann solution held to the bar the oracle itself meets
    (tests/test_oracle_physics.py).
  * sphgpu_run (the C++ Module layer and CLI) for 6 steps, and for 3 steps
    continued by a restart for 3 more, against the oracle's 6 uninterrupted
    steps, with the three-step bars of tests/test_gpu_parity.py.
  * energy.dat of sphgpu_run at the energyTime cadence of Solver::run
    (src/solver.cpp:340-345, src/output.cpp:66-90) against the oracle's sums
    at the same steps.
"""
from __future__ import annotations

import pathlib
import subprocess

import numpy as np
import pytest

from _cases import CASES
from _oracle import OracleEngine
from sphcode_amd import ics, snapshot
from test_oracle_physics import exact_sod

pytestmark = pytest.mark.gpu

ROOT = pathlib.Path(__file__).resolve().parents[1]
RUN = ROOT / "sphcode_amd" / "bin" / "sphgpu_run"


def _rms_err(a, b):
    return float(np.max(np.abs(a - b)) / (np.sqrt(np.mean(b * b)) + 1e-300))


@pytest.mark.parametrize("sph_type", [0, 1, 2])
def test_sod_tube_to_end_time(gpu, sph_type):
    from sphcode_amd import GPUEngine
    p, f = ics.sod(1600, 200)
    p.sph_type = sph_type
    assert len(f["mass"]) == 1800 and p.time_end == 0.2
    ref, g = OracleEngine(p), GPUEngine(p)
    steps = []
    for e in (ref, g):
        e.upload(f)
        e.initialize()
        n = 0
        while e.time()[0] < p.time_end:
            e.integrate()
            n += 1
        steps.append(n)
    ta, tb = ref.time(), g.time()
    a, b = ref.download(("pos", "dens", "vel", "ene")), g.download(("pos", "dens", "vel", "ene"))
    err = {k: _rms_err(b[k], a[k]) for k in ("pos", "dens", "vel", "ene")}
    x = b["pos"][:, 0]
    sel = (x > -0.22) & (x < 0.27)
    ex, _, _ = exact_sod(x[sel], tb[0])
    l1 = float(np.mean(np.abs(b["dens"][sel] - ex)))
    print(f"[sod] type={sph_type} steps oracle {steps[0]} gpu {steps[1]}, t {ta[0]:.10f} / {tb[0]:.10f}, "
          f"errors {err}, L1 vs exact {l1:.4e}", flush=True)
    assert steps[0] == steps[1]
    assert abs(tb[0] - ta[0]) <= 1e-9 * ta[0]
    for k, v in err.items():
        assert v <= 1e-6, f"{k}: {v:.3e}"
    assert l1 < 0.02
    g.close()


def _cli(tmp_path, pj, sin, sout, steps, restart=False, extra=()):
    cmd = [str(RUN), str(pj), "--particles", str(sin), "--output", str(sout), "--steps", str(steps), *extra]
    if restart:
        cmd.append("--restart")
    print("sphgpu_run", " ".join(cmd[1:]), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return snapshot.read(sout, with_hvs=True)


@pytest.mark.parametrize("case", ["sod_gsph2", "evrard_disph", "khi_disph"])
def test_cli_and_restart_against_oracle(gpu, tmp_path, case):
    p, f = CASES[case]()
    p.time_end = 1.0
    pj, sin = tmp_path / "p.json", tmp_path / "in.snap"
    snapshot.write_params(pj, p)
    snapshot.write(sin, f, int(p.dim))
    six, _, t6, dt6, _ = _cli(tmp_path, pj, sin, tmp_path / "six.snap", 6)
    _cli(tmp_path, pj, sin, tmp_path / "three.snap", 3)
    cont, _, tc, dtc, _ = _cli(tmp_path, pj, tmp_path / "three.snap", tmp_path / "cont.snap", 3, restart=True)
    ref = OracleEngine(p)
    ref.upload(f)
    ref.initialize()
    for _ in range(6):
        ref.integrate()
    t, dt, _ = ref.time()
    a = ref.download()
    for label, got, tg, dtg in (("cli", six, t6, dt6), ("restart", cont, tc, dtc)):
        assert abs(tg - t) <= 1e-9 * t and abs(dtg - dt) <= 1e-9 * dt, f"{case} {label}: t {tg} dt {dtg}"
        for k in ("sml", "dens", "pres"):
            v = got[k].reshape(a[k].shape)
            err = float(np.max(np.abs(v - a[k]) / np.abs(a[k])))
            assert err <= 1e-6, f"{case} {label}: {k} rel err {err:.3e}"
        for k in ("pos", "vel"):
            v = got[k].reshape(a[k].shape)
            assert _rms_err(v, a[k]) <= 1e-6, f"{case} {label}: {k}"
        print(f"[cli] {case} {label}: t={tg:.10g} matches the oracle's 6 steps", flush=True)


def test_energy_file_cadence(gpu, tmp_path):
    """energy.dat: the initial sums, then a line after every step that ends
    past the next energyTime mark (src/solver.cpp:318-345), with the
    kinetic, thermal, potential and total energy and the total momentum."""
    p, f = CASES["evrard_disph"]()
    p.time_end = 0.05
    p.time_energy = 0.01
    pj, sin = tmp_path / "p.json", tmp_path / "in.snap"
    snapshot.write_params(pj, p)
    snapshot.write(sin, f, int(p.dim))
    out_dir = tmp_path / "out"
    cmd = [str(RUN), str(pj), "--particles", str(sin), "--output-dir", str(out_dir)]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = (out_dir / "energy.dat").read_text().splitlines()
    assert lines[0].startswith("# time kinetic thermal potential total")
    rows = np.array([[float(v) for v in ln.split()] for ln in lines[1:]])
    ref = OracleEngine(p)
    ref.upload(f)
    ref.initialize()
    want = [np.concatenate([[ref.time()[0]], ref.energy()])]
    t_ene = p.time_energy
    while ref.time()[0] < p.time_end:
        ref.integrate()
        t = ref.time()[0]
        if t > t_ene:
            want.append(np.concatenate([[t], ref.energy()]))
            t_ene += p.time_energy
    want = np.array(want)
    print(f"[energy.dat] {len(rows)} lines, oracle {len(want)}", flush=True)
    assert rows.shape == want.shape
    scale = np.abs(want[:, 1:5]).max()
    assert np.allclose(rows[:, 0], want[:, 0], rtol=1e-9, atol=0)
    assert np.all(np.abs(rows[:, 1:5] - want[:, 1:5]) <= 1e-8 * scale)
    assert np.all(np.abs(rows[:, 5:] - want[:, 5:]) <= 1e-8 * scale)
