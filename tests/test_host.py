"""CPU checks of the C++ host layer: the JSON parameter reader follows the
reference's keys, defaults and errors (src/solver.cpp:155-299), parameter
files written from Python parse back to the same values, and snapshots
round-trip.  The reader is exercised through ``sphgpu_run --print-params``,
which touches no GPU."""
from __future__ import annotations

import json
import pathlib
import subprocess

import numpy as np
import pytest

from sphcode_amd import _abi, ics, snapshot

ROOT = pathlib.Path(__file__).resolve().parents[1]
RUN = ROOT / "sphcode_amd" / "bin" / "sphgpu_run"
HOST_LIB = ROOT / "sphcode_amd" / "lib" / "libsphhost.so"

REQUIRED = {"outputDirectory": "results/t", "endTime": 0.2, "gamma": 1.4}


def parse(tmp_path, doc) -> tuple[int, dict | None, str]:
    assert RUN.exists(), "run python -m sphcode_amd.build"
    f = tmp_path / "p.json"
    f.write_text(doc if isinstance(doc, str) else json.dumps(doc))
    r = subprocess.run([str(RUN), str(f), "--print-params"], capture_output=True, text=True, timeout=60)
    return r.returncode, (json.loads(r.stdout) if r.returncode == 0 else None), r.stderr


def test_host_library_built_and_linked():
    assert HOST_LIB.exists() and RUN.exists(), "run python -m sphcode_amd.build"
    out = subprocess.run(["readelf", "-d", str(HOST_LIB)], capture_output=True, text=True, check=True).stdout
    assert "libsphgpu.so" in out


def test_defaults(tmp_path):
    rc, p, err = parse(tmp_path, REQUIRED)
    assert rc == 0, err
    assert p["outputDirectory"] == "results/t"
    assert p["startTime"] == 0.0 and p["endTime"] == 0.2
    assert p["outputTime"] == pytest.approx(0.002) and p["energyTime"] == p["outputTime"]
    assert p["SPHType"] == "ssph" and p["kernel"] == "cubic_spline"
    assert (p["cflSound"], p["cflForce"]) == (0.3, 0.125)
    assert p["avAlpha"] == 1.0 and p["useBalsaraSwitch"] and not p["useTimeDependentAV"]
    assert (p["alphaMax"], p["alphaMin"], p["epsilonAV"]) == (2.0, 0.1, 0.2)
    assert not p["useArtificialConductivity"] and p["alphaAC"] == 1.0
    assert (p["maxTreeLevel"], p["leafParticleNumber"], p["neighborNumber"]) == (20, 1, 32)
    assert p["iterativeSmoothingLength"] and not p["periodic"] and not p["useGravity"]
    assert (p["G"], p["theta"]) == (1.0, 0.5) and p["use2ndOrderGSPH"]
    assert p["DIM"] == 3 and p["precision"] == "fp64" and p["neighborCapacity"] == 0


@pytest.mark.parametrize("key", ["outputDirectory", "endTime", "gamma"])
def test_required_keys(tmp_path, key):
    doc = dict(REQUIRED)
    del doc[key]
    rc, _, err = parse(tmp_path, doc)
    assert rc == 1 and f"No such node ({key})" in err


@pytest.mark.parametrize("doc,msg", [
    ({"startTime": 1.0}, "endTime < startTime"),
    ({"SPHType": "isph"}, "Unknown SPH type"),
    ({"kernel": "gaussian"}, "kernel is unknown."),
    ({"useTimeDependentAV": True, "alphaMax": 0.05}, "alphaMax < alphaMin"),
    ({"periodic": True, "DIM": 2, "rangeMax": [1, 1, 1], "rangeMin": [0, 0]}, "rangeMax != DIM"),
    # the reference reports a bad rangeMin with the rangeMax message
    ({"periodic": True, "DIM": 2, "rangeMax": [1, 1], "rangeMin": [0]}, "rangeMax != DIM"),
    ({"periodic": True, "DIM": 2}, "No such node (rangeMax)"),
    ({"neighborNumber": 1.5}, "conversion of data to type failed (neighborNumber)"),
    ({"useGravity": "yes"}, "conversion of data to type failed (useGravity)"),
    ({"DIM": 4}, "DIM must be 1, 2 or 3"),
    ({"precision": "fp16"}, "Unknown precision"),
    ({"DIM": 1, "kernel": "wendland"}, "Wendland C4 kernel is not defined for DIM = 1"),
])
def test_errors(tmp_path, doc, msg):
    rc, _, err = parse(tmp_path, {**REQUIRED, **doc})
    assert rc == 1 and msg in err, err


def test_malformed_json(tmp_path):
    rc, _, err = parse(tmp_path, '{"endTime": 0.2, "gamma": }')
    assert rc == 1 and "JSON parse error" in err


def test_conditional_keys(tmp_path):
    """Keys that the reference reads only under a switch keep their defaults
    when the switch is off (src/solver.cpp:221-230, 233-236, 286-290, 294-298)."""
    off = {"alphaMax": 9.0, "alphaMin": 8.0, "epsilonAV": 0.7, "alphaAC": 3.0, "G": 2.0, "theta": 0.1,
           "use2ndOrderGSPH": False}
    rc, p, err = parse(tmp_path, {**REQUIRED, **off})
    assert rc == 0, err
    assert (p["alphaMax"], p["alphaMin"], p["epsilonAV"], p["alphaAC"]) == (2.0, 0.1, 0.2, 1.0)
    assert (p["G"], p["theta"]) == (1.0, 0.5) and p["use2ndOrderGSPH"]
    on = {**off, "alphaMin": 0.5, "useTimeDependentAV": True, "useArtificialConductivity": True, "useGravity": 1,
          "SPHType": "gsph"}
    rc, p, err = parse(tmp_path, {**REQUIRED, **on})
    assert rc == 0, err
    assert (p["alphaMax"], p["alphaMin"], p["epsilonAV"], p["alphaAC"]) == (9.0, 0.5, 0.7, 3.0)
    assert (p["G"], p["theta"]) == (2.0, 0.1) and not p["use2ndOrderGSPH"]


def test_bool_and_number_forms(tmp_path):
    """Values are read as text: "1"/"0" are booleans, quoted numbers are numbers."""
    rc, p, err = parse(tmp_path, {**REQUIRED, "periodic": "1", "DIM": "2", "rangeMax": ["1.5", 2],
                                  "rangeMin": [-0.5, "0"], "useBalsaraSwitch": 0, "gamma": "1.6666666666666667"})
    assert rc == 0, err
    assert p["periodic"] and not p["useBalsaraSwitch"] and p["DIM"] == 2
    assert p["rangeMax"][:2] == [1.5, 2.0] and p["rangeMin"][:2] == [-0.5, 0.0]
    assert p["gamma"] == 1.6666666666666667


@pytest.mark.parametrize("make", [
    lambda: ics.sod(40, 5),
    lambda: ics.khi(n_low=8, n_high=11, jitter=0.05),
    lambda: ics.evrard(64, leaf=32),
    lambda: ics.sedov(n_side=4, n_hot=8, u_background=1e-2),
])
def test_params_to_json_round_trip(tmp_path, make):
    p, _ = make()
    p.time_end = 0.25
    p.precision = _abi.MIXED
    f = tmp_path / "rt.json"
    snapshot.write_params(f, p)
    r = subprocess.run([str(RUN), str(f), "--print-params"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    got = json.loads(r.stdout)
    want = snapshot.params_to_json(p)
    for k, v in want.items():
        if k in ("rangeMax", "rangeMin"):
            assert got[k][:p.dim] == v, k
        else:
            assert got[k] == v, k


def test_snapshot_round_trip(tmp_path):
    rng = np.random.default_rng(3)
    n, dim = 37, 2
    fields = {k: rng.normal(size=(n, dim)) for k in snapshot.VEC_FIELDS}
    fields.update({k: rng.normal(size=n) for k in snapshot.SCALAR_FIELDS})
    fields["id"] = rng.permutation(n).astype(np.int32)
    fields["neighbor"] = rng.integers(0, 100, n).astype(np.int32)
    f = tmp_path / "s.snap"
    snapshot.write(f, fields, dim, time=0.125, dt=1e-3)
    got, d, t, dt = snapshot.read(f)
    assert (d, t, dt) == (dim, 0.125, 1e-3)
    for k, v in fields.items():
        np.testing.assert_array_equal(got[k], v)
    assert f.stat().st_size == 48 + n * (8 * (4 * dim + 12) + 8)


MIRROR_SRC = ROOT / "tests" / "host_solver_mirror.cpp"


def build_mirror(out_dir: pathlib.Path) -> pathlib.Path:
    """Compile and link tests/host_solver_mirror.cpp (a solver written in the
    shape of the reference's Solver, src/solver.cpp:353-474) against
    libsphhost and libsphgpu."""
    exe = out_dir / "host_solver_mirror"
    lib = HOST_LIB.parent
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-Werror", "-Wno-unused-parameter",
           "-I", str(ROOT / "include"), "-I", str(ROOT / "sphcode_amd" / "host" / "include"),
           str(MIRROR_SRC), "-o", str(exe), "-L", str(lib), "-lsphhost", "-lsphgpu", f"-Wl,-rpath,{lib}"]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    return exe


def test_reference_shaped_solver_compiles_and_links(tmp_path):
    """A driver using the reference's Simulation surface -- get_particles()
    with vec_t positions, set_particle_num, add_vector_array, get_tree()
    with resize and make, get_periodic()->apply, get_dt, make_tree -- and
    the Module interface compiles and links against the host library."""
    exe = build_mirror(tmp_path)
    out = subprocess.run(["nm", "-C", "--undefined-only", str(exe)], capture_output=True, text=True, check=True).stdout
    for sym in ("sph::Simulation::get_particles()", "sph::BHTree::make(", "sph::Simulation::add_vector_array(",
                "sph::Simulation::make_tree()"):
        assert sym in out, sym
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "usage" in r.stderr


def test_host_kernels_match_oracle():
    """The host KernelFunction implementations (for code working on
    get_particles()) equal the oracle's kernels."""
    import sys
    sys.path.insert(0, str(ROOT / "tests"))
    from _oracle import kernel_eval
    src = r'''
#include <cstdio>
#include "sph/kernel_function.hpp"
#include "sph/periodic.hpp"
int main() {
    for (int dim = 1; dim <= 3; ++dim)
        for (int kern = 0; kern < 2; ++kern) {
            if (kern == 1 && dim == 1) continue;
            sph::SPHParameters p;
            p.dim = dim;
            p.kernel = kern ? sph::KernelType::WENDLAND : sph::KernelType::CUBIC_SPLINE;
            auto k = sph::make_kernel(p);
            for (double r : {0.0, 0.1, 0.37, 0.5, 0.81, 1.2}) {
                const double h = 1.3;
                sph::vec_t rij(r, 0.0, 0.0);
                std::printf("%d %d %.17g %.17g %.17g %.17g\n", dim, kern, r, k->w(r, h), k->dw(rij, r, h)[0], k->dhw(r, h));
            }
        }
    return 0;
}
'''
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        f = pathlib.Path(d) / "k.cpp"
        f.write_text(src)
        exe = pathlib.Path(d) / "k"
        subprocess.run(["g++", "-std=c++17", "-O2", "-I", str(ROOT / "include"), "-I",
                        str(ROOT / "sphcode_amd" / "host" / "include"), str(f), "-o", str(exe)], check=True)
        lines = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    for ln in filter(None, lines):
        dim, kern, r, w, dwx, dh = ln.split()
        dim, kern, r, w, dwx, dh = int(dim), int(kern), float(r), float(w), float(dwx), float(dh)
        ow, oc, odh = kernel_eval(dim, kern, r, 1.3)
        assert w == pytest.approx(ow, rel=1e-13, abs=1e-15)
        assert dwx == pytest.approx(oc * r, rel=1e-13, abs=1e-15)
        assert dh == pytest.approx(odh, rel=1e-12, abs=1e-14)
