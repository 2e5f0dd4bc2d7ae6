"""Mutation check of the oracle's pinning tests (run by hand, not by pytest).

Each entry of MUT is one plausible slip in oracle/sphref.cpp: a wrong
coefficient, weight, evaluation point or dropped term.  For each, this
script copies the repository to a scratch directory, applies the change,
rebuilds the oracle there and runs the oracle's CPU tests; every mutation
must make at least one test fail ("CAUGHT").

    python tests/mutation_check.py all
    python tests/mutation_check.py av_coef,hll_noroe tests/test_oracle_pins.py
"""
import os
import pathlib
import shutil
import subprocess
import sys
import tempfile

ROOT = str(pathlib.Path(__file__).resolve().parents[1])
MUT = {
 "av_coef": ("return -0.5 * bal * alpha * vsig * w * rho_inv;", "return -1.0 * bal * alpha * vsig * w * rho_inv;"),
 "av_vsig": ("const double vsig = a.sound + b.sound - 3.0 * w;", "const double vsig = a.sound + b.sound - 2.0 * w;"),
 "disph_balsara_w": ("const double wgt = type == SPHGPU_DISPH ? pj.mass * pj.ene : pj.mass;\n                divv -= wgt * dot(vij, dwij);\n                rot.add",
                     "const double wgt = pj.mass;\n                divv -= wgt * dot(vij, dwij);\n                rot.add"),
 "open_centre": ("const Vec d = s.box.rij(pi.pos, nd.com);", "const Vec d = s.box.rij(pi.pos, nd.center);"),
 "hll_noroe": ("const double s_l = std::min(L[0] - L[3], ut - ct);\n    const double s_r = std::max(R[0] + R[3], ut + ct);",
               "const double s_l = L[0] - L[3] + 0 * ut * ct;\n    const double s_r = R[0] + R[3];"),
 "muscl_delta": ("const double di = 0.5 * (1.0 - pi.sound * dt * rinv);\n                    const double dj = 0.5 * (1.0 - pj.sound * dt * rinv);",
                 "const double di = 0.5 + 0 * dt;\n                    const double dj = 0.5;"),
 "tdav_source": ("std::max(-divv, 0.0) * (amax - pi.alpha)) * dt;", "std::max(-divv, 0.0) * (amax)) * dt;"),
 "pre_vsig": ("const double vsig = pi.sound + pj.sound - 3.0 * dot(rij, pi.vel - pj.vel) / r;",
              "const double vsig = pi.sound + pj.sound - 2.0 * dot(rij, pi.vel - pj.vel) / r;"),
 "soft_hi": ("pi.phi -= G * q.mass * (soft_pot(r, pi.sml) + soft_pot(r, q.sml)) * 0.5;\n                pi.acc -= rij * (G * q.mass * (soft_acc(r, pi.sml) + soft_acc(r, q.sml)) * 0.5);",
             "pi.phi -= G * q.mass * (soft_pot(r, pi.sml) + soft_pot(r, pi.sml)) * 0.5;\n                pi.acc -= rij * (G * q.mass * (soft_acc(r, pi.sml) + soft_acc(r, pi.sml)) * 0.5);"),
 "gsph_gradp": ("s.grad_p[i] = (gd * pi.ene + gu) * (gamma - 1.0);", "s.grad_p[i] = (gd * pi.ene) * (gamma - 1.0);"),
 "ac_vsig": ("std::sqrt(2.0 * std::fabs(a.pres - b.pres) / (a.dens + b.dens))", "std::sqrt(1.0 * std::fabs(a.pres - b.pres) / (a.dens + b.dens))"),
 "ctl_omega": ("if(type == SPHGPU_SSPH) pi.gradh = 1.0 / (1.0 + hi / (D * rho) * drho);", "if(type == SPHGPU_SSPH) pi.gradh = 1.0;"),
 "ctl_fij": ("const double fij = 1.0 - pi.gradh / (pj.mass * pj.ene);", "const double fij = 1.0;"),
 # further slips
 "tdav_implicit": ("pi.alpha = (pi.alpha + dt * tau_inv * amin + si * dt * amax) / (1.0 + dt * tau_inv + si * dt);",
                   "pi.alpha = (pi.alpha + dt * tau_inv * amin + si * dt * amax) / (1.0 + dt * tau_inv);"),
 "ac_grav": ("std::fabs(dot(a.vel - b.vel, rij) / norm(rij))", "std::fabs(dot(a.vel - b.vel, rij))"),
 "balsara_eps": ("rot.magnitude() + 1e-4 * pi.sound / hi)", "rot.magnitude() + 1e-3 * pi.sound / hi)"),
 "av_rho": ("const double rho_inv = 2.0 / (a.dens + b.dens);", "const double rho_inv = 1.0 / (a.dens + b.dens);"),
 "av_alpha_avg": ("const double alpha = 0.5 * (a.alpha + b.alpha);", "const double alpha = a.alpha;"),
 "av_balsara_avg": ("const double bal = 0.5 * (a.balsara + b.balsara);", "const double bal = a.balsara;"),
 "disph_pinv": ("const double p_inv = (gamma - 1.0) / pi.pres;\n                divv *= p_inv;", "const double p_inv = 1.0 / pi.pres;\n                divv *= p_inv;"),
 "gsph_gradv": ("for(int c = 0; c < D; ++c) s.grad_v[c][i] = gv[c] * rinv;", "for(int c = 0; c < D; ++c) s.grad_v[c][i] = gv[c];"),
 "muscl_sign": ("L[1] = pj.dens + van_leer(dd, dd_j) * dj;", "L[1] = pj.dens - van_leer(dd, dd_j) * dj;"),
}


def run(name, tests):
    old, new = MUT[name]
    d = os.path.join(tempfile.gettempdir(), "sph_mutation", name)
    if os.path.exists(d):
        shutil.rmtree(d)
    shutil.copytree(ROOT, d, ignore=shutil.ignore_patterns(".git", "_work", "gpurun_out", "__pycache__", "lib"))
    os.makedirs(f"{d}/sphcode_amd/lib", exist_ok=True)
    for f in os.listdir(f"{ROOT}/sphcode_amd/lib"):
        shutil.copy(f"{ROOT}/sphcode_amd/lib/{f}", f"{d}/sphcode_amd/lib/{f}")
    src = open(f"{d}/oracle/sphref.cpp").read()
    assert src.count(old) == 1, (name, src.count(old))
    open(f"{d}/oracle/sphref.cpp", "w").write(src.replace(old, new))
    subprocess.run(["make", "-s", "-C", f"{d}/oracle", "-j8"], check=True, capture_output=True)
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", *tests], cwd=d,
                       capture_output=True, text=True)
    last = r.stdout.strip().splitlines()[-1] if r.stdout.strip() else r.stderr[-300:]
    print(f"{name:18s} {'CAUGHT' if r.returncode else 'SURVIVED'}  {last}", flush=True)
    shutil.rmtree(d)


if __name__ == "__main__":
    tests = sys.argv[2:] or ["tests/test_oracle_physics.py", "tests/test_oracle_tree.py", "tests/test_oracle_kernels.py", "tests/test_oracle_pins.py"]
    names = list(MUT) if sys.argv[1] == "all" else sys.argv[1].split(",")
    for n in names:
        run(n, tests)
