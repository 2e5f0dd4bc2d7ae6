"""CPU checks of the kernel derivatives used by the adjoint
(sphcode_amd/csrc/kderiv.hpp, compiled here with g++ as host code).

  * W, grad-W coefficient and dW/dh equal the oracle's kernels (which
    follow include/kernel/cubic_spline.hpp:21-52 and
    include/kernel/wendland_kernel.hpp:23-49) to round-off;
  * every higher derivative (d2W/dh2, d2W/dhdr, df/dr, df/dh) equals a
    central difference of the lower one, in both regions of the cubic
    spline and across the support;
  * the support edge: all values vanish at and beyond the support.
"""
from __future__ import annotations

import pathlib
import subprocess

import numpy as np
import pytest

from _oracle import kernel_eval

ROOT = pathlib.Path(__file__).resolve().parents[1]

DRIVER = r"""
#include <cstdio>
#include <cstdlib>
#include "kderiv.hpp"
using namespace sphgpu;
template <int D, int K> void run(double r, double h) {
    const auto o = kernel_derivs<D, K, double>(r, h);
    std::printf("%.17g %.17g %.17g %.17g %.17g %.17g %.17g\n", o.w, o.wh, o.whh, o.whr, o.f, o.fr, o.fh);
}
int main(int argc, char **argv) {
    const int D = std::atoi(argv[1]), K = std::atoi(argv[2]);
    for (int a = 3; a + 1 < argc; a += 2) {
        const double r = std::atof(argv[a]), h = std::atof(argv[a + 1]);
        if (D == 1 && K == 0) run<1, 0>(r, h);
        if (D == 2 && K == 0) run<2, 0>(r, h);
        if (D == 3 && K == 0) run<3, 0>(r, h);
        if (D == 2 && K == 1) run<2, 1>(r, h);
        if (D == 3 && K == 1) run<3, 1>(r, h);
    }
}
"""

COMBOS = [(1, 0), (2, 0), (3, 0), (2, 1), (3, 1)]


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    d = tmp_path_factory.mktemp("kd")
    src = d / "kd.cpp"
    src.write_text(DRIVER)
    exe = d / "kd"
    subprocess.run(["g++", "-std=c++17", "-O1", "-ffp-contract=off", "-I", str(ROOT / "sphcode_amd" / "csrc"),
                    str(src), "-o", str(exe)], check=True)
    return exe


def evaluate(driver, dim, kernel, pts):
    args = [str(driver), str(dim), str(kernel)]
    for r, h in pts:
        args += [repr(float(r)), repr(float(h))]
    out = subprocess.run(args, capture_output=True, text=True, check=True).stdout.split("\n")
    rows = [list(map(float, l.split())) for l in out if l.strip()]
    return np.array(rows)


def _points(kernel, h=0.7):
    support = h
    qs = np.concatenate([np.linspace(0.0, 0.98, 25), [0.3, 0.499, 0.501, 0.77, 0.999]])
    return [(q * support, h) for q in qs]


@pytest.mark.parametrize("dim,kernel", COMBOS)
def test_matches_oracle_kernels(driver, dim, kernel):
    pts = _points(kernel)
    got = evaluate(driver, dim, kernel, pts)
    for (r, h), row in zip(pts, got):
        w, c, dh = kernel_eval(dim, kernel, r, h)
        scale = max(abs(w), 1e-300)
        assert abs(row[0] - w) <= 1e-13 * scale + 1e-15, (r, h)
        if r > 0:
            assert abs(row[4] - c) <= 1e-12 * max(abs(c), 1.0), (r, h)
        assert abs(row[1] - dh) <= 1e-12 * max(abs(dh), 1.0), (r, h)


@pytest.mark.parametrize("dim,kernel", COMBOS)
def test_higher_derivatives_by_central_differences(driver, dim, kernel):
    pts = [(r, h) for r, h in _points(kernel) if r > 0.02]
    eps = 1e-6
    base = evaluate(driver, dim, kernel, pts)
    rp = evaluate(driver, dim, kernel, [(r + eps, h) for r, h in pts])
    rm = evaluate(driver, dim, kernel, [(r - eps, h) for r, h in pts])
    hp = evaluate(driver, dim, kernel, [(r, h + eps) for r, h in pts])
    hm = evaluate(driver, dim, kernel, [(r, h - eps) for r, h in pts])
    W, WH, WHH, WHR, F, FR, FH = range(7)
    d_r = (rp - rm) / (2 * eps)
    d_h = (hp - hm) / (2 * eps)
    r = np.array([p[0] for p in pts])
    scale = np.abs(base).max(axis=0) + 1e-30
    tol = 2e-6
    # dW/dr = r f
    np.testing.assert_allclose(d_r[:, W], r * base[:, F], atol=tol * scale[W])
    np.testing.assert_allclose(d_h[:, W], base[:, WH], atol=tol * scale[WH])
    np.testing.assert_allclose(d_h[:, WH], base[:, WHH], atol=tol * scale[WHH])
    np.testing.assert_allclose(d_r[:, WH], base[:, WHR], atol=tol * scale[WHR])
    np.testing.assert_allclose(d_r[:, F], base[:, FR], atol=tol * scale[FR])
    np.testing.assert_allclose(d_h[:, F], base[:, FH], atol=tol * scale[FH])


@pytest.mark.parametrize("dim,kernel", COMBOS)
def test_vanishes_at_support(driver, dim, kernel):
    h = 0.7
    got = evaluate(driver, dim, kernel, [(h, h), (1.3 * h, h)])
    assert np.all(np.abs(got) < 1e-13)
