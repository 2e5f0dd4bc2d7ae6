"""Pins the oracle's SPH kernels and gravitational softening.

Oracle functions pinned here (oracle/sphref.cpp):
  Cubic::w/dw/dhw, WendlandC4::w/dw/dhw   (include/kernel/*.hpp)
    - normalisation: the integral of W over its support is 1 in D = 1, 2, 3
      (D = 1 Wendland: sigma = 0, the reference's own value);
    - compact support: W, grad W, dW/dh vanish for r >= h;
    - grad W and dW/dh against central differences of W in r and h -- the
      check of test/kernel_test/kernel_test.cpp:20-45, 55-81, with a
      threshold instead of a printout;
    - W(0, h) against the closed forms sigma/H^D (cubic, H = h/2) and
      sigma/h^D (Wendland).
  soft_pot/soft_acc (src/bhtree.cpp:273-299)
    - g(r) r = -df/dr (force is the gradient of the potential);
    - g(r) = M(<r)/r^3 with M the mass fraction of a 3D cubic-spline cloud
      of support h: the Hernquist-Katz kernels are the field of the cubic
      spline density, which the kernel checks above pin independently;
    - f = 1/r and g = 1/r^3 for r >= h, f(0) = 1.4/e.
"""
from __future__ import annotations

import math

import numpy as np
import pytest
from scipy import integrate

from _oracle import kernel_eval, softening

CUBIC, WENDLAND = 0, 1
SHELL = {1: lambda r: 2.0, 2: lambda r: 2 * math.pi * r, 3: lambda r: 4 * math.pi * r * r}


@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("kernel", [CUBIC, WENDLAND])
@pytest.mark.parametrize("h", [0.37, 1.0, 2.5])
def test_normalisation(dim, kernel, h):
    val, _ = integrate.quad(lambda r: kernel_eval(dim, kernel, r, h)[0] * SHELL[dim](r), 0.0, h,
                            limit=200, epsabs=1e-13, epsrel=1e-12, points=[h / 2])
    expect = 0.0 if (kernel == WENDLAND and dim == 1) else 1.0
    assert abs(val - expect) < 1e-10, (dim, kernel, h, val)


@pytest.mark.parametrize("dim", [1, 2, 3])
@pytest.mark.parametrize("kernel", [CUBIC, WENDLAND])
def test_compact_support(dim, kernel):
    for r in (1.0, 1.0000001, 1.5, 3.0):
        w, c, dh = kernel_eval(dim, kernel, r, 1.0)
        assert w == 0.0 and c == 0.0 and dh == 0.0


@pytest.mark.parametrize("dim", [2, 3])
@pytest.mark.parametrize("kernel", [CUBIC, WENDLAND])
def test_derivatives_match_finite_differences(dim, kernel):
    """Mean abs error of the analytic derivatives against central
    differences at n = 6400 (kernel_test.cpp uses n = 100 ... 6400)."""
    n = 6400
    dx = 1.0 / n
    xs = (np.arange(n) + 0.5) * dx
    err_r = err_h = 0.0
    for x in xs:
        _, c, dh = kernel_eval(dim, kernel, x, 1.0)
        fd_r = (kernel_eval(dim, kernel, x + dx / 2, 1.0)[0] - kernel_eval(dim, kernel, x - dx / 2, 1.0)[0]) / dx
        fd_h = (kernel_eval(dim, kernel, x, 1.0 + dx / 2)[0] - kernel_eval(dim, kernel, x, 1.0 - dx / 2)[0]) / dx
        err_r += abs(c * x - fd_r)
        err_h += abs(dh - fd_h)
    assert err_r / n < 1e-6
    assert err_h / n < 1e-6


def test_derivatives_1d_cubic():
    n = 6400
    dx = 1.0 / n
    for x in (np.arange(n) + 0.5) * dx:
        _, c, dh = kernel_eval(1, CUBIC, x, 1.0)
        fd_r = (kernel_eval(1, CUBIC, x + dx / 2, 1.0)[0] - kernel_eval(1, CUBIC, x - dx / 2, 1.0)[0]) / dx
        fd_h = (kernel_eval(1, CUBIC, x, 1.0 + dx / 2)[0] - kernel_eval(1, CUBIC, x, 1.0 - dx / 2)[0]) / dx
        assert abs(c * x - fd_r) < 1e-5 and abs(dh - fd_h) < 1e-5


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_central_values(dim):
    sig_cubic = {1: 2 / 3, 2: 10 / (7 * math.pi), 3: 1 / math.pi}[dim]
    sig_c4 = {1: 0.0, 2: 9 / math.pi, 3: 495 / (32 * math.pi)}[dim]
    for h in (0.5, 1.0, 3.0):
        w, _, _ = kernel_eval(dim, CUBIC, 0.0, h)
        assert w == pytest.approx(sig_cubic / (h / 2) ** dim, rel=1e-14)
        w, _, _ = kernel_eval(dim, WENDLAND, 0.0, h)
        assert w == pytest.approx(sig_c4 / h ** dim, rel=1e-14, abs=0)


def test_softening_force_is_gradient_of_potential():
    """Holds for u = 2r/h >= 1; the inner branch is the quirk below."""
    h = 0.8
    for r in np.linspace(h / 2 + 1e-3, 1.5, 300):
        d = 1e-6
        dfdr = (softening(r + d, h)[0] - softening(r - d, h)[0]) / (2 * d)
        f, g = softening(r, h)
        assert abs(g * r + dfdr) < 1e-6 * max(1.0, abs(dfdr)), r


def test_softening_inner_potential_quirk():
    """For u < 1 the reference's potential kernel (src/bhtree.cpp:279,
    src/gravity_force.cpp:21) is 1.4/e - (1.4/e - f_true)/4, where f_true is
    the potential whose gradient is the (pinned) force: f_true(r) =
    f(e) + integral_r^e g(s) s ds.  The published Hernquist & Katz kernel
    has 2 u^2 (...) where the reference has 0.5 u^2 (...); phi, and so the
    potential energy, carry that factor for pairs closer than h/2 (forces
    do not).  The restatement keeps it."""
    h = 0.9
    e = h / 2
    f_e = softening(e * (1 + 1e-13), h)[0]  # mid branch at u = 1 is consistent
    for r in np.linspace(0.0, e * 0.999, 40):
        tail, _ = integrate.quad(lambda s: softening(s, h)[1] * s, r, e, epsabs=1e-14, epsrel=1e-13)
        f_true = f_e + tail
        f_ref = softening(r, h)[0]
        assert f_ref == pytest.approx(1.4 / e - (1.4 / e - f_true) / 4, rel=1e-10), r


def test_softening_is_cubic_spline_cloud():
    """g(r) = M(<r)/r^3, M = mass fraction of the 3D cubic spline of support h."""
    h = 1.3
    for r in np.linspace(0.05, 1.6, 60):
        m, _ = integrate.quad(lambda s: 4 * math.pi * s * s * kernel_eval(3, CUBIC, s, h)[0], 0.0, min(r, h),
                              epsabs=1e-13, epsrel=1e-12, points=[h / 2] if r > h / 2 else None)
        _, g = softening(r, h)
        assert g == pytest.approx(m / r ** 3, rel=1e-9), r


def test_softening_limits():
    h = 0.6
    e = h / 2
    f0, _ = softening(0.0, h)
    assert f0 == pytest.approx(1.4 / e, rel=1e-15)
    for r in (h, 1.2 * h, 5.0):
        f, g = softening(r, h)
        assert f == pytest.approx(1 / r, rel=1e-14) and g == pytest.approx(1 / r ** 3, rel=1e-14)
    # continuity: force at u = 1 and u = 2, potential at u = 2
    for r in (e, 2 * e):
        lo, hi = softening(r * (1 - 1e-12), h), softening(r * (1 + 1e-12), h)
        assert lo[1] == pytest.approx(hi[1], rel=1e-9)
    lo, hi = softening(2 * e * (1 - 1e-12), h), softening(2 * e * (1 + 1e-12), h)
    assert lo[0] == pytest.approx(hi[0], rel=1e-9)
