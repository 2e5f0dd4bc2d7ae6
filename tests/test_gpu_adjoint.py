"""Gradient checks of the differentiable rollout (FP64).

The loss is L = sum m |v_K|^2 after Solver::initialize and K steps of fixed
dt.  The adjoint's directional derivatives <dL/dx0, d>, <dL/dv0, d>,
<dL/du0, d> along random directions must equal central differences of L:
  * of the GPU forward itself (K = 1, 2, 4), relative error <= 1e-5;
  * of the CPU oracle's forward (K = 2), which pins the gradient to the
    reference algorithm, relative error <= 1e-5.
Also: the rollout forward equals sphgpu_integrate with the same dt, bit for
bit; the context is left at the final state; unsupported configurations
are refused.
"""
from __future__ import annotations

import numpy as np
import pytest

from _cases import CASES
from _oracle import OracleEngine
from sphcode_amd import _abi, ics

pytestmark = pytest.mark.gpu

DT = 0.01


def _case(n_side=12, balsara=True, kernel=_abi.CUBIC_SPLINE):
    p, f = ics.rollout(n_side=n_side)
    p.av_use_balsara_switch = 1 if balsara else 0
    p.kernel = kernel
    return p, f


def _loss(fields):
    return float(np.sum(fields["mass"][:, None] * fields["vel"] ** 2))


def _gpu_loss(p, f, steps, dt=DT):
    from sphcode_amd import GPUEngine
    g = GPUEngine(p)
    g.upload(f)
    g.rollout(steps, dt)
    out = g.download(("mass", "vel"))
    g.close()
    return _loss(out)


def _oracle_loss(p, f, steps):
    o = OracleEngine(p)
    o.upload(f)
    o.initialize()
    for _ in range(steps):
        o.integrate(DT)
    out = o.download(("mass", "vel"))
    o.close()
    return _loss(out)


def _gradient(p, f, steps, dt=DT):
    from sphcode_amd import GPUEngine
    g = GPUEngine(p)
    g.upload(f)
    g.rollout(steps, dt)
    out = g.download(("mass", "vel"))
    seed = 2.0 * out["mass"][:, None] * out["vel"]
    gx, gv, gu = g.rollout_adjoint(dl_dv=seed)
    g.close()
    return gx, gv, gu


def _perturbed(f, field, d, eps):
    g = dict(f)
    g[field] = f[field] + eps * d
    if field == "pos":
        g[field] = np.mod(g[field], 1.0)
    return g


def _fd(loss_fn, p, f, field, d, eps, steps, **kw):
    lp = loss_fn(p, _perturbed(f, field, d, eps), steps, **kw)
    lm = loss_fn(p, _perturbed(f, field, d, -eps), steps, **kw)
    return (lp - lm) / (2 * eps)


# Small steps: the viscosity term Pi_ij is only continuous (not smooth) at
# v_ij . r_ij = 0, so a larger step lets a few pairs cross that kink and
# biases the difference quotient (observed: 1e-4 .. 1e-3 relative at 1e-6,
# 1e-7 .. 1e-6 at 1e-7 and 1e-8; without viscosity every step agrees).
EPS = {"pos": 1e-7, "vel": 1e-7, "ene": 1e-7}


@pytest.mark.parametrize("steps", [1, 2, 4])
@pytest.mark.parametrize("field", ["pos", "vel", "ene"])
def test_gradient_matches_gpu_finite_differences(gpu, steps, field):
    p, f = _case()
    gx, gv, gu = _gradient(p, f, steps)
    grad = {"pos": gx, "vel": gv, "ene": gu}[field]
    rng = np.random.default_rng(7)
    d = rng.normal(size=np.shape(f[field]))
    fd = _fd(_gpu_loss, p, f, field, d, EPS[field], steps)
    ad = float(np.sum(grad * d))
    print(f"steps={steps} field={field} adjoint={ad:.12e} fd={fd:.12e}", flush=True)
    assert abs(ad - fd) <= 1e-5 * abs(fd) + 1e-12


@pytest.mark.parametrize("variant", ["no_balsara", "wendland"])
def test_gradient_variants(gpu, variant):
    p, f = _case(balsara=variant != "no_balsara",
                 kernel=_abi.WENDLAND if variant == "wendland" else _abi.CUBIC_SPLINE)
    gx, gv, gu = _gradient(p, f, 2)
    rng = np.random.default_rng(11)
    for field, grad in (("pos", gx), ("vel", gv), ("ene", gu)):
        d = rng.normal(size=np.shape(f[field]))
        fd = _fd(_gpu_loss, p, f, field, d, EPS[field], 2)
        ad = float(np.sum(grad * d))
        print(f"{variant} field={field} adjoint={ad:.12e} fd={fd:.12e}", flush=True)
        assert abs(ad - fd) <= 1e-5 * abs(fd) + 1e-12


@pytest.mark.parametrize("dim,n_side,kernel", [(1, 64, _abi.CUBIC_SPLINE), (2, 48, _abi.CUBIC_SPLINE),
                                                (2, 48, _abi.WENDLAND)])
def test_gradient_lower_dimensions(gpu, dim, n_side, kernel):
    """1D and 2D periodic boxes (the 2D curl is a scalar; no Balsara switch
    in 1D, src/pre_interaction.cpp:116), dt = 2e-3 (within the CFL limit)."""
    p, f = ics.rollout(n_side=n_side, dim=dim)
    p.kernel = kernel
    if kernel == _abi.WENDLAND:
        p.neighbor_number = 32
    dt = 2e-3
    gx, gv, gu = _gradient(p, f, 3, dt=dt)
    rng = np.random.default_rng(13)
    for field, grad in (("pos", gx), ("vel", gv), ("ene", gu)):
        d = rng.normal(size=np.shape(f[field]))
        fd = _fd(_gpu_loss, p, f, field, d, EPS[field], 3, dt=dt)
        ad = float(np.sum(grad * d))
        print(f"dim={dim} kernel={kernel} field={field} adjoint={ad:.12e} fd={fd:.12e}", flush=True)
        assert abs(ad - fd) <= 1e-5 * abs(fd) + 1e-12


@pytest.mark.parametrize("field", ["pos", "vel", "ene"])
def test_gradient_matches_oracle_finite_differences(gpu, field):
    p, f = _case()
    gx, gv, gu = _gradient(p, f, 2)
    grad = {"pos": gx, "vel": gv, "ene": gu}[field]
    d = np.random.default_rng(5).normal(size=np.shape(f[field]))
    fd = _fd(_oracle_loss, p, f, field, d, EPS[field], 2)
    ad = float(np.sum(grad * d))
    print(f"oracle field={field} adjoint={ad:.12e} fd={fd:.12e}", flush=True)
    assert abs(ad - fd) <= 1e-5 * abs(fd) + 1e-12


def test_rollout_equals_integrate_and_leaves_final_state(gpu):
    from sphcode_amd import GPUEngine
    p, f = _case()
    a = GPUEngine(p)
    a.upload(f)
    a.rollout(3, DT)
    ra = a.download()
    b = GPUEngine(p)
    b.upload(f)
    b.initialize()
    for _ in range(3):
        b.integrate(DT)
    rb = b.download()
    for k in ("pos", "vel", "ene", "sml", "dens", "acc", "dene"):
        np.testing.assert_array_equal(ra[k], rb[k], err_msg=k)
    assert a.time()[0] == pytest.approx(3 * DT)
    a.rollout_adjoint(dl_dv=np.ones((a.n, 3)))
    rc = a.download()
    for k in ("pos", "vel", "ene", "acc", "dene", "sml", "dens"):
        np.testing.assert_array_equal(rc[k], ra[k], err_msg=k)
    # the context continues from there like the plain integrator
    a.integrate(DT)
    b.integrate(DT)
    np.testing.assert_array_equal(a.download(("vel",))["vel"], b.download(("vel",))["vel"])


def test_adjoint_is_linear_in_the_seed(gpu):
    from sphcode_amd import GPUEngine
    p, f = _case(n_side=8)
    g = GPUEngine(p)
    g.upload(f)
    g.rollout(2, DT)
    rng = np.random.default_rng(3)
    s1, s2 = rng.normal(size=(g.n, 3)), rng.normal(size=g.n)
    a1 = g.rollout_adjoint(dl_dv=s1)
    a2 = g.rollout_adjoint(dl_du=s2)
    a12 = g.rollout_adjoint(dl_dv=2.0 * s1, dl_du=-3.0 * s2)
    for x1, x2, x12 in zip(a1, a2, a12):
        np.testing.assert_allclose(x12, 2.0 * x1 - 3.0 * x2, rtol=1e-10, atol=1e-12 * np.abs(x12).max())


@pytest.mark.parametrize("change", ["disph", "tdav", "gravity", "no_iteration"])
def test_unsupported_configurations(gpu, change):
    from sphcode_amd import GPUEngine, SPHError
    p, f = _case(n_side=6)
    if change == "disph":
        p.sph_type = _abi.DISPH
    elif change == "tdav":
        p.av_use_time_dependent = 1
    elif change == "gravity":
        p.gravity_is_valid = 1
    else:
        p.iterative_sml = 0
    g = GPUEngine(p)
    g.upload(f)
    with pytest.raises(SPHError) as e:
        g.rollout(2, DT)
    assert e.value.code == _abi.E_INVALID


@pytest.mark.parametrize("steps", [1, 4])
def test_mixed_gradient_matches_fp64(gpu, steps):
    """Mixed precision (FP32 pair arithmetic in the forward and in the
    adjoint passes, FP64 accumulation): every gradient field within 1e-4
    of the FP64 gradient, relative in the 2-norm and per component against
    the field's RMS."""
    p, f = _case()
    g64 = _gradient(p, f, steps)
    pm = p.copy()
    pm.precision = _abi.MIXED
    gmx = _gradient(pm, f, steps)
    for name, a, b in zip(("pos", "vel", "ene"), gmx, g64):
        rel = np.linalg.norm(a - b) / np.linalg.norm(b)
        worst = np.max(np.abs(a - b)) / np.sqrt(np.mean(b * b))
        print(f"mixed steps={steps} {name}: norm rel {rel:.3e}, max/rms {worst:.3e}", flush=True)
        assert rel <= 1e-4 and worst <= 1e-3


def test_adjoint_without_rollout_is_a_state_error(gpu):
    from sphcode_amd import GPUEngine, SPHError
    p, f = _case(n_side=6)
    g = GPUEngine(p)
    g.upload(f)
    with pytest.raises(SPHError) as e:
        g.rollout_adjoint(dl_dv=np.ones((g.n, 3)))
    assert e.value.code == _abi.E_STATE


def test_torch_autograd_function(gpu):
    """The autograd binding returns the adjoint's gradients for a torch loss."""
    import torch
    from sphcode_amd import GPUEngine
    from sphcode_amd.autograd import rollout
    p, f = _case(n_side=8)
    eng = GPUEngine(p)
    pos = torch.tensor(f["pos"], requires_grad=True)
    vel = torch.tensor(f["vel"], requires_grad=True)
    ene = torch.tensor(f["ene"], requires_grad=True)
    m = torch.tensor(f["mass"])
    xk, vk, uk = rollout(eng, f, pos, vel, ene, 2, DT)
    loss = (m[:, None] * vk ** 2).sum() + 0.5 * (m * uk).sum()
    loss.backward()
    # the same through the C ABI directly
    g = GPUEngine(p)
    g.upload(f)
    g.rollout(2, DT)
    out = g.download(("vel", "ene", "mass"))
    gx, gv, gu = g.rollout_adjoint(dl_dv=2.0 * out["mass"][:, None] * out["vel"], dl_du=0.5 * out["mass"])
    np.testing.assert_array_equal(pos.grad.numpy(), gx)
    np.testing.assert_array_equal(vel.grad.numpy(), gv)
    np.testing.assert_array_equal(ene.grad.numpy(), gu)
