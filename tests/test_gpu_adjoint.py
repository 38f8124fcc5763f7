"""Discrete adjoint (dns_adjoint_*) checks.

1. Dot-product test per step: <TLM(d), lam> == <d, ADJ(lam)> with the
   tangent-linear model restated in numpy (tests/numpy_step.py), random
   state, direction and adjoint seed, including the stale-velocity part.
2. The full gradient of the final enstrophy against central finite
   differences (eps = 1e-6) of the CPU oracle's FP64 runs along seeded
   divergence-free directions: <grad, d> within 1e-6 relative.
3. Checkpoint interval does not change the gradient.
4. The torch.autograd.Function wrapper returns the same gradient."""
import numpy as np
import pytest

import dns_amd
from conftest import random_field, rel
from numpy_step import inner, r2c, solenoidal_field, step_tlm

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,seed", [(16, 1), (32, 2), (48, 3)])
def test_step_dot_product(n, seed):
    rng = np.random.default_rng(seed)
    nu, dt = 0.02, 0.01
    u = random_field(n, seed)
    Uh = r2c(u) + 1e-3 * (rng.standard_normal((3, n, n, n // 2 + 1)) + 1j * rng.standard_normal((3, n, n, n // 2 + 1)))
    us = u + 0.1 * random_field(n, seed + 7)
    dUh = r2c(random_field(n, seed + 1))
    dus = random_field(n, seed + 2)
    lUh = rng.standard_normal(Uh.shape) + 1j * rng.standard_normal(Uh.shape)
    lus = rng.standard_normal(us.shape)
    _, _, tUh, tus = step_tlm(Uh, us, dUh, dus, nu, dt)
    lhs = inner(tUh, tus, lUh, lus)
    print(f"adjoint step n={n}", flush=True)
    with dns_amd.Solver(n, nu, dt) as s:
        aUh, aus = s.adjoint_step(Uh, us, lUh, lus)
    rhs = inner(dUh, dus, aUh, aus)
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), abs(rhs)), (lhs, rhs)
    # the stale-velocity coupling alone
    _, _, tUh, tus = step_tlm(Uh, us, 0 * dUh, dus, nu, dt)
    lhs2 = inner(tUh, tus, lUh, lus)
    rhs2 = float(np.dot(dus.ravel(), aus.ravel()))
    assert abs(lhs2 - rhs2) <= 1e-12 * max(abs(lhs2), abs(rhs2))


def _fd(oracle, n, nu, dt, nsteps, u0, d, eps):
    vals = []
    for sgn in (1, -1):
        o = oracle.solver(n, nu, dt)
        o.load_velocity(u0 + sgn * eps * d)
        o.step(nsteps)
        vals.append(o.diagnostics()[1])
        o.close()
    return (float(vals[0]) - float(vals[1])) / (2 * eps)


@pytest.mark.parametrize("n,nsteps,every", [(16, 12, 5), (32, 10, 4)])
def test_gradient_vs_central_differences(oracle, n, nsteps, every):
    nu, dt = 1.0 / 200, 2e-3
    u0 = solenoidal_field(n, 11)
    u0 *= np.sqrt(0.5 / (0.5 * np.mean(np.sum(u0 * u0, axis=0))))
    print(f"gradient n={n} steps={nsteps}", flush=True)
    with dns_amd.Solver(n, nu, dt) as s:
        J = s.adjoint_forward(u0, nsteps, every)
        g = s.adjoint_backward()
    o = oracle.solver(n, nu, dt)
    o.load_velocity(u0)
    o.step(nsteps)
    assert abs(J - float(o.diagnostics()[1])) < 1e-12 * J
    o.close()
    for k in range(3):
        d = solenoidal_field(n, 100 + k)
        fd = _fd(oracle, n, nu, dt, nsteps, u0, d, 1e-6)
        ad = float(np.sum(g * d))
        assert abs(ad - fd) <= 1e-6 * abs(fd), (k, ad, fd)


def test_checkpoint_interval_invariance():
    n, nu, dt, nsteps = 16, 0.01, 5e-3, 9
    u0 = solenoidal_field(n, 5)
    grads = []
    for every in (1, 4, 9):
        with dns_amd.Solver(n, nu, dt) as s:
            s.adjoint_forward(u0, nsteps, every)
            grads.append(s.adjoint_backward())
    assert rel(grads[1], grads[0]) == 0.0 and rel(grads[2], grads[0]) == 0.0


def test_torch_autograd_wrapper():
    torch = pytest.importorskip("torch")
    from dns_amd.autograd import final_enstrophy
    n, nu, dt, nsteps = 16, 0.01, 5e-3, 6
    u0 = solenoidal_field(n, 9)
    with dns_amd.Solver(n, nu, dt) as s:
        J0 = s.adjoint_forward(u0, nsteps, 3)
        g0 = s.adjoint_backward()
    for dev in ("cpu", "cuda"):
        x = torch.tensor(u0, dtype=torch.float64, device=dev, requires_grad=True)
        J = final_enstrophy(x, nu, dt, nsteps, every=3)
        (2.5 * J).backward()
        assert abs(J.item() - J0) < 1e-14 * J0
        assert rel(x.grad.cpu().numpy(), 2.5 * g0) < 1e-14
