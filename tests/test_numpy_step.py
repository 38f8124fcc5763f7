"""Pins the numpy restatement used by the adjoint dot-product tests:
its forward step equals the CPU oracle's (itself pinned to the reference's
golden trajectories), and its tangent-linear model equals central finite
differences of that forward step."""
import numpy as np

from conftest import random_field, rel
from numpy_step import r2c, step_tlm


def test_forward_matches_oracle(oracle):
    n, nu, dt = 16, 0.02, 0.01
    u = random_field(n, 4)
    o = oracle.solver(n, nu, dt)
    o.load_velocity(u)
    Uh, us = o.get_spectral(), o.get_stale_velocity()
    o.step(1)
    U1, us1, _, _ = step_tlm(Uh, us, 0 * Uh, 0 * us, nu, dt)
    assert rel(U1, o.get_spectral()) < 1e-13
    assert rel(us1, o.get_stale_velocity()) < 1e-13


def test_tangent_matches_finite_differences():
    n, nu, dt = 16, 0.02, 0.01
    rng = np.random.default_rng(3)
    u = random_field(n, 5)
    Uh, us = r2c(u), u + 0.1 * random_field(n, 6)
    dUh = r2c(random_field(n, 7)) + 0.01 * (rng.standard_normal(Uh.shape) + 1j * rng.standard_normal(Uh.shape))
    dus = random_field(n, 8)
    _, _, tU, ts = step_tlm(Uh, us, dUh, dus, nu, dt)
    eps = 1e-5
    Up, sp, _, _ = step_tlm(Uh + eps * dUh, us + eps * dus, 0 * dUh, 0 * dus, nu, dt)
    Um, sm, _, _ = step_tlm(Uh - eps * dUh, us - eps * dus, 0 * dUh, 0 * dus, nu, dt)
    assert rel((Up - Um) / (2 * eps), tU) < 1e-8
    assert rel((sp - sm) / (2 * eps), ts) < 1e-8
