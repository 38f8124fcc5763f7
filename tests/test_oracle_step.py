"""Pins the CPU oracle's RK4 step (oracle/orc_solver.c) against the
reference's own golden trajectories (0128/*, 0256/* of the reference,
excerpted in tests/golden/) and against closed forms:
  - Taylor-Green E(0) = 1/8, Z(0) = 3/8 (poc/tgv.mac of the reference);
  - the linear (Stokes) limit: a shear flow whose u x omega is a pure
    gradient decays mode by mode with the RK4 amplification polynomial
    R(z) = 1 - z + z^2/2 - z^3/6 + z^4/24, z = nu dt |k|^2;
  - a Beltrami (ABC) flow, curl u = k u, so u x omega = 0 at every stage:
    the whole state decays by R(nu dt k^2) per step (pins the curl and the
    cross product: a swapped component or index leaves a nonzero N);
  - the divergence k.U of any field evolves mode by mode by R(nu dt |k|^2)
    per step whatever the nonlinear term does, because the projection
    removes the divergent part of N (pins the projection).
The golden runs also pin the stale-velocity rule of stage 1 (main.c:344-353):
a solver that refreshes u at stage 1 drifts by ~3e-13 within 10 steps at
n=128 and fails the 1e-13 bound used here."""
import math
import os
import subprocess

import numpy as np
import pytest

from conftest import (ORACLE_CLI, beltrami_field, divergence_hat, load_golden, parse_lines,
                      rk4_amplification, tgv_field, wavenumbers, write_raw)


def test_tgv_initial_diagnostics(oracle):
    n = 32
    s = oracle.solver(n, 0.01, 0.01)
    s.load_velocity(tgv_field(n)[:3])
    e, z = s.diagnostics()
    assert abs(e - 0.125) < 1e-15 and abs(z - 0.375) < 1e-15


def _run_cli(tmp_path, n, nu, dt, T):
    raw = os.path.join(tmp_path, f"tgv{n}.raw")
    if not os.path.exists(raw):
        write_raw(raw, tgv_field(n)[:3])
    r = subprocess.run([ORACLE_CLI, "-i", raw, "-n", nu, "-s", str(dt), "-t", str(T)],
                       capture_output=True, text=True, check=True)
    return parse_lines(r.stdout)


@pytest.mark.parametrize("key,T", [("128/100", 0.05), ("128/1600", 0.05), ("256/1600", 0.025)])
def test_oracle_cli_matches_golden(tmp_path, oracle, key, T):
    g = load_golden()[key]
    rows = _run_cli(str(tmp_path), g["n"], g["nu"], g["dt"], T)
    gold = {r[0]: (float(r[1]), float(r[2]), float(r[3])) for r in g["rows"]}
    assert len(rows) >= 2
    for step, t, e, z in rows:
        gt, ge, gz = gold[step]
        assert t == gt
        assert abs(e - ge) / ge < 1e-13, (step, e, ge)
        assert abs(z - gz) / gz < 1e-13, (step, z, gz)


def test_stokes_limit_rk4_polynomial(oracle):
    n, nu, dt, m = 16, 0.05, 0.1, 7
    x = 2 * math.pi * np.arange(n) / n
    u = np.zeros((3, n, n, n))
    u[0] = (np.sin(x)[None, :, None] + 0.5 * np.cos(2 * x)[None, :, None]) * np.ones((n, 1, n))
    s = oracle.solver(n, nu, dt)
    s.load_velocity(u)
    e0, _ = s.diagnostics()
    s.step(m)
    uh = s.get_spectral()
    for ky, amp in ((1, 0.5), (2, 0.25)):
        zz = nu * dt * ky * ky
        R = 1 - zz + zz ** 2 / 2 - zz ** 3 / 6 + zz ** 4 / 24
        got = np.abs(uh[0, 0, ky, 0]) / n ** 3
        assert abs(got - amp * R ** m) < 1e-14, (ky, got, amp * R ** m)
    # everything else stays at round-off
    uh[0, 0, [1, 2, n - 1, n - 2], 0] = 0
    assert np.max(np.abs(uh)) / n ** 3 < 1e-15


def test_beltrami_flow_decays_by_rk4_polynomial(oracle):
    n, nu, dt, k, m = 16, 0.02, 0.05, 2, 12
    s = oracle.solver(n, nu, dt)
    s.load_velocity(beltrami_field(n, k))
    u0 = s.get_spectral()
    s.step(m)
    want = u0 * rk4_amplification(nu * dt * k * k) ** m
    assert np.max(np.abs(s.get_spectral() - want)) / np.max(np.abs(u0)) < 1e-13


def test_divergence_decays_by_rk4_polynomial(oracle):
    n, nu, dt, m = 16, 0.02, 0.02, 6
    rng = np.random.default_rng(17)
    u = rng.standard_normal((3, n, n, n))  # not solenoidal
    s = oracle.solver(n, nu, dt)
    s.load_velocity(u)
    d0 = divergence_hat(s.get_spectral())
    s.step(m)
    kx, ky, kz = wavenumbers(n)
    want = d0 * rk4_amplification(nu * dt * (kx ** 2 + ky ** 2 + kz ** 2)) ** m
    got = divergence_hat(s.get_spectral())
    assert np.max(np.abs(got - want)) / np.max(np.abs(d0)) < 1e-13
