"""GPU RK4 step (FP64) vs the CPU oracle on identical inputs, vs the
reference's golden trajectory, and the mixed form's error budget.

Bars: FP64 state and diagnostics agree with the oracle to 1e-12 relative
(max norm); mixed energy within 1e-5 relative of the FP64 GPU run."""
import numpy as np
import pytest

import dns_amd
from conftest import (beltrami_field, divergence_hat, load_golden, random_field, rel, rk4_amplification,
                      tgv_field, wavenumbers)

pytestmark = pytest.mark.gpu


def _pair(oracle, n, nu, dt, uvw, precision="fp64"):
    g = dns_amd.Solver(n, nu, dt, precision)
    o = oracle.solver(n, nu, dt)
    g.load_velocity(uvw)
    o.load_velocity(uvw)
    return g, o


@pytest.mark.parametrize("n", [16, 32, 48, 64])
def test_load_matches_oracle(oracle, n):
    print(f"load n={n}", flush=True)
    u = random_field(n, seed=n)
    g, o = _pair(oracle, n, 0.01, 0.01, u)
    assert rel(g.get_spectral(), o.get_spectral()) < 1e-14
    assert rel(g.get_stale_velocity(), o.get_stale_velocity()) == 0.0
    ge, gz = g.diagnostics()
    oe, oz = o.diagnostics()
    assert abs(ge - oe) / oe < 1e-14 and abs(gz - oz) / oz < 1e-14
    g.close()


@pytest.mark.parametrize("n,seed", [(16, 1), (32, 2), (48, 3), (64, 4), (96, 5)])
def test_steps_match_oracle(oracle, n, seed):
    print(f"steps n={n}", flush=True)
    u = random_field(n, seed=seed)
    nu, dt = 0.02, 0.005
    g, o = _pair(oracle, n, nu, dt, u)
    for k in (1, 2, 5):
        g.step(k)
        o.step(k)
        assert rel(g.get_spectral(), o.get_spectral()) < 1e-12, k
        assert rel(g.get_stale_velocity(), o.get_stale_velocity()) < 1e-12, k
        ge, gz = g.diagnostics()
        oe, oz = o.diagnostics()
        assert abs(ge - oe) / oe < 1e-13 and abs(gz - oz) / oz < 1e-13
    g.close()


def test_tgv_initial_diagnostics():
    with dns_amd.Solver(64, 0.01, 0.01) as g:
        g.load_velocity(tgv_field(64)[:3])
        e, z = g.diagnostics()
    assert abs(e - 0.125) < 1e-16 and abs(z - 0.375) < 1e-15


def test_golden_trajectory_128_re100():
    gold = load_golden()["128/100"]
    n, nu, dt = gold["n"], float(gold["nu"]), gold["dt"]
    rows = {r[0]: (float(r[2]), float(r[3])) for r in gold["rows"]}
    print("golden 128/100, 100 steps", flush=True)
    with dns_amd.Solver(n, nu, dt) as g:
        g.load_velocity(tgv_field(n)[:3])
        for step in range(0, 101, 10):
            if step:
                g.step(10)
            e, z = g.diagnostics()
            ge, gz = rows[step]
            assert abs(e - ge) / ge < 1e-12, (step, e, ge)
            assert abs(z - gz) / gz < 1e-12, (step, z, gz)


def test_pressure_and_snapshot_match_oracle(oracle):
    n, nu, dt = 32, 0.02, 0.01
    u = random_field(n, seed=11)
    g, o = _pair(oracle, n, nu, dt, u)
    snap0 = g.snapshot()
    assert np.all(snap0[3] == 0)
    g.set_pressure_capture(True)
    g.step(3)
    o.step(3)
    so, sg = o.snapshot(), g.snapshot()
    assert rel(sg[:3], so[:3]) < 1e-12
    assert rel(sg[3], so[3]) < 1e-11
    g.close()


def test_state_roundtrip_and_restart(oracle):
    n = 32
    u = random_field(n, seed=5)
    g, o = _pair(oracle, n, 0.02, 0.01, u)
    o.step(2)
    g.set_spectral(o.get_spectral())
    g.set_stale_velocity(o.get_stale_velocity())
    assert rel(g.get_spectral(), o.get_spectral()) == 0.0
    g.step(2)
    o.step(2)
    assert rel(g.get_spectral(), o.get_spectral()) < 1e-12
    g.close()


def test_mixed_energy_error_tgv():
    n, nu, dt = 64, 0.01, 0.01
    print("mixed vs fp64 n=64, 100 steps", flush=True)
    with dns_amd.Solver(n, nu, dt, "fp64") as a, dns_amd.Solver(n, nu, dt, "mixed") as b:
        u = tgv_field(n)[:3]
        a.load_velocity(u)
        b.load_velocity(u)
        worst = 0.0
        for _ in range(10):
            a.step(10)
            b.step(10)
            ea, za = a.diagnostics()
            eb, zb = b.diagnostics()
            worst = max(worst, abs(eb - ea) / ea)
        print("mixed max rel energy error", worst)
        assert worst < 1e-5


def test_mixed_isotropic_field_max_norm(tmp_path):
    """Mixed vs FP64 on the seeded isotropic field of dns-ic (k0=4, E=0.5):
    energy and enstrophy within 1e-5 relative, and the physical velocity
    (the stale velocity after the step) within 1e-5 of max|u| in max norm."""
    import subprocess
    from conftest import DNS_IC
    n, nu, dt = 64, 1 / 800, 2e-3
    raw = tmp_path / "iso64.raw"
    subprocess.run([DNS_IC, "-t", "iso", "-N", str(n), "-o", str(raw)], check=True, capture_output=True)
    u = np.fromfile(raw, dtype="<f8").reshape(4, n, n, n)[:3].copy()
    print("mixed vs fp64 isotropic n=64, 100 steps", flush=True)
    with dns_amd.Solver(n, nu, dt, "fp64") as a, dns_amd.Solver(n, nu, dt, "mixed") as b:
        a.load_velocity(u)
        b.load_velocity(u)
        we = wz = 0.0
        for _ in range(10):
            a.step(10)
            b.step(10)
            ea, za = a.diagnostics()
            eb, zb = b.diagnostics()
            we = max(we, abs(eb - ea) / ea)
            wz = max(wz, abs(zb - za) / za)
        ua, ub = a.get_stale_velocity(), b.get_stale_velocity()
        wu = np.abs(ub - ua).max() / np.abs(ua).max()
        print(f"mixed max rel error: E {we:.2e} Z {wz:.2e} u (max norm) {wu:.2e}")
        assert we < 1e-5 and wz < 1e-5 and wu < 1e-5


def test_mixed_running_sum_in_double():
    """DNS_MIXED_ACC64 differs from DNS_MIXED only in the running sum of the
    stage increments (double instead of float): the two runs differ (the
    double sum is in use) by no more than the float rounding of that sum,
    and the double-sum run is no further from the FP64 run than the float
    one by more than that rounding."""
    n, nu, dt = 64, 0.01, 0.004
    u = random_field(n, 7)
    out = {}
    for p in ("fp64", "mixed", "mixed_acc64"):
        with dns_amd.Solver(n, nu, dt, p) as s:
            s.load_velocity(u)
            lines = [s.diagnostics()]
            for _ in range(5):
                s.step(10)
                lines.append(s.diagnostics())
            out[p] = (np.array(lines), s.get_spectral())
    (e64, _), (ef, sf), (ed, sd) = out["fp64"], out["mixed"], out["mixed_acc64"]
    assert not np.array_equal(sf, sd)
    rel = np.max(np.abs(sf - sd)) / np.max(np.abs(sd))
    print(f"float vs double running sum after 50 steps: spectral max-norm {rel:.2e}", flush=True)
    assert rel <= 1e-6  # measured 9.1e-9
    for name, e in (("mixed", ef), ("mixed_acc64", ed)):
        de = np.max(np.abs(e[:, 0] - e64[:, 0]) / e64[:, 0])
        print(f"{name}: max rel energy dev vs fp64 {de:.2e}", flush=True)
        assert de <= 1e-5


@pytest.mark.parametrize("precision", ["fp64", "mixed", "mixed_acc64"])
def test_state_transfer_continues_bitwise(precision):
    """A state taken out with get_spectral / get_stale_velocity and put into
    a second solver with set_spectral / set_stale_velocity continues bit for
    bit (the library's internal state layout and, in mixed precision, its
    float copy of the spectral state are rebuilt by the setters)."""
    n, nu, dt = 48, 0.01, 0.004
    u = random_field(n, 41)
    with dns_amd.Solver(n, nu, dt, precision) as a, dns_amd.Solver(n, nu, dt, precision) as b:
        a.load_velocity(u)
        a.step(5)
        b.load_velocity(random_field(n, 42))
        b.step(2)
        b.set_spectral(a.get_spectral())
        b.set_stale_velocity(a.get_stale_velocity())
        assert np.array_equal(b.get_spectral(), a.get_spectral())
        a.step(6)
        b.step(6)
        assert np.array_equal(b.get_spectral(), a.get_spectral())
        assert np.array_equal(b.get_stale_velocity(), a.get_stale_velocity())


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-13), ("mixed", 1e-6), ("mixed_acc64", 1e-6)])
def test_beltrami_flow_decays_by_rk4_polynomial(precision, tol):
    """curl u = k u, so u x omega = 0: the state decays by R(nu dt k^2)."""
    n, nu, dt, k, m = 32, 0.02, 0.05, 3, 20
    print(f"beltrami {precision} n={n}", flush=True)
    with dns_amd.Solver(n, nu, dt, precision) as s:
        s.load_velocity(beltrami_field(n, k))
        u0 = s.get_spectral()
        s.step(m)
        want = u0 * rk4_amplification(nu * dt * k * k) ** m
        assert np.max(np.abs(s.get_spectral() - want)) / np.max(np.abs(u0)) < tol


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-13), ("mixed", 1e-6), ("mixed_acc64", 1e-6)])
def test_divergence_decays_by_rk4_polynomial(precision, tol):
    """The projection removes the divergent part of N, so k.U of any field
    evolves by R(nu dt |k|^2) per step."""
    n, nu, dt, m = 32, 0.02, 0.01, 8
    rng = np.random.default_rng(18)
    u = rng.standard_normal((3, n, n, n))  # not solenoidal
    print(f"divergence {precision} n={n}", flush=True)
    with dns_amd.Solver(n, nu, dt, precision) as s:
        s.load_velocity(u)
        d0 = divergence_hat(s.get_spectral())
        s.step(m)
        kx, ky, kz = wavenumbers(n)
        want = d0 * rk4_amplification(nu * dt * (kx ** 2 + ky ** 2 + kz ** 2)) ** m
        got = divergence_hat(s.get_spectral())
        assert np.max(np.abs(got - want)) / np.max(np.abs(d0)) < tol


def test_unsupported_grid_size_raises():
    with pytest.raises(dns_amd.DnsError) as e:
        dns_amd.Solver(100, 0.01, 0.01)
    assert e.value.code == -1 and "unsupported grid size n=100" in str(e.value)
