"""The large grids' own kernel instances on the GPU.

Every grid size compiles its own launch shapes: 16-column float tiles with
16 elements per thread and four-position epilogue batches from n=256, the
radix-3 schedules of 192, 384 and 768, the two-wave z pencils of n=1024.
The small-n parity tests do not run these.  Here: FP64 steps against the
CPU oracle up to n=512, mixed against FP64 on the same grids, and the
Beltrami closed form at n=768 and n=1024 (the oracle would take minutes
there; at n=1024 `bench.py --n 1024` also compares FP64 and mixed, whose
curl paths differ at that size: energy within ~4e-13).

Inputs are sums of separable Fourier modes (cheap to build at 512^3),
including modes beyond the 2/3 cutoff and on the Nyquist planes."""
import numpy as np
import pytest

import dns_amd
from conftest import beltrami_field, helical_spectrum, rel, rk4_amplification

pytestmark = pytest.mark.gpu


def modes_field(n, seed, nmodes=6):
    """Velocity made of `nmodes` separable modes per component,
    A sin|cos(kx x) sin|cos(ky y) sin|cos(kz z), wavenumbers up to n/2."""
    rng = np.random.default_rng(seed)
    x = 2 * np.pi * np.arange(n) / n
    u = np.zeros((3, n, n, n))
    for c in range(3):
        for m in range(nmodes):
            # the first mode of each component sits on the Nyquist plane of one
            # axis, the second beyond the 2/3 cutoff, the rest anywhere
            k = rng.integers(0, n // 3, size=3)
            if m == 0:
                k[c] = n // 2
            elif m == 1:
                k[(c + 1) % 3] = n // 2 - 2
            f = [np.cos(k[a] * x) if rng.integers(2) else np.sin(k[a] * x) for a in range(3)]
            if not np.any(f[0]) or not np.any(f[1]) or not np.any(f[2]):
                continue  # sin of a Nyquist wavenumber vanishes on the grid
            u[c] += rng.uniform(0.5, 1.0) * f[0][:, None, None] * f[1][None, :, None] * f[2][None, None, :]
    return u


@pytest.mark.parametrize("n", [128, 192, 256, 384, 512])
def test_fp64_steps_match_oracle(oracle, n):
    nu, dt = 1.0 / 400, 1e-3
    print(f"fp64 vs oracle n={n}", flush=True)
    u = modes_field(n, seed=n)
    with dns_amd.Solver(n, nu, dt, "fp64") as g:
        o = oracle.solver(n, nu, dt)
        try:
            g.load_velocity(u)
            o.load_velocity(u)
            del u
            g.step(2)
            o.step(2)
            assert rel(g.get_spectral(), o.get_spectral()) < 1e-12
            assert rel(g.get_stale_velocity(), o.get_stale_velocity()) < 1e-12
            ge, gz = g.diagnostics()
            oe, oz = o.diagnostics()
            assert abs(ge - oe) / oe < 1e-13 and abs(gz - oz) / oz < 1e-13
        finally:
            o.close()


@pytest.mark.parametrize("n", [256, 384, 512])
def test_mixed_tracks_fp64(n):
    nu, dt = 1.0 / 400, 1e-3
    print(f"mixed vs fp64 n={n}", flush=True)
    u = modes_field(n, seed=7 * n)
    with dns_amd.Solver(n, nu, dt, "fp64") as a, dns_amd.Solver(n, nu, dt, "mixed") as b:
        a.load_velocity(u)
        b.load_velocity(u)
        del u
        for _ in range(2):
            a.step(5)
            b.step(5)
            ea, za = a.diagnostics()
            eb, zb = b.diagnostics()
            assert abs(eb - ea) / ea < 1e-6 and abs(zb - za) / za < 1e-5
        assert rel(b.get_spectral(), a.get_spectral()) < 1e-5
        # the physical velocity is stored in float and this field carries
        # modes up to the Nyquist planes, where float rounding of the
        # transforms weighs most (n=384: 1.1e-5 of max |u| after 10 steps)
        assert rel(b.get_stale_velocity(), a.get_stale_velocity()) < 5e-5


def beltrami_broadcast(n, k, a=1.0, b=0.7, c=0.4):
    """The ABC field of conftest.beltrami_field built from 1-D factors."""
    x = 2 * np.pi * np.arange(n) / n
    s, co = np.sin(k * x), np.cos(k * x)
    u = np.empty((3, n, n, n))
    u[0] = a * s[None, None, :] + c * co[None, :, None]
    u[1] = b * s[:, None, None] + a * co[None, None, :]
    u[2] = c * s[None, :, None] + b * co[:, None, None]
    return u


def test_beltrami_broadcast_is_beltrami_field():
    assert np.array_equal(beltrami_broadcast(12, 2), beltrami_field(12, 2))


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-13), ("mixed", 1e-6)])
def test_beltrami_n768(precision, tol):
    """curl u = k u, so u x omega = 0 and every mode decays by the RK4
    polynomial R = R(nu dt k^2) per step: energy and enstrophy (computed on
    the device) scale by R^2 per step on the radix-3 grid n=768."""
    n, nu, dt, k, m = 768, 0.002, 0.01, 7, 3
    print(f"beltrami {precision} n={n}", flush=True)
    with dns_amd.Solver(n, nu, dt, precision) as s:
        s.load_velocity(beltrami_broadcast(n, k))
        e0, z0 = s.diagnostics()
        assert abs(z0 - k * k * e0) <= 1e-13 * z0   # a single shell |k| = k
        s.step(m)
        e, z = s.diagnostics()
        f = rk4_amplification(nu * dt * k * k) ** (2 * m)
        assert abs(e - e0 * f) / (e0 * f) < tol and abs(z - z0 * f) / (z0 * f) < tol


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-13), ("mixed", 1e-6)])
def test_beltrami_n1024(precision, tol):
    """The n=1024 instances (FP64: 16-element y passes, 4-column x tiles that
    also transform omega_x, about 160 GB of device memory; both: z pencils
    over two waves): energy and enstrophy decay by R^2 per step."""
    n, nu, dt, k, m = 1024, 0.002, 0.01, 9, 2
    print(f"beltrami {precision} n={n}", flush=True)
    with dns_amd.Solver(n, nu, dt, precision) as s:
        u = beltrami_broadcast(n, k)
        s.load_velocity(u)
        del u
        e0, z0 = s.diagnostics()
        assert abs(z0 - k * k * e0) <= 1e-13 * z0
        s.step(m)
        e, z = s.diagnostics()
        f = rk4_amplification(nu * dt * k * k) ** (2 * m)
        assert abs(e - e0 * f) / (e0 * f) < tol and abs(z - z0 * f) / (z0 * f) < tol


# |k|^2 = 25625 = 160^2 + 5^2 = 128^2 + 96^2 + 5^2 = 155^2 + 40^2
HELICAL = [(3, 4, 160), (-4, 3, 160), (5, 0, 160), (128, 5, 96), (-5, 128, 96), (128, 96, 5),
                (-96, 128, 5), (160, 3, 4), (4, 160, 3), (96, 5, 128), (155, 0, 40), (0, 155, 40)]


@pytest.mark.parametrize("n,precision,tol", [(1024, "fp64", 1e-12), (1024, "mixed", 1e-5), (972, "fp64", 1e-12)])
def test_helical_beltrami(n, precision, tol):
    """A Beltrami field of twelve helical modes with |k|^2 = 25625 at kz = 3,
    4, 5, 40, 96, 128 and 160 (the ABC fields above have kz = 0 and 9 only).
    The z pencils of n=1024 (128 threads) and n=972 (81 threads) span two
    waves and rebuild the mirrored half of each pencil from these rows, in
    threads of both waves (positions n - kz), and u x omega vanishes only if
    every point is right: then every coefficient of the state is multiplied
    by the RK4 polynomial R(nu dt |k|^2) per step.  The time step keeps the
    advective number kmax max|u| dt (about 342 * 5.3 * 0.001 = 1.8) inside
    RK4's stability bound of about 2.8: at dt = 0.01 (about 18) rounding
    noise at the dealias edge grows by orders of magnitude per step, in FP64
    (from 1e-16) as in mixed precision (from 1e-7, blowing up by step 4)."""
    nu, dt, m = 2e-4, 1e-3, 3
    print(f"helical beltrami {precision} n={n}", flush=True)
    spec, kk = helical_spectrum(n, HELICAL, seed=1024)
    cells = [(kx % n, ky % n, kz) for kx, ky, kz in HELICAL]
    f = rk4_amplification(nu * dt * kk) ** m

    def check(s):
        s.load_velocity(u)
        e0, z0 = s.diagnostics()
        assert abs(z0 - kk * e0) <= 1e-12 * z0
        s.step(m)
        e, z = s.diagnostics()
        a = s.get_spectral()
        # expected state: the initial coefficients times R^m, zero elsewhere
        for i, j, k in cells:
            a[:, i, j, k] -= spec[:, i, j, k] * f
        scale = max(float(np.abs(spec[:, i, j, k]).max()) for i, j, k in cells) * f
        err = max(float(np.abs(a[c]).max()) for c in range(3)) / scale
        de = abs(e - e0 * f ** 2) / (e0 * f ** 2)
        dz = abs(z - z0 * f ** 2) / (z0 * f ** 2)
        print(f"state {err:.3e} energy {de:.3e} enstrophy {dz:.3e}", flush=True)
        assert err < tol and de < tol and dz < tol

    # the physical field through the library's plain FP64 c2r, not the fused
    # z pass (a float c2r would leave rounding noise whose curl, at high
    # wavenumbers, no longer cancels in u x omega)
    phys, _ = helical_spectrum(n, HELICAL, seed=1024, scale=1.0)
    u = np.empty((3, n, n, n))
    with dns_amd.Solver(n, nu, dt, "fp64") as s:
        for c in range(3):
            u[c] = s.irfft3(phys[c])
        del phys
        if precision == "fp64":
            check(s)
            return
    with dns_amd.Solver(n, nu, dt, precision) as s:
        check(s)
