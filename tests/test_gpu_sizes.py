"""Every even grid size n = 2^a 3^b from 8 to 1024 (the reference accepts
any n whose file size checks out, main.c:163-171; FFTW transforms any
length).  The sizes beyond the original twelve, with their own radix
schedules, pencil widths and z-pass blocks:

* small (8..162): FP64 transforms against numpy pocketfft; FP64 steps
  against the CPU oracle; mixed against FP64; every launch shape bitwise
  identical.
* 216..486: one FP64 step against the CPU oracle on a field of separable
  modes (Nyquist planes and modes beyond the 2/3 cutoff included).
* 576..972: the 3-D r2c of a separable field against the products of its
  1-D spectra (a closed form: every coefficient is a product of three 1-D
  DFT values), and the Beltrami flow's decay by the RK4 polynomial in FP64
  and mixed precision.
* n=864 (spectral state rows interleaved in pairs): the state of a
  separable field through dns_get_spectral / dns_set_spectral against its
  closed form.
* `dns-ic -l 3` (tgv.py -l 3: n = 8) through the command line, 1000 steps,
  against the oracle command line."""
import math
import os
import subprocess

import numpy as np
import pytest

import dns_amd
from conftest import DNS_CLI, DNS_IC, ORACLE_CLI, parse_lines, random_field, rel, rk4_amplification
from test_gpu_large import beltrami_broadcast, modes_field

pytestmark = pytest.mark.gpu

SMALL = [8, 12, 18, 24, 36, 54, 72, 108, 144, 162]
MEDIUM = [216, 288, 324, 432, 486]
LARGE = [576, 648, 864, 972]


def _smooth23(n):
    for p in (2, 3):
        while n % p == 0:
            n //= p
    return n == 1


def test_size_list_matches_library():
    import re
    with open(os.path.join(os.path.dirname(dns_amd.__file__), "csrc", "grid_sizes.h")) as f:
        listed = sorted(int(m) for m in re.findall(r"X\((\d+)\)", f.read()))
    assert listed == [n for n in range(8, 1025, 2) if _smooth23(n)]
    assert set(SMALL + MEDIUM + LARGE) <= set(listed)


@pytest.mark.parametrize("n", SMALL)
def test_transforms_fp64(n):
    rng = np.random.default_rng(n)
    a = rng.standard_normal((n, n, n))
    spec = rng.standard_normal((n, n, n // 2 + 1)) + 1j * rng.standard_normal((n, n, n // 2 + 1))
    print(f"transforms fp64 n={n}", flush=True)
    with dns_amd.Solver(n, 0.01, 0.01, "fp64") as s:
        got = s.rfft3(a)
        back = s.irfft3(spec)
    tol = 1e-15 * 3 * math.log2(n)
    assert rel(got, np.fft.rfftn(a)) < tol
    assert rel(back, np.fft.irfftn(spec, s=(n, n, n), axes=(0, 1, 2)) * n ** 3) < tol


@pytest.mark.parametrize("n", SMALL)
def test_steps_match_oracle(oracle, n):
    u = random_field(n, seed=n + 1)
    nu, dt = 0.02, 0.005
    print(f"steps vs oracle n={n}", flush=True)
    with dns_amd.Solver(n, nu, dt, "fp64") as g:
        o = oracle.solver(n, nu, dt)
        try:
            g.load_velocity(u)
            o.load_velocity(u)
            for k in (1, 4):
                g.step(k)
                o.step(k)
                assert rel(g.get_spectral(), o.get_spectral()) < 1e-12, k
                assert rel(g.get_stale_velocity(), o.get_stale_velocity()) < 1e-12, k
                ge, gz = g.diagnostics()
                oe, oz = o.diagnostics()
                assert abs(ge - oe) / oe < 1e-13 and abs(gz - oz) / oz < 1e-13
        finally:
            o.close()


@pytest.mark.parametrize("n", SMALL)
def test_mixed_tracks_fp64(n):
    u = random_field(n, seed=3 * n)
    nu, dt = 0.02, 0.005
    print(f"mixed vs fp64 n={n}", flush=True)
    with dns_amd.Solver(n, nu, dt, "fp64") as a, dns_amd.Solver(n, nu, dt, "mixed") as b:
        a.load_velocity(u)
        b.load_velocity(u)
        a.step(10)
        b.step(10)
        ea, za = a.diagnostics()
        eb, zb = b.diagnostics()
        assert abs(eb - ea) / ea < 1e-6 and abs(zb - za) / za < 1e-5
        assert rel(b.get_spectral(), a.get_spectral()) < 1e-5


@pytest.mark.parametrize("n", [24, 72, 108])
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_launch_shapes_bitwise_identical(n, precision):
    u = random_field(n, 5 * n)
    ref = None
    with dns_amd.Solver(n, 0.01, 0.004, precision) as s:
        assert s.launch_shapes() == 4
        for shape in range(4):
            s.set_launch_shape(shape)
            s.load_velocity(u)
            s.step(3)
            got = s.get_spectral()
            if ref is None:
                ref = got
            assert np.array_equal(got, ref), shape


def test_batch_of_new_sizes_matches_single_runs():
    ns = [18, 36, 18, 54]
    nus = [0.01, 0.005, 0.002, 0.004]
    with dns_amd.Batch(ns, nus, 2e-3, "fp64") as b:
        for m, n in enumerate(ns):
            b.load_velocity(m, random_field(n, 40 + m))
        b.step(5)
        for m, n in enumerate(ns):
            with dns_amd.Solver(n, nus[m], 2e-3, "fp64") as s:
                s.load_velocity(random_field(n, 40 + m))
                s.step(5)
                assert np.array_equal(b.get_spectral(m), s.get_spectral()), m


@pytest.mark.parametrize("n", MEDIUM)
def test_medium_step_matches_oracle(oracle, n):
    nu, dt = 1.0 / 400, 1e-3
    print(f"one step vs oracle n={n}", flush=True)
    u = modes_field(n, seed=n)
    with dns_amd.Solver(n, nu, dt, "fp64") as g:
        o = oracle.solver(n, nu, dt)
        try:
            g.load_velocity(u)
            o.load_velocity(u)
            del u
            g.step(1)
            o.step(1)
            assert rel(g.get_spectral(), o.get_spectral()) < 1e-12
            assert rel(g.get_stale_velocity(), o.get_stale_velocity()) < 1e-12
            ge, gz = g.diagnostics()
            oe, oz = o.diagnostics()
            assert abs(ge - oe) / oe < 1e-13 and abs(gz - oz) / oz < 1e-13
        finally:
            o.close()


def _separable(n, seed, nterms=5):
    """A real field sum_m A_m f_m(x) g_m(y) h_m(z) (sin or cos of integer
    wavenumbers up to n/2) and the dict {(i, j, k): coefficient} of its
    half-spectrum r2c, from the 1-D spectra of the factors."""
    rng = np.random.default_rng(seed)
    x = 2 * np.pi * np.arange(n) / n
    u = np.zeros((n, n, n))
    coef = {}
    for m in range(nterms):
        amp = rng.uniform(0.5, 1.0)
        fs = []
        for a in range(3):
            k = n // 2 if m == a else int(rng.integers(0, n // 2))
            f = np.cos(k * x) if (m == a or rng.integers(2)) else np.sin(k * x)
            fs.append(f)
        u += amp * fs[0][:, None, None] * fs[1][None, :, None] * fs[2][None, None, :]
        F0, F1, F2 = np.fft.fft(fs[0]), np.fft.fft(fs[1]), np.fft.rfft(fs[2])
        for i in np.flatnonzero(np.abs(F0) > 1e-9 * n):
            for j in np.flatnonzero(np.abs(F1) > 1e-9 * n):
                for k in np.flatnonzero(np.abs(F2) > 1e-9 * n):
                    coef[(i, j, k)] = coef.get((i, j, k), 0) + amp * F0[i] * F1[j] * F2[k]
    return u, coef


@pytest.mark.parametrize("n", LARGE)
def test_large_rfft3_closed_form(n):
    u, coef = _separable(n, seed=n)
    print(f"rfft3 closed form n={n}", flush=True)
    with dns_amd.Solver(n, 0.01, 0.01, "fp64") as s:
        got = s.rfft3(u)
    del u
    scale = max(abs(v) for v in coef.values())
    for (i, j, k), v in coef.items():
        assert abs(got[i, j, k] - v) < 1e-13 * scale, (i, j, k)
        got[i, j, k] = 0
    assert np.max(np.abs(got)) < 1e-13 * scale


@pytest.mark.parametrize("n", LARGE)
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-13), ("mixed", 1e-6)])
def test_large_beltrami(n, precision, tol):
    nu, dt, k, m = 0.002, 0.01, 5, 2
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


def test_interleaved_state_rows_round_trip():
    """n=864 keeps its spectral state with x rows interleaved in pairs (a
    pencil takes 72 threads; x tiles are 7 columns wide).  The state of a
    loaded separable field, read back through dns_get_spectral, holds the
    closed-form coefficients at their (i, j, k) positions; written back with
    dns_set_spectral it reads back bitwise; and the device's enstrophy sum,
    which weights every position by its |k|^2, equals the sum over the
    closed form."""
    n = 864
    u, coef = _separable(n, seed=7)
    print(f"interleaved state rows n={n}", flush=True)
    with dns_amd.Solver(n, 0.01, 0.01, "fp64") as s:
        s.load_velocity(np.stack([u, u, u]))
        del u
        e, z = s.diagnostics()
        spec = s.get_spectral()
        scale = max(abs(v) for v in coef.values())
        for (i, j, k), v in coef.items():
            assert np.all(np.abs(spec[:, i, j, k] - v) < 1e-13 * scale), (i, j, k)
        s.set_spectral(spec)
        back = s.get_spectral()
        assert np.array_equal(back, spec)
        del back, spec
    ew = ze = 0.0
    for (i, j, k), v in coef.items():   # plain half-spectrum sums over n^6 (main.c:257-266)
        kx, ky = (i if i < n // 2 else i - n), (j if j < n // 2 else j - n)
        a = 3 * abs(v) ** 2
        ew += a
        ze += (kx * kx + ky * ky + k * k) * a
    assert abs(e - ew / n ** 6) / e < 1e-12 and abs(z - ze / n ** 6) / z < 1e-12, (e, ew / n ** 6, z, ze / n ** 6)


def test_cli_level3_matches_oracle(tmp_path):
    """tgv.py -l 3 gives n = 8 (tgv.py:16); the drop-in runs it like the
    reference: 1000 steps of -n 0.01 -s 0.01 -t 9.995."""
    raw = tmp_path / "tgv8.raw"
    subprocess.run([DNS_IC, "-l", "3", "-o", str(raw)], check=True, capture_output=True, timeout=60)
    assert os.path.getsize(raw) == 4 * 8 ** 3 * 8
    args = ["-i", str(raw), "-n", "0.01", "-s", "0.01", "-t", "9.995"]
    g = subprocess.run([DNS_CLI, *args], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    o = subprocess.run([ORACLE_CLI, *args], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert g.returncode == 0 and o.returncode == 0, (g.stderr, o.stderr)
    a, b = parse_lines(g.stdout), parse_lines(o.stdout)
    assert a[0] == (0, 0.0, 0.125, 0.375) and a[-1][0] == 1000 and len(a) == len(b)
    for (sa, ta, ea, za), (sb, tb, eb, zb) in zip(a, b):
        assert sa == sb and ta == tb
        assert abs(ea - eb) <= 1e-12 * eb and abs(za - zb) <= 1e-12 * zb, sa
