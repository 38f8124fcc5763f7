"""Pins the CPU oracle's FFTs (oracle/orc_fft.c) -- the FFTW substitute --
with checks that do not reuse its own formulas: brute-force DFT sums (the
method of poc/one.c and poc/three.c of the reference), numpy pocketfft
(same layout and sign convention as FFTW's r2c/c2r, main.c:221-223),
single-mode closed forms, round trip and Parseval."""
import math

import numpy as np
import pytest

from conftest import rel


def brute_rfft3(a):
    """Direct triple sum, as poc/three.c:46-66 does it."""
    n = a.shape[0]
    idx = np.arange(n)
    w = np.exp(-2j * math.pi * np.outer(idx, idx) / n)
    full = np.einsum("ia,jb,kc,abc->ijk", w, w, w, a.astype(complex))
    return full[:, :, : n // 2 + 1]


@pytest.mark.parametrize("n", [2, 4, 6, 8, 12])
def test_rfft3_brute_force(oracle, n):
    rng = np.random.default_rng(n)
    a = rng.standard_normal((n, n, n))
    assert rel(oracle.rfft3(a), brute_rfft3(a)) < 1e-13


@pytest.mark.parametrize("n", [16, 24, 48, 64, 96])
def test_rfft3_irfft3_vs_numpy(oracle, n):
    rng = np.random.default_rng(100 + n)
    a = rng.standard_normal((n, n, n))
    assert rel(oracle.rfft3(a), np.fft.rfftn(a)) < 2e-15 * math.log2(n) * 3
    s = np.fft.rfftn(a) + 0.3 * (rng.standard_normal((n, n, n // 2 + 1)) + 1j * rng.standard_normal((n, n, n // 2 + 1)))
    # non-Hermitian input: FFTW/pocketfft c2r = c2c on x,y then hc2r on z,
    # which drops Im of the kz=0 and kz=n/2 coefficients
    ref = np.fft.irfftn(s, s=(n, n, n), axes=(0, 1, 2)) * n ** 3
    assert rel(oracle.irfft3(s), ref) < 2e-15 * math.log2(n) * 3


@pytest.mark.parametrize("n", [2, 3, 4, 6, 8, 12, 16, 18, 27, 32, 48, 96, 128, 192, 256, 512, 1024])
@pytest.mark.parametrize("sign", [-1, 1])
def test_fft1_vs_numpy(oracle, n, sign):
    rng = np.random.default_rng(n)
    a = rng.standard_normal((5, n)) + 1j * rng.standard_normal((5, n))
    ref = np.fft.fft(a) if sign < 0 else np.fft.ifft(a) * n
    assert rel(oracle.fft1(a, sign), ref) < 1e-15 * max(1, math.log2(n)) * 2


def test_single_mode_closed_form(oracle):
    n = 16
    x = 2 * math.pi * np.arange(n) / n
    u = np.sin(x)[:, None, None] * np.ones((1, n, n))  # sin x, x = i dx
    s = oracle.rfft3(u)
    expect = np.zeros_like(s)
    expect[1, 0, 0] = -0.5j * n ** 3
    expect[n - 1, 0, 0] = 0.5j * n ** 3
    assert np.max(np.abs(s - expect)) < 1e-12 * n ** 3
    w = np.cos(3 * x)[None, None, :] * np.ones((n, n, 1))  # cos 3z on the half axis
    s = oracle.rfft3(w)
    expect = np.zeros_like(s)
    expect[0, 0, 3] = 0.5 * n ** 3
    assert np.max(np.abs(s - expect)) < 1e-12 * n ** 3


@pytest.mark.parametrize("n", [8, 32, 48])
def test_round_trip_and_parseval(oracle, n):
    rng = np.random.default_rng(7 * n)
    a = rng.standard_normal((n, n, n))
    s = oracle.rfft3(a)
    assert rel(oracle.irfft3(s) / n ** 3, a) < 1e-14
    w = np.full(n // 2 + 1, 2.0)
    w[0] = 1.0
    w[n // 2] = 1.0
    e_spec = np.sum(w * np.abs(s) ** 2) / n ** 3
    assert abs(e_spec - np.sum(a * a)) / np.sum(a * a) < 1e-13
