"""GPU 3-D transforms (the FFTW replacement) vs numpy pocketfft, which
shares FFTW's layout and conventions (main.c:221-223).  FP64 must agree to
round-off; the mixed form (FP32 kernels, twiddles rounded from double)
to FP32 accuracy."""
import math

import numpy as np
import pytest

import dns_amd
from conftest import rel

pytestmark = pytest.mark.gpu

SIZES = [16, 32, 48, 64, 96, 128, 192, 256]


@pytest.mark.parametrize("n", SIZES)
def test_rfft3_fp64(n):
    rng = np.random.default_rng(n)
    a = rng.standard_normal((n, n, n))
    print(f"rfft3 fp64 n={n}", flush=True)
    with dns_amd.Solver(n, 0.01, 0.01, "fp64") as s:
        got = s.rfft3(a)
    assert rel(got, np.fft.rfftn(a)) < 1e-15 * 3 * math.log2(n)


@pytest.mark.parametrize("n", SIZES)
def test_irfft3_fp64_non_hermitian(n):
    rng = np.random.default_rng(1000 + n)
    spec = rng.standard_normal((n, n, n // 2 + 1)) + 1j * rng.standard_normal((n, n, n // 2 + 1))
    print(f"irfft3 fp64 n={n}", flush=True)
    with dns_amd.Solver(n, 0.01, 0.01, "fp64") as s:
        got = s.irfft3(spec)
    ref = np.fft.irfftn(spec, s=(n, n, n), axes=(0, 1, 2)) * n ** 3
    assert rel(got, ref) < 1e-15 * 3 * math.log2(n)


@pytest.mark.parametrize("n", [32, 48, 128, 256])
def test_transforms_mixed(n):
    rng = np.random.default_rng(2000 + n)
    a = rng.standard_normal((n, n, n))
    print(f"fft mixed n={n}", flush=True)
    with dns_amd.Solver(n, 0.01, 0.01, "mixed") as s:
        got = s.rfft3(a)
        back = s.irfft3(np.fft.rfftn(a)) / n ** 3
    assert rel(got, np.fft.rfftn(a)) < 2e-7 * math.log2(n)
    assert rel(back, a) < 2e-7 * math.log2(n)


def test_single_mode_closed_form():
    n = 64
    x = 2 * math.pi * np.arange(n) / n
    u = np.sin(x)[:, None, None] * np.ones((1, n, n))
    with dns_amd.Solver(n, 0.01, 0.01, "fp64") as s:
        got = s.rfft3(u)
    expect = np.zeros_like(got)
    expect[1, 0, 0] = -0.5j * n ** 3
    expect[n - 1, 0, 0] = 0.5j * n ** 3
    assert np.max(np.abs(got - expect)) < 1e-12 * n ** 3
