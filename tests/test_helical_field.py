"""The helical Beltrami field of conftest.helical_spectrum, on the CPU: its
inverse transform is the sum of the modes evaluated point by point, and
its curl (numpy's spectral derivative) is |k| times itself, so u x omega
vanishes at every grid point."""
import numpy as np

from conftest import helical_spectrum, wavenumbers

# |k|^2 = 26 with kz > 0
KS = [(1, 3, 4), (3, -4, 1), (-4, 1, 3), (4, 3, 1), (1, 0, 5), (-3, -1, 4), (0, -1, 5), (-1, 4, 3)]


def test_helical_spectrum_is_the_sum_of_its_modes():
    n = 16
    spec, kk = helical_spectrum(n, KS, seed=3)
    assert kk == 26
    u = np.stack([np.fft.irfftn(spec[c], s=(n, n, n), axes=(0, 1, 2)) for c in range(3)])
    x = 2 * np.pi * np.arange(n) / n
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    want = np.zeros_like(u)
    for kx, ky, kz in KS:
        a = 2 * spec[:, kx % n, ky % n, kz] / n ** 3
        ph = np.exp(1j * (kx * X + ky * Y + kz * Z))
        for c in range(3):
            want[c] += (a[c] * ph).real
    assert np.max(np.abs(u - want)) < 1e-13 * np.max(np.abs(want))


def test_helical_spectrum_is_beltrami():
    n = 32
    spec, kk = helical_spectrum(n, KS, seed=5)
    kx, ky, kz = wavenumbers(n)
    ox = 1j * (ky * spec[2] - kz * spec[1])
    oy = 1j * (kz * spec[0] - kx * spec[2])
    oz = 1j * (kx * spec[1] - ky * spec[0])
    u = np.stack([np.fft.irfftn(spec[c], s=(n, n, n), axes=(0, 1, 2)) for c in range(3)])
    w = np.stack([np.fft.irfftn(o, s=(n, n, n), axes=(0, 1, 2)) for o in (ox, oy, oz)])
    scale = np.max(np.abs(u))
    assert np.max(np.abs(w - np.sqrt(kk) * u)) < 1e-12 * np.sqrt(kk) * scale
    cross = np.cross(u, w, axis=0)
    assert np.max(np.abs(cross)) < 1e-11 * np.sqrt(kk) * scale ** 2
    # divergence-free and not degenerate: every mode carries energy
    div = kx * spec[0] + ky * spec[1] + kz * spec[2]
    assert np.max(np.abs(div)) < 1e-9 * np.max(np.abs(spec))
    for k in KS:
        assert np.linalg.norm(spec[:, k[0] % n, k[1] % n, k[2]]) > 0.4 * n ** 3 / 2
