"""Shared fixtures: the CPU oracle (ctypes over oracle/_build/liboracle.so),
Taylor-Green and random input fields, the golden trajectories."""
import ctypes
import json
import math
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "_build", "liboracle.so")
ORACLE_CLI = os.path.join(ORACLE_DIR, "_build", "dns_cpu")
DNS_CLI = os.path.join(ROOT, "dns_amd", "bin", "dns")
DNS_IC = os.path.join(ROOT, "dns_amd", "bin", "dns-ic")
GOLDEN = os.path.join(ROOT, "tests", "golden", "tgv_trajectories.json")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an AMD GPU (MI355X) and the built HIP library")


@pytest.fixture(scope="session", autouse=True)
def _torch_runtime_first(request):
    """When GPU tests run, initialise torch's HIP runtime at the start of the
    session, before libdns_hip's: the torch tests (autograd wrapper, device
    buffers) then find their device however many library contexts came and
    went before them (once, torch's late initialisation reported no HIP
    device after the adjoint tests had run)."""
    if any(item.get_closest_marker("gpu") for item in request.session.items):
        try:
            import torch
        except ImportError:
            return
        if torch.cuda.is_available():
            torch.cuda.init()


def _build_oracle():
    if not (os.path.exists(ORACLE_LIB) and os.path.exists(ORACLE_CLI)):
        subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


class Oracle:
    """ctypes view of oracle/oracle.h."""

    def __init__(self):
        _build_oracle()
        L = ctypes.CDLL(ORACLE_LIB)
        P, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        L.orc_fft1.argtypes = [I, P, ctypes.c_long, I]
        L.orc_rfft3.argtypes = [I, P, P]
        L.orc_irfft3.argtypes = [I, P, P]
        L.orc_create.restype = P
        L.orc_create.argtypes = [I, D, D]
        L.orc_destroy.argtypes = [P]
        L.orc_load.argtypes = [P, P]
        L.orc_step.argtypes = [P, ctypes.c_long]
        L.orc_diag.argtypes = [P, ctypes.POINTER(ctypes.c_longdouble), ctypes.POINTER(ctypes.c_longdouble)]
        L.orc_diag_d.argtypes = [P, ctypes.POINTER(D), ctypes.POINTER(D)]
        for f in ("orc_get_spectral", "orc_set_spectral", "orc_get_ustale", "orc_set_ustale",
                  "orc_get_phat", "orc_snapshot"):
            getattr(L, f).argtypes = [P, P]
        self.L = L

    def rfft3(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        n = a.shape[0]
        out = np.empty((n, n, n // 2 + 1), dtype=np.complex128)
        assert self.L.orc_rfft3(n, a.ctypes.data, out.ctypes.data) == 0
        return out

    def irfft3(self, s):
        s = np.ascontiguousarray(s, dtype=np.complex128)
        n = s.shape[0]
        out = np.empty((n, n, n), dtype=np.float64)
        assert self.L.orc_irfft3(n, s.ctypes.data, out.ctypes.data) == 0
        return out

    def fft1(self, a, sign):
        a = np.ascontiguousarray(a, dtype=np.complex128).copy()
        assert self.L.orc_fft1(a.shape[-1], a.ctypes.data, a.size // a.shape[-1], sign) == 0
        return a

    def solver(self, n, nu, dt):
        return OracleSolver(self, n, nu, dt)


class OracleSolver:
    def __init__(self, orc, n, nu, dt):
        self.o, self.n = orc, n
        self.p = orc.L.orc_create(n, nu, dt)

    def close(self):
        if self.p:
            self.o.L.orc_destroy(self.p)
            self.p = None

    def __del__(self):
        self.close()

    def load_velocity(self, uvw):
        uvw = np.ascontiguousarray(uvw, dtype=np.float64)
        self.o.L.orc_load(self.p, uvw.ctypes.data)

    def step(self, k=1):
        self.o.L.orc_step(self.p, k)

    def diagnostics(self):
        e, z = ctypes.c_longdouble(), ctypes.c_longdouble()
        self.o.L.orc_diag(self.p, ctypes.byref(e), ctypes.byref(z))
        return e.value, z.value

    def get_spectral(self):
        n = self.n
        out = np.empty((3, n, n, n // 2 + 1), dtype=np.complex128)
        self.o.L.orc_get_spectral(self.p, out.ctypes.data)
        return out

    def set_spectral(self, a):
        a = np.ascontiguousarray(a, dtype=np.complex128)
        self.o.L.orc_set_spectral(self.p, a.ctypes.data)

    def get_stale_velocity(self):
        n = self.n
        out = np.empty((3, n, n, n), dtype=np.float64)
        self.o.L.orc_get_ustale(self.p, out.ctypes.data)
        return out

    def set_stale_velocity(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        self.o.L.orc_set_ustale(self.p, a.ctypes.data)

    def get_phat(self):
        n = self.n
        out = np.empty((n, n, n // 2 + 1), dtype=np.complex128)
        self.o.L.orc_get_phat(self.p, out.ctypes.data)
        return out

    def snapshot(self):
        n = self.n
        out = np.empty((4, n, n, n), dtype=np.float64)
        self.o.L.orc_snapshot(self.p, out.ctypes.data)
        return out


@pytest.fixture(scope="session")
def oracle():
    return Oracle()


def tgv_field(n):
    """Taylor-Green vortex as tgv.py:14-27 of the reference builds it."""
    dx = 2 * math.pi / n
    i = np.arange(n)
    s, c = np.sin(dx * i), np.cos(dx * i)
    a = np.zeros((4, n, n, n))
    a[0] = s[:, None, None] * c[None, :, None] * c[None, None, :]
    a[1] = (-c)[:, None, None] * s[None, :, None] * c[None, None, :]
    return a


def random_field(n, seed, nyquist=True):
    """Smooth-ish random real velocity (not divergence-free), with content
    in the Nyquist planes unless nyquist=False."""
    rng = np.random.default_rng(seed)
    u = rng.standard_normal((3, n, n, n))
    uh = np.fft.rfftn(u, axes=(1, 2, 3))
    k = np.fft.fftfreq(n, 1.0 / n)
    kz = np.arange(n // 2 + 1)
    kk = k[:, None, None] ** 2 + k[None, :, None] ** 2 + kz[None, None, :] ** 2
    uh *= np.exp(-kk / (0.08 * n * n))[None]
    if not nyquist:
        uh[:, n // 2, :, :] = 0
        uh[:, :, n // 2, :] = 0
        uh[:, :, :, n // 2] = 0
    return np.fft.irfftn(uh, s=(n, n, n), axes=(1, 2, 3))


def write_raw(path, uvw, p=None):
    n = uvw.shape[-1]
    a = np.zeros((4, n, n, n))
    a[:3] = uvw[:3]
    if p is not None:
        a[3] = p
    a.tofile(path)


def load_golden():
    with open(GOLDEN) as f:
        return json.load(f)


def parse_lines(text):
    rows = []
    for line in text.splitlines():
        parts = line.split()
        if len(parts) == 4 and parts[0].lstrip("-").isdigit():
            rows.append((int(parts[0]), float(parts[1]), float(parts[2]), float(parts[3])))
    return rows


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def rk4_amplification(z):
    """RK4 amplification factor of u' = -u over one step of size z."""
    return 1 - z + z ** 2 / 2 - z ** 3 / 6 + z ** 4 / 24


def beltrami_field(n, k=1, a=1.0, b=0.7, c=0.4):
    """ABC (Beltrami) flow with wavenumber k: curl u = k u, so u x omega = 0
    and the flow decays by viscosity alone."""
    x = 2 * np.pi * np.arange(n) / n
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    return np.stack([a * np.sin(k * Z) + c * np.cos(k * Y),
                     b * np.sin(k * X) + a * np.cos(k * Z),
                     c * np.sin(k * Y) + b * np.cos(k * X)])


def helical_spectrum(n, ks, seed, scale=None):
    """Unnormalised half spectrum [3][n][n][n/2+1] (numpy rfftn convention;
    coefficient scale * A h / 2 at k, scale n^3 by default, 1 for a spectrum
    whose unnormalised c2r is u itself) of u = sum over ks of
    Re(A h e^{i k.x}): h the eigenvector of the
    Fourier curl i k x . with eigenvalue +|k|, A a seeded complex amplitude.
    Every k has kz > 0 and the same |k|, so curl u = |k| u (a Beltrami field
    with modes at many kz).  The array is allocated lazily (np.zeros): only
    the pages of the modes are written."""
    ks = [tuple(int(v) for v in k) for k in ks]
    shells = {k[0] ** 2 + k[1] ** 2 + k[2] ** 2 for k in ks}
    assert len(shells) == 1 and all(0 < k[2] < n // 2 for k in ks) and len(set(ks)) == len(ks)
    kk = shells.pop()
    rng = np.random.default_rng(seed)
    spec = np.zeros((3, n, n, n // 2 + 1), dtype=np.complex128)
    for k in ks:
        kx, ky, kz = k
        curl = 1j * np.array([[0, -kz, ky], [kz, 0, -kx], [-ky, kx, 0]], dtype=float)
        w, v = np.linalg.eig(curl)
        h = v[:, int(np.argmin(np.abs(w - math.sqrt(kk))))]
        a = (rng.uniform(0.5, 1.0) * np.exp(2j * np.pi * rng.uniform())) * h / np.linalg.norm(h)
        spec[:, kx % n, ky % n, kz] = (float(n) ** 3 if scale is None else scale) * a / 2
    return spec, kk


def wavenumbers(n):
    """kx, ky, kz of the half-spectrum layout [n][n][n/2+1] (Nyquist of x
    and y at -n/2, main.c:224-230)."""
    w = np.arange(n)
    w = np.where(w < n // 2, w, w - n).astype(float)
    return w[:, None, None], w[None, :, None], np.arange(n // 2 + 1, dtype=float)[None, None, :]


def divergence_hat(uh):
    """k . U over the stored half spectrum (without the factor i)."""
    n = uh.shape[1]
    kx, ky, kz = wavenumbers(n)
    return kx * uh[0] + ky * uh[1] + kz * uh[2]
