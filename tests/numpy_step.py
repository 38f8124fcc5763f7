"""Test infrastructure: numpy restatement of one RK4 step of main.c:334-420
(pocketfft transforms, same layout and conventions as FFTW's r2c/c2r) and
its tangent-linear model, used by the adjoint dot-product tests.

State: (Uh [3,n,n,n/2+1] complex, us [3,n,n,n] real stale velocity)."""
import numpy as np


def waves(n):
    kx = np.fft.fftfreq(n, 1.0 / n)            # Nyquist -n/2 (main.c:229-230)
    kz = np.arange(n // 2 + 1, dtype=float)     # kz in [0, n/2] (main.c:224-228)
    KX, KY, KZ = np.meshgrid(kx, kx, kz, indexing="ij")
    kmax = 2.0 / 3.0 * (n // 2 + 1)
    mask = (np.abs(KX) < kmax) & (np.abs(KY) < kmax) & (np.abs(KZ) < kmax)
    return KX, KY, KZ, KX * KX + KY * KY + KZ * KZ, mask


def c2r(X, n):
    return np.fft.irfftn(X, s=(n, n, n), axes=(-3, -2, -1)) * n ** 3


def r2c(x):
    return np.fft.rfftn(x, axes=(-3, -2, -1))


def curl(U, K):
    KX, KY, KZ = K[:3]
    return 1j * np.stack([KY * U[2] - KZ * U[1], KZ * U[0] - KX * U[2], KX * U[1] - KY * U[0]])


def cross(a, b):
    return np.stack([a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]])


def project(D, K):
    KX, KY, KZ, KK = K[:4]
    P = np.where(KK > 0, (D[0] * KX + D[1] * KY + D[2] * KZ) / np.where(KK > 0, KK, 1), 0)
    return D - np.stack([P * KX, P * KY, P * KZ])


A_RK = (1 / 6.0, 1 / 3.0, 1 / 3.0, 1 / 6.0)
B_RK = (0.5, 0.5, 1.0)


def step_tlm(Uh, us, dUh, dus, nu, dt):
    """One step and its tangent: returns (Uh', us', dUh', dus')."""
    n = us.shape[-1]
    K = waves(n)
    KK, mask = K[3], K[4]
    s = 1.0 / n ** 3
    Ur, dUr = Uh, dUh
    acc, dacc = Uh.copy(), dUh.copy()
    for r in range(4):
        if r == 0:
            u, du = us, dus
        else:
            u, du = s * c2r(Ur, n), s * c2r(dUr, n)
        w, dw = s * c2r(curl(Ur, K), n), s * c2r(curl(dUr, K), n)
        N = cross(u, w)
        dN = cross(du, w) + cross(u, dw)
        dU = project(dt * mask * r2c(N), K) - nu * dt * KK * Ur
        ddU = project(dt * mask * r2c(dN), K) - nu * dt * KK * dUr
        acc = acc + A_RK[r] * dU
        dacc = dacc + A_RK[r] * ddU
        if r < 3:
            Ur, dUr = Uh + B_RK[r] * dU, dUh + B_RK[r] * ddU
        if r == 3:
            us_new, dus_new = u, du
    return acc, us_new, dacc, dus_new


def inner(a_h, a_s, b_h, b_s):
    """<a, b> = Re sum conj(a_h) b_h + sum a_s b_s (the adjoint's inner product)."""
    return float(np.real(np.vdot(a_h, b_h)) + np.dot(a_s.ravel(), b_s.ravel()))


def solenoidal_field(n, seed, k0=4.0):
    """Seeded random divergence-free real velocity (for test directions)."""
    rng = np.random.default_rng(seed)
    KX, KY, KZ, KK, _ = waves(n)
    g = rng.standard_normal((3, n, n, n // 2 + 1)) + 1j * rng.standard_normal((3, n, n, n // 2 + 1))
    g = project(g, (KX, KY, KZ, KK))
    amp = np.sqrt(KK) * np.exp(-(KK / (k0 * k0)))
    g *= amp
    g[:, n // 2] = 0
    g[:, :, n // 2] = 0
    g[:, :, :, n // 2] = 0
    return c2r(g, n) / n ** 3
