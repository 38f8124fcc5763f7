"""dns_amd: MI355X-native RK4 step of the pseudo-spectral Navier-Stokes
solver `dns` (slitvinov/dns).

The product is the C-ABI library ``libdns_hip.so`` (include/dns.h) and the
``dns`` command line built on it.  This module is a thin ctypes binding of
that library for Python callers (tests, benchmarks, torch autograd); it
never falls back to any other implementation -- if the library is missing
it raises.
"""
import ctypes
import os

import numpy as np

__all__ = ["Solver", "Batch", "DnsError", "lib", "LIB_PATH", "FP64", "MIXED", "abi_functions"]

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
# DNS_LIB selects another build of the library (e.g. to compare two builds)
LIB_PATH = os.environ.get("DNS_LIB") or os.path.join(PKG, "libdns_hip.so")
HEADER = os.path.join(ROOT, "include", "dns.h")
FP64, MIXED, MIXED_ACC64 = 0, 1, 2
PRECISIONS = {"fp64": FP64, "mixed": MIXED, "mixed_acc64": MIXED_ACC64}
KERNEL_SLOTS = 24

DNS_OK, DNS_EINVAL, DNS_ENODEV, DNS_ENOMEM, DNS_EHIP, DNS_ENONFINITE = 0, -1, -2, -3, -4, -5

_lib = None


class DnsError(RuntimeError):
    def __init__(self, code, where):
        self.code = code
        detail = lib().dns_last_error().decode()
        msg = f"{where}: {lib().dns_strerror(code).decode()} ({code})"
        super().__init__(msg + (f": {detail}" if detail else ""))


def lib():
    """Load libdns_hip.so (built in-tree by ``python -m dns_amd.build``)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: run `python -m dns_amd.build` first")
    L = ctypes.CDLL(LIB_PATH)
    P, D, I, LL = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_int, ctypes.c_long
    sig = {
        "dns_create": (I, [ctypes.POINTER(P), I, ctypes.c_double, ctypes.c_double, I, I]),
        "dns_destroy": (None, [P]),
        "dns_get_n": (I, [P]),
        "dns_load_velocity": (I, [P, P]),
        "dns_step": (I, [P, LL]),
        "dns_set_pressure_capture": (I, [P, I]),
        "dns_diagnostics": (I, [P, D, D]),
        "dns_snapshot": (I, [P, P]),
        "dns_snapshot_begin": (I, [P]),
        "dns_snapshot_wait": (I, [P, ctypes.POINTER(ctypes.POINTER(ctypes.c_double))]),
        "dns_get_spectral": (I, [P, P]),
        "dns_set_spectral": (I, [P, P]),
        "dns_get_stale_velocity": (I, [P, P]),
        "dns_set_stale_velocity": (I, [P, P]),
        "dns_rfft3": (I, [P, P, P]),
        "dns_irfft3": (I, [P, P, P]),
        "dns_device_bytes": (ctypes.c_longlong, [P]),
        "dns_set_kernel_timing": (I, [P, I]),
        "dns_kernel_timing": (I, [P, D, ctypes.POINTER(LL)]),
        "dns_kernel_slot_name": (ctypes.c_char_p, [I]),
        "dns_adjoint_forward": (I, [P, P, LL, LL, D]),
        "dns_adjoint_backward": (I, [P, ctypes.c_double, P]),
        "dns_adjoint_step": (I, [P, P, P, P, P, P, P]),
        "dns_launch_shapes": (I, [P]),
        "dns_set_launch_shape": (I, [P, I]),
        "dns_get_launch_shape": (I, [P]),
        "dns_autotune": (I, [P, I, ctypes.POINTER(I)]),
        "dns_batch_autotune": (I, [P, I]),
        "dns_batch_launch_shape": (I, [P, I]),
        "dns_batch_create": (I, [ctypes.POINTER(P), I, P, P, P, I, I]),
        "dns_batch_destroy": (None, [P]),
        "dns_batch_load_velocity": (I, [P, I, P]),
        "dns_batch_step": (I, [P, LL]),
        "dns_batch_diagnostics": (I, [P, I, D, D]),
        "dns_batch_get_spectral": (I, [P, I, P]),
        "dns_batch_get_stale_velocity": (I, [P, I, P]),
        "dns_batch_member_n": (I, [P, I]),
        "dns_batch_device_bytes": (ctypes.c_longlong, [P]),
        "dns_strerror": (ctypes.c_char_p, [I]),
        "dns_last_error": (ctypes.c_char_p, []),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def abi_functions(header=HEADER):
    """Names of the functions declared in include/dns.h."""
    import re
    text = open(header).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dns_[a-z0-9_]+)\s*\(", text)))


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def _check(rc, where):
    if rc != DNS_OK:
        raise DnsError(rc, where)


class Solver:
    """One solver context on one GPU (see include/dns.h)."""

    def __init__(self, n, nu, dt, precision=FP64, device=0):
        if isinstance(precision, str):
            precision = PRECISIONS[precision]
        self.n, self.nu, self.dt, self.precision = int(n), float(nu), float(dt), precision
        self._ctx = ctypes.c_void_p()
        _check(lib().dns_create(ctypes.byref(self._ctx), self.n, self.nu, self.dt, precision, device),
               "dns_create")

    @property
    def nf(self):
        return self.n // 2 + 1

    def close(self):
        if self._ctx:
            lib().dns_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def device_bytes(self):
        return lib().dns_device_bytes(self._ctx)

    def set_kernel_timing(self, on=True):
        _check(lib().dns_set_kernel_timing(self._ctx, int(bool(on))), "dns_set_kernel_timing")

    def kernel_timing(self):
        """{slot name: (total ms, launches)} since timing was switched on."""
        ms = (ctypes.c_double * KERNEL_SLOTS)()
        cnt = (ctypes.c_long * KERNEL_SLOTS)()
        _check(lib().dns_kernel_timing(self._ctx, ms, cnt), "dns_kernel_timing")
        return {lib().dns_kernel_slot_name(i).decode(): (ms[i], cnt[i]) for i in range(KERNEL_SLOTS)}

    def launch_shapes(self):
        return lib().dns_launch_shapes(self._ctx)

    def set_launch_shape(self, shape):
        _check(lib().dns_set_launch_shape(self._ctx, int(shape)), "dns_set_launch_shape")

    def launch_shape(self):
        return lib().dns_get_launch_shape(self._ctx)

    def autotune(self, steps=3):
        best = ctypes.c_int()
        _check(lib().dns_autotune(self._ctx, int(steps), ctypes.byref(best)), "dns_autotune")
        return best.value

    def load_velocity(self, uvw):
        uvw = np.ascontiguousarray(uvw, dtype=np.float64)
        assert uvw.shape == (3, self.n, self.n, self.n) or uvw.size >= 3 * self.n ** 3
        _check(lib().dns_load_velocity(self._ctx, _ptr(uvw)), "dns_load_velocity")

    def step(self, nsteps=1):
        _check(lib().dns_step(self._ctx, int(nsteps)), "dns_step")

    def set_pressure_capture(self, on=True):
        _check(lib().dns_set_pressure_capture(self._ctx, int(bool(on))), "dns_set_pressure_capture")

    def diagnostics(self, check=True):
        e, z = ctypes.c_double(), ctypes.c_double()
        rc = lib().dns_diagnostics(self._ctx, ctypes.byref(e), ctypes.byref(z))
        if check or rc != DNS_ENONFINITE:
            _check(rc, "dns_diagnostics")
        return e.value, z.value

    def snapshot(self):
        out = np.empty((4, self.n, self.n, self.n), dtype=np.float64)
        _check(lib().dns_snapshot(self._ctx, _ptr(out)), "dns_snapshot")
        return out

    def snapshot_begin(self):
        _check(lib().dns_snapshot_begin(self._ctx), "dns_snapshot_begin")

    def snapshot_wait(self):
        """A copy of the pinned buffer of the last snapshot_begin."""
        p = ctypes.POINTER(ctypes.c_double)()
        _check(lib().dns_snapshot_wait(self._ctx, ctypes.byref(p)), "dns_snapshot_wait")
        n = self.n
        return np.ctypeslib.as_array(p, shape=(4, n, n, n)).copy()

    def get_spectral(self):
        out = np.empty((3, self.n, self.n, self.nf), dtype=np.complex128)
        _check(lib().dns_get_spectral(self._ctx, _ptr(out)), "dns_get_spectral")
        return out

    def set_spectral(self, a):
        a = np.ascontiguousarray(a, dtype=np.complex128)
        assert a.shape == (3, self.n, self.n, self.nf)
        _check(lib().dns_set_spectral(self._ctx, _ptr(a)), "dns_set_spectral")

    def get_stale_velocity(self):
        out = np.empty((3, self.n, self.n, self.n), dtype=np.float64)
        _check(lib().dns_get_stale_velocity(self._ctx, _ptr(out)), "dns_get_stale_velocity")
        return out

    def set_stale_velocity(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.shape == (3, self.n, self.n, self.n)
        _check(lib().dns_set_stale_velocity(self._ctx, _ptr(a)), "dns_set_stale_velocity")

    def adjoint_forward(self, u0, nsteps, every=20):
        """Run nsteps from u0 keeping checkpoints; returns the final enstrophy."""
        u0 = np.ascontiguousarray(u0, dtype=np.float64)
        assert u0.shape == (3, self.n, self.n, self.n)
        z = ctypes.c_double()
        _check(lib().dns_adjoint_forward(self._ctx, _ptr(u0), int(nsteps), int(every), ctypes.byref(z)),
               "dns_adjoint_forward")
        return z.value

    def adjoint_backward(self, seed=1.0):
        """seed * d(final enstrophy)/d(u0) after adjoint_forward."""
        g = np.empty((3, self.n, self.n, self.n), dtype=np.float64)
        _check(lib().dns_adjoint_backward(self._ctx, float(seed), _ptr(g)), "dns_adjoint_backward")
        return g

    def adjoint_step(self, uh, us, lam_uh, lam_us):
        """Adjoint of one step at state (uh, us) applied to (lam_uh, lam_us)."""
        n, nf = self.n, self.nf
        a = [np.ascontiguousarray(x, dtype=t) for x, t in
             ((uh, np.complex128), (us, np.float64), (lam_uh, np.complex128), (lam_us, np.float64))]
        assert a[0].shape == a[2].shape == (3, n, n, nf) and a[1].shape == a[3].shape == (3, n, n, n)
        o1 = np.empty((3, n, n, nf), dtype=np.complex128)
        o2 = np.empty((3, n, n, n), dtype=np.float64)
        _check(lib().dns_adjoint_step(self._ctx, *[_ptr(x) for x in a], _ptr(o1), _ptr(o2)), "dns_adjoint_step")
        return o1, o2

    def rfft3(self, real):
        real = np.ascontiguousarray(real, dtype=np.float64)
        assert real.shape == (self.n,) * 3
        out = np.empty((self.n, self.n, self.nf), dtype=np.complex128)
        _check(lib().dns_rfft3(self._ctx, _ptr(real), _ptr(out)), "dns_rfft3")
        return out

    def irfft3(self, spec):
        spec = np.ascontiguousarray(spec, dtype=np.complex128)
        assert spec.shape == (self.n, self.n, self.nf)
        out = np.empty((self.n,) * 3, dtype=np.float64)
        _check(lib().dns_irfft3(self._ctx, _ptr(spec), _ptr(out)), "dns_irfft3")
        return out


class Batch:
    """An ensemble of independent simulations (include/dns.h, dns_batch_*)."""

    def __init__(self, ns, nus, dts, precision=FP64, device=0):
        if isinstance(precision, str):
            precision = PRECISIONS[precision]
        self.ns = [int(v) for v in ns]
        m = len(self.ns)
        n_arr = np.ascontiguousarray(self.ns, dtype=np.int32)
        nu_arr = np.ascontiguousarray(np.broadcast_to(np.asarray(nus, dtype=np.float64), (m,)))
        dt_arr = np.ascontiguousarray(np.broadcast_to(np.asarray(dts, dtype=np.float64), (m,)))
        self._b = ctypes.c_void_p()
        _check(lib().dns_batch_create(ctypes.byref(self._b), m, _ptr(n_arr), _ptr(nu_arr), _ptr(dt_arr),
                                      precision, device), "dns_batch_create")

    def close(self):
        if self._b:
            lib().dns_batch_destroy(self._b)
            self._b = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __len__(self):
        return len(self.ns)

    def device_bytes(self):
        return lib().dns_batch_device_bytes(self._b)

    def load_velocity(self, m, uvw):
        uvw = np.ascontiguousarray(uvw, dtype=np.float64)
        n = self.ns[m]
        assert uvw.shape == (3, n, n, n)
        _check(lib().dns_batch_load_velocity(self._b, m, _ptr(uvw)), "dns_batch_load_velocity")

    def step(self, nsteps=1):
        _check(lib().dns_batch_step(self._b, int(nsteps)), "dns_batch_step")

    def autotune(self, steps=3):
        _check(lib().dns_batch_autotune(self._b, int(steps)), "dns_batch_autotune")
        return [lib().dns_batch_launch_shape(self._b, m) for m in range(len(self.ns))]

    def diagnostics(self, m):
        e, z = ctypes.c_double(), ctypes.c_double()
        _check(lib().dns_batch_diagnostics(self._b, m, ctypes.byref(e), ctypes.byref(z)), "dns_batch_diagnostics")
        return e.value, z.value

    def get_spectral(self, m):
        n = self.ns[m]
        out = np.empty((3, n, n, n // 2 + 1), dtype=np.complex128)
        _check(lib().dns_batch_get_spectral(self._b, m, _ptr(out)), "dns_batch_get_spectral")
        return out

    def get_stale_velocity(self, m):
        n = self.ns[m]
        out = np.empty((3, n, n, n), dtype=np.float64)
        _check(lib().dns_batch_get_stale_velocity(self._b, m, _ptr(out)), "dns_batch_get_stale_velocity")
        return out
