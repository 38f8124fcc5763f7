"""dns_diagnostics_noalloc (the allocation-free variant of dns_diagnostics)
against dns_diagnostics, through the C ABI: the same energy and enstrophy,
bit for bit, for the spectral state of a context copied to caller-owned
device memory in the host API's x-major layout; two calls in a row, the
second on a smaller grid (the static scratch keeps the first call's
partials, which the second must not see); sizes beyond the scratch's limit
are refused before any device work."""
import ctypes

import numpy as np
import pytest

import dns_amd
from conftest import random_field

pytestmark = pytest.mark.gpu


def _noalloc(torch, spec, n):
    L = dns_amd.lib()
    f = L.dns_diagnostics_noalloc
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    dev = torch.from_numpy(np.ascontiguousarray(spec).view(np.float64)).to("cuda")
    out = torch.full((2,), float("nan"), dtype=torch.float64, device="cuda")
    stream = torch.cuda.current_stream()
    rc = f(dev.data_ptr(), n, ctypes.c_void_p(stream.cuda_stream), out.data_ptr())
    stream.synchronize()
    assert rc == 0, dns_amd.lib().dns_last_error()
    e, z = out.cpu().numpy()
    return float(e), float(z)


def test_noalloc_diagnostics_equal_context_diagnostics():
    torch = pytest.importorskip("torch")
    for n, seed in ((64, 1), (32, 2)):  # the second call on a smaller grid
        print(f"noalloc diagnostics n={n}", flush=True)
        with dns_amd.Solver(n, 0.02, 0.005) as s:
            s.load_velocity(random_field(n, seed))
            s.step(3)
            want = s.diagnostics()
            spec = s.get_spectral()
        got = _noalloc(torch, spec, n)
        assert got == want, (n, got, want)


def test_noalloc_refuses_sizes_beyond_scratch():
    L = dns_amd.lib()
    f = L.dns_diagnostics_noalloc
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    for n in (2048, 0, 7):
        # non-null dummy pointers: the size check comes before any device work
        assert f(ctypes.c_void_p(16), n, None, ctypes.c_void_p(16)) == -1
        assert b"dns_diagnostics_noalloc" in L.dns_last_error()
