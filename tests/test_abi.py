"""The C-ABI library loads, exports every function include/dns.h declares,
and fails cleanly (no crash, DNS_ENODEV / DNS_EINVAL) where no GPU is
visible.  No compute call is made here."""
import ctypes
import subprocess

import pytest

import dns_amd


def test_header_symbols_exported():
    names = dns_amd.abi_functions()
    assert len(names) >= 15
    out = subprocess.run(["nm", "-D", "--defined-only", dns_amd.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T " in line}
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    L = dns_amd.lib()
    for n in names:
        assert hasattr(L, n)


def test_strerror_table():
    L = dns_amd.lib()
    assert L.dns_strerror(0) == b"success"
    assert L.dns_strerror(-5) == b"non-finite energy or enstrophy"
    assert L.dns_strerror(-99) == b"unknown error"


def test_null_arguments_rejected():
    L = dns_amd.lib()
    assert L.dns_create(None, 64, 0.01, 0.01, 0, 0) == dns_amd.DNS_EINVAL
    assert L.dns_step(None, 1) == dns_amd.DNS_EINVAL
    e = ctypes.c_double()
    assert L.dns_diagnostics(None, ctypes.byref(e), ctypes.byref(e)) == dns_amd.DNS_EINVAL
    L.dns_destroy(None)


def _gpu_visible():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.mark.skipif(_gpu_visible(), reason="checks the no-GPU path")
def test_create_without_gpu_reports_enodev():
    with pytest.raises(dns_amd.DnsError) as ei:
        dns_amd.Solver(64, 0.01, 0.01)
    assert ei.value.code in (dns_amd.DNS_ENODEV, dns_amd.DNS_EHIP)
