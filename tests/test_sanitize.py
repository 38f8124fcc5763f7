"""Host C under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only).

Builds, with gcc, instrumented copies of the oracle command line
(oracle/dns_cpu.c + orc_fft.c + orc_solver.c), of the drop-in `dns`
(dns_amd/cli/dns.c, linked against libdns_hip.so) and of `dns-ic`
(dns_amd/cli/dns_ic.c), then runs what needs no GPU: full oracle runs with
snapshots at n = 8, 12 and 16 (radix 2, 3 and 4 paths of its FFT), every
argument error of both command lines, and the Taylor-Green generator.  Any
sanitizer report fails the test (halt_on_error, non-zero exit)."""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, ORACLE_DIR, random_field, tgv_field, write_raw
from test_cli import CASES

SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]
ENV = {"ASAN_OPTIONS": "halt_on_error=1:detect_leaks=1:abort_on_error=0",
       "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1", "OMP_NUM_THREADS": "4"}
PKG = os.path.join(ROOT, "dns_amd")
INC = os.path.join(ROOT, "include")


@pytest.fixture(scope="module")
def bins(tmp_path_factory):
    d = tmp_path_factory.mktemp("san")
    out = {}
    cc = ["gcc", "-std=gnu99", *SAN]
    subprocess.run([*cc, "-fopenmp", f"-I{ORACLE_DIR}", os.path.join(ORACLE_DIR, "dns_cpu.c"),
                    os.path.join(ORACLE_DIR, "orc_fft.c"), os.path.join(ORACLE_DIR, "orc_solver.c"), "-lm",
                    "-o", str(d / "dns_cpu")], check=True, capture_output=True)
    out["dns_cpu"] = str(d / "dns_cpu")
    lib = os.path.join(PKG, "libdns_hip.so")
    if os.path.exists(lib):
        for name, src in (("dns", "dns.c"), ("dns-ic", "dns_ic.c")):
            subprocess.run([*cc, "-pthread", f"-I{INC}", os.path.join(PKG, "cli", src), f"-L{PKG}", "-ldns_hip",
                            f"-Wl,-rpath,{PKG}", "-lm", "-o", str(d / name)], check=True, capture_output=True)
            out[name] = str(d / name)
    return out


def _run(exe, args, cwd):
    env = dict(os.environ)
    env.update(ENV)
    return subprocess.run([exe, *args], cwd=cwd, capture_output=True, text=True, env=env, timeout=300)


def _clean(r):
    assert "runtime error" not in r.stderr, r.stderr
    assert "Sanitizer" not in r.stderr, r.stderr


@pytest.mark.parametrize("n", [8, 12, 16])
def test_oracle_cli_run_with_dumps(tmp_path, bins, n):
    u = random_field(n, 5 + n)
    write_raw(tmp_path / "in.raw", u, p=np.ones((n, n, n)))
    r = _run(bins["dns_cpu"], ["-v", "-d", "-i", "in.raw", "-n", "0.02", "-s", "0.01", "-t", "0.205"], str(tmp_path))
    _clean(r)
    assert r.returncode == 0
    assert [int(l.split()[0]) for l in r.stdout.splitlines()] == [0, 10, 20]
    assert (tmp_path / "00000020.raw").stat().st_size == 4 * 8 * n ** 3


@pytest.mark.parametrize("args,msg", CASES)
def test_argument_errors(tmp_path, bins, args, msg):
    for exe in ("dns_cpu", "dns"):
        if exe not in bins:
            pytest.skip("libdns_hip.so is not built")
        r = _run(bins[exe], args, str(tmp_path))
        _clean(r)
        assert r.returncode == 1 and r.stderr == msg


def test_wrong_size_and_help(tmp_path, bins):
    np.zeros(4 * 10 ** 3 + 1).tofile(tmp_path / "bad.raw")
    for exe in ("dns_cpu", "dns"):
        if exe not in bins:
            pytest.skip("libdns_hip.so is not built")
        r = _run(bins[exe], ["-t", "1", "-n", "0.01", "-s", "0.1", "-i", "bad.raw"], str(tmp_path))
        _clean(r)
        assert r.returncode == 1 and r.stderr == "dns: error: wrong file 'bad.raw'\n"
        r = _run(bins[exe], ["-h"], str(tmp_path))
        _clean(r)
        assert r.returncode == 1 and r.stderr.startswith("Usage: dns")


def test_dns_ic_taylor_green(tmp_path, bins):
    if "dns-ic" not in bins:
        pytest.skip("libdns_hip.so is not built")
    for n in (8, 12):
        r = _run(bins["dns-ic"], ["-N", str(n), "-o", f"tgv{n}.raw"], str(tmp_path))
        _clean(r)
        assert r.returncode == 0
        a = np.fromfile(tmp_path / f"tgv{n}.raw").reshape(4, n, n, n)
        assert np.max(np.abs(a - tgv_field(n))) < 4e-16
    r = _run(bins["dns-ic"], ["-N", "0", "-o", "x.raw"], str(tmp_path))
    _clean(r)
    assert r.returncode != 0
