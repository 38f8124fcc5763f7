"""Command-line boundary of the drop-in (main.c:58-173): the GPU `dns`
binary and the CPU oracle `dns_cpu` must give the reference's messages and
exit codes for every argument/file error, without touching a GPU.  Also
checks the Taylor-Green generator `dns-ic` against tgv.py's closed form."""
import os
import signal
import subprocess

import numpy as np
import pytest

from conftest import DNS_CLI, DNS_IC, ORACLE_CLI, tgv_field, write_raw

CASES = [
    ([], "dns: error: -t is not set or invalid\n"),
    (["-t", "1"], "dns: error: -n is not set or invalid\n"),
    (["-t", "1", "-n", "0.01"], "dns: error: -s is not set or invalid\n"),
    (["-t", "1", "-n", "0.01", "-s", "0.1"], "dns: error: -i is not set\n"),
    (["-i"], "dns: error: -i needs an argument\n"),
    (["-n"], "dns: error: -n needs an argument\n"),
    (["-s"], "dns: error: -s needs an argument\n"),
    (["-t"], "dns: error: -t needs an argument\n"),
    (["-n", "0.0x1"], "dns: error: '0.0x1' is not a number\n"),
    (["-t", "abc"], "dns: error: 'abc' is not a number\n"),
    (["-q"], "dns: error: unknown option '-q'\n"),
    (["-t", "1", "-n", "0.01", "-s", "0.1", "-i", "/nonexistent/x.raw"],
     "dns: error: fail to open '/nonexistent/x.raw'\n"),
]


@pytest.mark.parametrize("exe", [DNS_CLI, ORACLE_CLI], ids=["dns", "dns_cpu"])
@pytest.mark.parametrize("args,msg", CASES)
def test_argument_errors(exe, args, msg):
    r = subprocess.run([exe, *args], capture_output=True, text=True)
    assert r.returncode == 1
    assert r.stderr == msg
    assert r.stdout == ""


def test_unknown_precision_is_refused(tmp_path):
    """DNS_PRECISION (fp64, mixed, mixed_acc64) is checked before any device
    call; any other value ends the run with an error line."""
    p = tmp_path / "tgv.raw"
    write_raw(p, tgv_field(8)[:3])
    env = dict(os.environ, DNS_PRECISION="half")
    r = subprocess.run([DNS_CLI, "-t", "1", "-n", "0.01", "-s", "0.1", "-i", str(p)], capture_output=True,
                       text=True, env=env)
    assert r.returncode == 1
    assert r.stderr == "dns: error: DNS_PRECISION must be fp64, mixed or mixed_acc64\n"
    assert r.stdout == ""


@pytest.mark.parametrize("exe", [DNS_CLI, ORACLE_CLI], ids=["dns", "dns_cpu"])
def test_wrong_file_size(tmp_path, exe):
    p = tmp_path / "bad.raw"
    np.zeros(4 * 10 ** 3 + 1).tofile(p)
    r = subprocess.run([exe, "-t", "1", "-n", "0.01", "-s", "0.1", "-i", str(p)], capture_output=True, text=True)
    assert r.returncode == 1
    assert r.stderr == f"dns: error: wrong file '{p}'\n"


@pytest.mark.parametrize("exe", [DNS_CLI, ORACLE_CLI], ids=["dns", "dns_cpu"])
def test_help(exe):
    r = subprocess.run([exe, "-h"], capture_output=True, text=True)
    assert r.returncode == 1
    assert r.stderr.startswith("Usage: dns [-v] [-d] -i <input.raw> -n <viscosity> -t <end time> -s <time step>\n")
    assert "  dns -i tgv.raw -n 0.01 -t 1.0 -s 0.001 -v\n" in r.stderr


def test_only_second_char_of_flag_counts():
    # main.c:59 switches on argv[k][1]: "-terse" is "-t"
    r = subprocess.run([DNS_CLI, "-terse"], capture_output=True, text=True)
    assert r.stderr == "dns: error: -t needs an argument\n"


@pytest.mark.parametrize("n", [8, 16, 24])
def test_dns_ic_taylor_green(tmp_path, n):
    p = tmp_path / "tgv.raw"
    subprocess.run([DNS_IC, "-N", str(n), "-o", str(p)], check=True, capture_output=True)
    a = np.fromfile(p).reshape(4, n, n, n)
    ref = tgv_field(n)
    assert np.max(np.abs(a - ref)) < 4e-16
    assert np.all(a[2:] == 0)


def test_dns_ic_level(tmp_path):
    p = tmp_path / "tgv.raw"
    r = subprocess.run([DNS_IC, "-l", "3", "-o", str(p)], check=True, capture_output=True, text=True)
    assert os.path.getsize(p) == 4 * 8 ** 3 * 8
    assert "n=8" in r.stderr


def test_oracle_traps_nonfinite_input(tmp_path):
    """main.c:49-50 enables FP traps: an Inf in the input makes the initial
    r2c compute inf - inf, so the process dies with SIGFPE before its first
    line.  (The GPU binary reproduces this from its non-finite check;
    tests/test_gpu_cli.py compares the two.)"""
    u = tgv_field(16)
    u[0, 3, 4, 5] = np.inf
    p = tmp_path / "inf.raw"
    write_raw(p, u[:3])
    r = subprocess.run([ORACLE_CLI, "-i", str(p), "-n", "0.001", "-s", "0.01", "-t", "1"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == -signal.SIGFPE
    assert r.stdout == ""


def test_oracle_traps_overflow(tmp_path):
    """A violently unstable time step overflows within a few steps: the
    line of step 0 is printed, then SIGFPE."""
    p = tmp_path / "tgv.raw"
    write_raw(p, tgv_field(16)[:3])
    r = subprocess.run([ORACLE_CLI, "-i", str(p), "-n", "0.001", "-s", "5", "-t", "1000"],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == -signal.SIGFPE
    assert r.stdout == "         0  0.0000000000000000e+00  1.2500000000000000e-01  3.7500000000000000e-01\n"
