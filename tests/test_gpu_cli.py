"""End-to-end drop-in checks on the GPU: the `dns` binary against the CPU
oracle binary `dns_cpu` on identical inputs (stdout lines, -d snapshot and
XDMF files), and the isotropic initial condition of `dns-ic`.

Tolerance: FP64 energy/enstrophy within 1e-12 relative of the oracle over
the whole run; snapshot fields within 1e-12 of max."""
import os
import signal
import subprocess

import numpy as np
import pytest

from conftest import DNS_CLI, DNS_IC, ORACLE_CLI, parse_lines, write_raw, tgv_field
from dump_check import check_dump_dir

pytestmark = pytest.mark.gpu


def _run(exe, args, cwd, env=None):
    e = dict(os.environ)
    e.update(env or {})
    r = subprocess.run([exe, *args], cwd=cwd, capture_output=True, text=True, env=e, timeout=600)
    assert r.returncode == 0, r.stderr
    return r


def _compare(a, b, tol=1e-12):
    assert len(a) == len(b) and len(a) > 0
    for (sa, ta, ea, za), (sb, tb, eb, zb) in zip(a, b):
        assert sa == sb and ta == tb
        assert abs(ea - eb) <= tol * abs(eb), (sa, ea, eb)
        assert abs(za - zb) <= tol * abs(zb), (sa, za, zb)


def test_config0_tgv64_1000_steps(tmp_path):
    """Taylor-Green n=64, nu=0.01, dt=0.01, 1000 steps (-t 9.995)."""
    raw = tmp_path / "tgv64.raw"
    _run(DNS_IC, ["-l", "6", "-o", str(raw)], tmp_path)
    args = ["-i", str(raw), "-n", "0.01", "-s", "0.01", "-t", "9.995"]
    print("config0: dns (GPU)", flush=True)
    g = parse_lines(_run(DNS_CLI, args, tmp_path).stdout)
    print("config0: dns_cpu (oracle)", flush=True)
    o = parse_lines(_run(ORACLE_CLI, args, tmp_path).stdout)
    assert g[0] == (0, 0.0, 0.125, 0.375)
    assert g[-1][0] == 1000
    _compare(g, o)


@pytest.mark.parametrize("precision", ["mixed", "mixed_acc64"])
def test_mixed_cli_within_budget(tmp_path, precision):
    raw = tmp_path / "tgv64.raw"
    _run(DNS_IC, ["-l", "6", "-o", str(raw)], tmp_path)
    args = ["-i", str(raw), "-n", "0.01", "-s", "0.01", "-t", "1.995"]
    a = parse_lines(_run(DNS_CLI, args, tmp_path).stdout)
    b = parse_lines(_run(DNS_CLI, args, tmp_path, {"DNS_PRECISION": precision}).stdout)
    assert [x[0] for x in a] == [x[0] for x in b]
    for (s1, _, e1, _z1), (s2, _, e2, _z2) in zip(a, b):
        assert abs(e1 - e2) / e1 < 1e-5


def test_dump_files_match_oracle(tmp_path):
    n = 32
    u = np.zeros((3, n, n, n))
    rng = np.random.default_rng(9)
    from conftest import random_field
    u[:] = random_field(n, 9)
    gdir, odir = tmp_path / "g", tmp_path / "o"
    gdir.mkdir()
    odir.mkdir()
    raw = tmp_path / "in.raw"
    write_raw(raw, u, p=rng.standard_normal((n, n, n)))
    args = ["-d", "-i", str(raw), "-n", "0.02", "-s", "0.01", "-t", "0.105"]
    g = parse_lines(_run(DNS_CLI, args, gdir).stdout)
    o = parse_lines(_run(ORACLE_CLI, args, odir).stdout)
    _compare(g, o)
    assert sorted(os.listdir(gdir)) == sorted(os.listdir(odir)) == \
        ["00000000.raw", "00000010.raw", "a.00000000.xdmf2", "a.00000001.xdmf2"]
    for f in ("a.00000000.xdmf2", "a.00000001.xdmf2"):
        assert (gdir / f).read_bytes() == (odir / f).read_bytes()
    # the sidecars and file set against main.c:269-330 directly
    check_dump_dir(str(gdir), n, 0.01, 10, u, tol=1e-14)
    for f in ("00000000.raw", "00000010.raw"):
        a = np.fromfile(gdir / f).reshape(4, n, n, n)
        b = np.fromfile(odir / f).reshape(4, n, n, n)
        assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(b))
    # step-0 snapshot: U,V,W round trip of the input, P = 0 (input P ignored)
    a = np.fromfile(gdir / "00000000.raw").reshape(4, n, n, n)
    assert np.max(np.abs(a[:3] - u)) < 1e-14 * np.max(np.abs(u))
    assert np.all(a[3] == 0)


def _iso(tmp_path, n, seed, name):
    raw = tmp_path / name
    _run(DNS_IC, ["-t", "iso", "-N", str(n), "-S", str(seed), "-o", str(raw)], tmp_path)
    return np.fromfile(raw).reshape(4, n, n, n)


def test_isotropic_initial_condition(tmp_path):
    n = 64
    a = _iso(tmp_path, n, 7, "a.raw")
    b = _iso(tmp_path, n, 7, "b.raw")
    c = _iso(tmp_path, n, 8, "c.raw")
    assert np.array_equal(a, b)
    assert not np.array_equal(a, c)
    u = a[:3]
    assert np.all(a[3] == 0)
    assert abs(0.5 * np.mean(np.sum(u * u, axis=0)) - 0.5) < 1e-12
    uh = np.fft.rfftn(u, axes=(1, 2, 3))
    k = np.fft.fftfreq(n, 1.0 / n)
    kz = np.arange(n // 2 + 1)
    div = k[:, None, None] * uh[0] + k[None, :, None] * uh[1] + kz[None, None, :] * uh[2]
    assert np.max(np.abs(div)) < 1e-11 * np.max(np.abs(uh))
    # Nyquist planes and mean are empty
    assert np.max(np.abs(uh[:, n // 2])) < 1e-12 * np.max(np.abs(uh))
    assert np.max(np.abs(uh[:, :, n // 2])) < 1e-12 * np.max(np.abs(uh))
    assert np.max(np.abs(uh[:, :, :, n // 2])) < 1e-12 * np.max(np.abs(uh))
    assert np.max(np.abs(uh[:, 0, 0, 0])) < 1e-12 * np.max(np.abs(uh))
    # shell spectrum follows k^4 exp(-2 (k/4)^2) in shape (peak at k0 = 4)
    kk = np.sqrt(k[:, None, None] ** 2 + k[None, :, None] ** 2 + kz[None, None, :] ** 2)
    w = np.where(kz == 0, 1.0, 2.0)[None, None, :]
    e = 0.5 * w * np.sum(np.abs(uh) ** 2, axis=0) / n ** 6
    shells = np.bincount(np.rint(kk).astype(int).ravel(), weights=e.ravel())[:20]
    model = np.arange(20.0) ** 4 * np.exp(-2 * (np.arange(20.0) / 4) ** 2)
    assert 3 <= int(np.argmax(shells)) <= 4
    corr = np.corrcoef(shells[1:16], model[1:16])[0, 1]
    assert corr > 0.97, corr


def test_isotropic_run_matches_oracle(tmp_path, oracle):
    import dns_amd
    n, nu, dt = 64, 1.0 / 800, 2e-3
    u = _iso(tmp_path, n, 20261202, "iso.raw")[:3]
    g = dns_amd.Solver(n, nu, dt)
    o = oracle.solver(n, nu, dt)
    g.load_velocity(u)
    o.load_velocity(u)
    for _ in range(3):
        g.step(10)
        o.step(10)
        ge, gz = g.diagnostics()
        oe, oz = o.diagnostics()
        assert abs(ge - oe) / oe < 1e-12 and abs(gz - oz) / oz < 1e-12
    g.close()


def _run_any(exe, args, cwd):
    return subprocess.run([exe, *args], cwd=cwd, capture_output=True, text=True, timeout=600)


def test_nonfinite_input_raises_sigfpe(tmp_path):
    """An Inf in the input: the reference (FP traps, main.c:49-50) and the
    oracle die with SIGFPE before the first line; dns detects the
    non-finite energy of step 0 and raises SIGFPE itself."""
    u = tgv_field(16)
    u[0, 3, 4, 5] = np.inf
    p = tmp_path / "inf.raw"
    write_raw(p, u[:3])
    args = ["-i", str(p), "-n", "0.001", "-s", "0.01", "-t", "1"]
    print("nonfinite input: dns and dns_cpu", flush=True)
    g, o = _run_any(DNS_CLI, args, tmp_path), _run_any(ORACLE_CLI, args, tmp_path)
    assert o.returncode == g.returncode == -signal.SIGFPE
    assert g.stdout == o.stdout == ""


@pytest.mark.parametrize("precision", ["fp64", "mixed", "mixed_acc64"])
def test_unstable_run_raises_sigfpe(tmp_path, precision):
    """dt = 5 at n=16 overflows by step 3: the same stdout (the line of
    step 0) and the same signal as the oracle, whose traps fire at the
    first overflow; dns finds the non-finite state at the next line."""
    p = tmp_path / "tgv.raw"
    write_raw(p, tgv_field(16)[:3])
    args = ["-i", str(p), "-n", "0.001", "-s", "5", "-t", "1000"]
    print(f"unstable run: dns ({precision}) and dns_cpu", flush=True)
    env = dict(os.environ, DNS_PRECISION=precision)
    g = subprocess.run([DNS_CLI, *args], cwd=tmp_path, capture_output=True, text=True, timeout=600, env=env)
    o = _run_any(ORACLE_CLI, args, tmp_path)
    assert o.returncode == g.returncode == -signal.SIGFPE
    assert g.stdout == o.stdout
    assert len(g.stdout.splitlines()) == 1


@pytest.mark.skipif(hasattr(os, "geteuid") and os.geteuid() == 0, reason="root may write a read-only directory")
def test_unwritable_dump_directory(tmp_path):
    """-d in a directory the process cannot write: dns reports the file it
    could not open and exits 1 (the reference does not check this fopen)."""
    p = tmp_path / "tgv.raw"
    write_raw(p, tgv_field(16)[:3])
    ro = tmp_path / "ro"
    ro.mkdir()
    os.chmod(ro, 0o555)
    try:
        print("unwritable dump directory: dns", flush=True)
        r = _run_any(DNS_CLI, ["-d", "-i", str(p), "-n", "0.01", "-s", "0.01", "-t", "0.05"], ro)
    finally:
        os.chmod(ro, 0o755)
    assert r.returncode == 1
    assert r.stderr == "dns: error: fail to open '00000000.raw'\n"
    assert r.stdout.splitlines()[0].split()[0] == "0"


def test_verbose_lines(tmp_path):
    """-v: "dns: device: <ordinal>" (in place of the reference's thread
    count) and "dns: n = <n>" (main.c:159-160,172-173), then the step loop's
    wall time; stdout is unchanged by -v."""
    p = tmp_path / "tgv.raw"
    write_raw(p, tgv_field(16)[:3])
    args = ["-i", str(p), "-n", "0.01", "-s", "0.01", "-t", "0.195"]
    print("verbose: dns -v", flush=True)
    v = _run(DNS_CLI, ["-v", *args], tmp_path)
    q = _run(DNS_CLI, args, tmp_path)
    assert v.stdout == q.stdout and len(v.stdout.splitlines()) == 3
    lines = v.stderr.splitlines()
    assert lines[0] == "dns: device: 0" and lines[1] == "dns: n = 16"
    head, _, rest = lines[2].partition(" steps in ")
    assert head == "dns: 20" and rest.endswith(" steps/s)"), lines[2]
    assert q.stderr == ""


def test_unsupported_grid_size(tmp_path):
    """Grids are even n = 2^a 3^b, 8..1024: a 100^3 file is read (its size
    passes main.c:168's check) and then refused by the library."""
    p = tmp_path / "n100.raw"
    np.zeros(4 * 100 ** 3).tofile(p)
    print("unsupported grid n=100: dns", flush=True)
    r = _run_any(DNS_CLI, ["-i", str(p), "-n", "0.01", "-s", "0.01", "-t", "0.1"], tmp_path)
    assert r.returncode == 1
    assert r.stderr == "dns: error: invalid argument (unsupported grid size n=100)\n"
    assert r.stdout == ""


def test_mixed_dump_files_close_to_fp64(tmp_path):
    """-d under DNS_PRECISION=mixed: the same files as FP64 (names, XDMF
    bytes), raw velocity within 1e-5 and pressure within 1e-4 of the FP64
    dump's max (the pressure comes from the float projection)."""
    n = 32
    from conftest import random_field
    raw = tmp_path / "in.raw"
    write_raw(raw, random_field(n, 19))
    gdir, mdir = tmp_path / "g", tmp_path / "m"
    gdir.mkdir()
    mdir.mkdir()
    args = ["-d", "-i", str(raw), "-n", "0.02", "-s", "0.01", "-t", "0.105"]
    print("dumps: dns fp64 and mixed", flush=True)
    _run(DNS_CLI, args, gdir)
    _run(DNS_CLI, args, mdir, env={"DNS_PRECISION": "mixed"})
    assert sorted(os.listdir(gdir)) == sorted(os.listdir(mdir))
    for f in ("a.00000000.xdmf2", "a.00000001.xdmf2"):
        assert (gdir / f).read_bytes() == (mdir / f).read_bytes()
    for f in ("00000000.raw", "00000010.raw"):
        a = np.fromfile(gdir / f).reshape(4, n, n, n)
        b = np.fromfile(mdir / f).reshape(4, n, n, n)
        assert np.max(np.abs(b[:3] - a[:3])) <= 1e-5 * np.max(np.abs(a[:3])), f
        assert np.max(np.abs(b[3] - a[3])) <= 1e-4 * max(np.max(np.abs(a[3])), 1e-300) or not np.any(a[3]), f
