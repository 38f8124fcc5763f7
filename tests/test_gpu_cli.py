"""End-to-end drop-in checks on the GPU: the `dns` binary against the CPU
oracle binary `dns_cpu` on identical inputs (stdout lines, -d snapshot and
XDMF files), and the isotropic initial condition of `dns-ic`.

Tolerance: FP64 energy/enstrophy within 1e-12 relative of the oracle over
the whole run; snapshot fields within 1e-12 of max."""
import os
import subprocess

import numpy as np
import pytest

from conftest import DNS_CLI, DNS_IC, ORACLE_CLI, parse_lines, write_raw, tgv_field

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


def test_mixed_cli_within_budget(tmp_path):
    raw = tmp_path / "tgv64.raw"
    _run(DNS_IC, ["-l", "6", "-o", str(raw)], tmp_path)
    args = ["-i", str(raw), "-n", "0.01", "-s", "0.01", "-t", "1.995"]
    a = parse_lines(_run(DNS_CLI, args, tmp_path).stdout)
    b = parse_lines(_run(DNS_CLI, args, tmp_path, {"DNS_PRECISION": "mixed"}).stdout)
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
