"""Snapshot files of the oracle command line (`dns_cpu -d`) against the
independent expectations of dump_check.py (main.c:269-330), and three
mistakes planted in a copy of the oracle's writer that the checker must
catch: a wrong Seek stride, the raw file named by tstep instead of idump,
and the geometry type spelt correctly."""
import os
import subprocess

import pytest

from conftest import ORACLE_CLI, ORACLE_DIR, tgv_field, write_raw
from dump_check import check_dump_dir

N, DT, T = 8, 0.01, 0.105  # prints steps 0 and 10, then stops at step 11


def run_dump(exe, d):
    u = tgv_field(N)
    write_raw(os.path.join(d, "in.raw"), u[:3])
    r = subprocess.run([exe, "-d", "-i", "in.raw", "-n", "0.01", "-s", str(DT), "-t", str(T)], cwd=d,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert [int(l.split()[0]) for l in r.stdout.splitlines()] == [0, 10]
    return u


def test_oracle_dump_files(tmp_path):
    u = run_dump(ORACLE_CLI, str(tmp_path))
    check_dump_dir(str(tmp_path), N, DT, 10, u)
    # the sidecar of dump 1 names 00000001.raw, which is not the file written
    assert not (tmp_path / "00000001.raw").exists()
    assert (tmp_path / "00000010.raw").exists()


MUTANTS = {
    "seek_stride": ("offset += n * n * n * (long)sizeof(double);", "offset += n * n * (long)sizeof(double);"),
    "raw_named_by_tstep": ("names[v], offset, n, n, n, idump);", "names[v], offset, n, n, n, idump * 10);"),
    "spelling": ('GeometryType=\\"ORIGIn_DXDYDZ\\"', 'GeometryType=\\"ORIGIN_DXDYDZ\\"'),
}


@pytest.mark.parametrize("name", sorted(MUTANTS))
def test_checker_catches_planted_mistake(tmp_path, name):
    old, new = MUTANTS[name]
    src = open(os.path.join(ORACLE_DIR, "dns_cpu.c")).read()
    assert src.count(old) == 1, f"mutation site of {name} not found"
    mdir = tmp_path / "src"
    mdir.mkdir()
    (mdir / "dns_cpu.c").write_text(src.replace(old, new))
    exe = str(tmp_path / "dns_cpu_mut")
    subprocess.run(["gcc", "-std=gnu99", "-O1", "-fopenmp", f"-I{ORACLE_DIR}", str(mdir / "dns_cpu.c"),
                    os.path.join(ORACLE_DIR, "orc_fft.c"), os.path.join(ORACLE_DIR, "orc_solver.c"), "-lm",
                    "-o", exe], check=True, capture_output=True)
    run = tmp_path / "run"
    run.mkdir()
    u = run_dump(exe, str(run))
    with pytest.raises(AssertionError):
        check_dump_dir(str(run), N, DT, 10, u)
