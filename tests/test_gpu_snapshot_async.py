"""Asynchronous snapshots (dns_snapshot_begin / dns_snapshot_wait) and the
command line's -d path built on them.

* The buffer of an asynchronous snapshot is bitwise the synchronous
  snapshot of the same state, although steps run between begin and wait.
* Consecutive asynchronous snapshots each hold their own state.
* `dns -d` writes raw files bitwise equal to the synchronous snapshots of a
  library run of the same steps (pressure captured at the last step before
  each line, zeros at step 0), with main.c's stdout lines; the writer thread
  reports an unwritable file as main.c does (also covered by the command
  line tests)."""
import os
import subprocess

import numpy as np
import pytest

import dns_amd
from conftest import DNS_CLI, parse_lines, random_field, write_raw

pytestmark = pytest.mark.gpu


def test_async_equals_sync_with_steps_in_between():
    n, nu, dt = 48, 0.02, 0.01
    u = random_field(n, 21)
    with dns_amd.Solver(n, nu, dt) as a, dns_amd.Solver(n, nu, dt) as b:
        for s in (a, b):
            s.load_velocity(u)
            s.set_pressure_capture(True)
            s.step(3)
        want = b.snapshot()
        a.snapshot_begin()
        a.step(7)  # runs while the copy proceeds; the snapshot is of step 3
        got = a.snapshot_wait()
        assert np.array_equal(got, want)
        b.step(7)
        a.snapshot_begin()
        got2 = a.snapshot_wait()
        assert np.array_equal(got2, b.snapshot())
        assert not np.array_equal(got2, got)
        # the stepping after the first begin is unaffected by the snapshot
        assert np.array_equal(a.get_spectral(), b.get_spectral())


def test_cli_dump_files_equal_library_snapshots(tmp_path):
    n, nu, dt = 32, 0.02, 0.01
    u = random_field(n, 22)
    raw = tmp_path / "in.raw"
    write_raw(raw, u, p=np.ones((n, n, n)))
    r = subprocess.run([DNS_CLI, "-d", "-i", str(raw), "-n", repr(nu), "-s", repr(dt), "-t", "0.295"],
                       cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    rows = parse_lines(r.stdout)
    assert [x[0] for x in rows] == [0, 10, 20, 30]
    names = sorted(os.listdir(tmp_path))
    assert names == ["00000000.raw", "00000010.raw", "00000020.raw", "00000030.raw", "a.00000000.xdmf2",
                     "a.00000001.xdmf2", "a.00000002.xdmf2", "a.00000003.xdmf2", "in.raw"]
    with dns_amd.Solver(n, nu, dt) as s:
        s.load_velocity(u)
        s.set_pressure_capture(True)
        for k in range(4):
            if k:
                s.step(10)
            want = s.snapshot()
            got = np.fromfile(tmp_path / f"{10 * k:08d}.raw").reshape(4, n, n, n)
            assert np.array_equal(got, want), k
            e, z = s.diagnostics()
            assert rows[k][2] == e and rows[k][3] == z
