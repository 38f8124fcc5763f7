"""roctx ranges: with DNS_ROCTX set the library opens the roctx library and
marks every step and RK stage it launches (rocprofv3 --marker-trace shows
them); the results are those of a run without ranges, bit for bit."""
import json
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CHILD = r"""
import json, sys
sys.path.insert(0, sys.argv[1])
import numpy as np
import dns_amd
n = 32
x = 2 * np.pi * np.arange(n) / n
u = np.zeros((3, n, n, n))
u[0] = np.sin(x)[:, None, None] * np.cos(x)[None, :, None] * np.cos(x)[None, None, :]
u[1] = -np.cos(x)[:, None, None] * np.sin(x)[None, :, None] * np.cos(x)[None, None, :]
with dns_amd.Solver(n, 0.01, 0.004, "fp64") as s:
    s.load_velocity(u)
    s.set_kernel_timing(True)   # direct launches: the ranges are pushed per step
    s.step(3)
    s.set_kernel_timing(False)
    s.step(10)                  # graph replays (ranges mark the capture)
    e, z = s.diagnostics()
print("RESULT", json.dumps([e, z]))
"""


def _run(env_extra):
    env = dict(os.environ, **env_extra)
    r = subprocess.run([sys.executable, "-c", CHILD, ROOT], env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    line = [l for l in r.stdout.splitlines() if l.startswith("RESULT")]
    return json.loads(line[0][7:])


def test_roctx_ranges_leave_results_unchanged():
    print("roctx on/off, n=32", flush=True)
    plain = _run({"DNS_ROCTX": "0"})
    marked = _run({"DNS_ROCTX": "1"})
    assert plain == marked
