"""The adjoint workload at its full size: gradient of the final enstrophy
with respect to the initial velocity for the n=128 isotropic field of
`dns-ic` (seed 20261202, k0=4, kinetic energy 0.5), nu=1/800, dt=2e-3, 200
RK4 steps, checkpoint every 20 steps, FP64.

1. Dot-product test of one step at n=128 against the numpy tangent-linear
   model, at the workload's initial state and parameters (<= 1e-12).
2. <grad, d> against central differences (eps=1e-6) of the GPU FP64 forward
   along 8 seeded divergence-free directions (`dns-ic` isotropic fields,
   seeds 777..784), within 1e-6 relative.
3. Two of those directions against central differences of the CPU oracle
   (`dns_cpu`, the OpenMP C restatement of main.c), within 1e-6 relative;
   the perturbed runs' final enstrophy of the GPU and of the oracle agree
   within 1e-12 relative (the forward over 200 steps at this workload).

The oracle runs (4 x 200 steps at 128^3) go to 4 processes of 4 OpenMP
threads each, started first, while the GPU work proceeds."""
import concurrent.futures as cf
import os
import subprocess

import numpy as np
import pytest

import dns_amd
from conftest import DNS_IC, ORACLE_CLI, parse_lines, write_raw
from numpy_step import inner, r2c, step_tlm

pytestmark = pytest.mark.gpu

N, NU, DT, STEPS, EVERY, EPS = 128, 1.0 / 800, 2e-3, 200, 20, 1e-6
# -t such that the run stops right after step 200 (t accumulates dt)
T_END = "0.399"
CPU_DIRS = 2
GPU_DIRS = 8


def _iso(tmp_path, seed):
    raw = tmp_path / f"iso{seed}.raw"
    subprocess.run([DNS_IC, "-t", "iso", "-N", str(N), "-S", str(seed), "-o", str(raw)], check=True,
                   capture_output=True, timeout=120)
    return np.fromfile(raw, dtype="<f8").reshape(4, N, N, N)[:3].copy()


def _oracle_final_enstrophy(tmp_path, name):
    env = dict(os.environ, OMP_NUM_THREADS="4")
    r = subprocess.run([ORACLE_CLI, "-i", name, "-n", repr(NU), "-s", repr(DT), "-t", T_END], cwd=tmp_path,
                       capture_output=True, text=True, env=env, timeout=900)
    assert r.returncode == 0, r.stderr
    rows = parse_lines(r.stdout)
    assert rows[-1][0] == STEPS, rows[-1]
    return rows[-1][3]


def test_adjoint_config_n128(tmp_path):
    u0 = _iso(tmp_path, 20261202)
    dirs = [_iso(tmp_path, 777 + k) for k in range(GPU_DIRS)]
    # oracle runs first: they take longest
    jobs = {}
    pool = cf.ThreadPoolExecutor(2 * CPU_DIRS)
    for k in range(CPU_DIRS):
        for sgn in (1, -1):
            name = f"pert{k}{'p' if sgn > 0 else 'm'}.raw"
            write_raw(tmp_path / name, u0 + sgn * EPS * dirs[k])
            jobs[(k, sgn)] = pool.submit(_oracle_final_enstrophy, tmp_path, name)
    try:
        print("adjoint config: forward/backward n=128, 200 steps, every 20", flush=True)
        with dns_amd.Solver(N, NU, DT, "fp64") as s:
            J = s.adjoint_forward(u0, STEPS, EVERY)
            g = s.adjoint_backward()
            gpu_pm = {}
            for k in range(GPU_DIRS):
                for sgn in (1, -1):
                    s.load_velocity(u0 + sgn * EPS * dirs[k])
                    s.step(STEPS)
                    gpu_pm[(k, sgn)] = s.diagnostics()[1]
            s.load_velocity(u0)
            s.step(STEPS)
            assert s.diagnostics()[1] == J
        assert np.isfinite(g).all() and J > 0
        for k in range(GPU_DIRS):
            fd = (gpu_pm[(k, 1)] - gpu_pm[(k, -1)]) / (2 * EPS)
            ad = float(np.sum(g * dirs[k]))
            print(f"direction {k}: adjoint {ad:.12e} GPU central difference {fd:.12e}", flush=True)
            assert abs(ad - fd) <= 1e-6 * abs(fd), (k, ad, fd)
        print("adjoint config: waiting for the oracle runs", flush=True)
        for k in range(CPU_DIRS):
            cpu = {sgn: jobs[(k, sgn)].result() for sgn in (1, -1)}
            for sgn in (1, -1):
                assert abs(gpu_pm[(k, sgn)] - cpu[sgn]) <= 1e-12 * abs(cpu[sgn]), (k, sgn, gpu_pm[(k, sgn)], cpu[sgn])
            fd = (cpu[1] - cpu[-1]) / (2 * EPS)
            ad = float(np.sum(g * dirs[k]))
            print(f"direction {k}: adjoint {ad:.12e} oracle central difference {fd:.12e}", flush=True)
            assert abs(ad - fd) <= 1e-6 * abs(fd), (k, ad, fd)
    finally:
        pool.shutdown(wait=True)


def test_adjoint_step_dot_product_n128(tmp_path):
    """One step at the workload's initial state: <TLM d, lam> = <d, ADJ lam>."""
    u0 = _iso(tmp_path, 20261202)
    rng = np.random.default_rng(128)
    Uh = r2c(u0)
    us = u0
    dUh = r2c(_iso(tmp_path, 901))
    dus = _iso(tmp_path, 902)
    lUh = rng.standard_normal(Uh.shape) + 1j * rng.standard_normal(Uh.shape)
    lus = rng.standard_normal(us.shape)
    _, _, tUh, tus = step_tlm(Uh, us, dUh, dus, NU, DT)
    lhs = inner(tUh, tus, lUh, lus)
    print("adjoint step n=128", flush=True)
    with dns_amd.Solver(N, NU, DT) as s:
        aUh, aus = s.adjoint_step(Uh, us, lUh, lus)
    rhs = inner(dUh, dus, aUh, aus)
    assert abs(lhs - rhs) <= 1e-12 * max(abs(lhs), abs(rhs)), (lhs, rhs)
