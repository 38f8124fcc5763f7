"""Batched ensembles (dns_batch_*): every member must reproduce its own
single-context run (same n, nu, dt, precision) bit for bit -- the batch
only adds a grid dimension to the same kernels and replays them from a
captured HIP graph."""
import numpy as np
import pytest

import dns_amd
from conftest import random_field, tgv_field

pytestmark = pytest.mark.gpu


def _ic(n, m):
    return tgv_field(n)[:3] if m % 2 == 0 else random_field(n, seed=100 + m)


@pytest.mark.parametrize("precision", ["fp64", "mixed", "mixed_acc64"])
def test_members_match_single_runs(precision):
    ns = [32, 48, 32, 64, 48, 32]
    nus = [0.01, 0.02, 1.0 / 400, 0.005, 0.0125, 1.0 / 1600]
    dts = [0.005, 0.004, 0.005, 0.003, 0.004, 0.002]
    print(f"batch {precision} members={len(ns)}", flush=True)
    with dns_amd.Batch(ns, nus, dts, precision) as b:
        for m, n in enumerate(ns):
            b.load_velocity(m, _ic(n, m))
        b.step(3)
        b.step(4)
        for m, n in enumerate(ns):
            with dns_amd.Solver(n, nus[m], dts[m], precision) as s:
                s.load_velocity(_ic(n, m))
                s.step(7)
                assert np.array_equal(b.get_spectral(m), s.get_spectral()), m
                assert np.array_equal(b.get_stale_velocity(m), s.get_stale_velocity()), m
                assert b.diagnostics(m) == s.diagnostics(), m


def test_ensemble_config_shape():
    """The 24-member ensemble shape: n in {64,96,128,192} x Re in
    {100,...,3000}, half Taylor-Green, half random; a few steps, then two
    members are rerun alone."""
    res = [100, 200, 400, 800, 1600, 3000]
    ns, nus = [], []
    for n in (64, 96, 128, 192):
        for r in res:
            ns.append(n)
            nus.append(1.0 / r)
    dt = 1e-3
    print("ensemble 24 members", flush=True)
    with dns_amd.Batch(ns, nus, dt, "mixed") as b:
        for m, n in enumerate(ns):
            b.load_velocity(m, _ic(n, m))
        b.step(5)
        for m in (3, 22):
            with dns_amd.Solver(ns[m], nus[m], dt, "mixed") as s:
                s.load_velocity(_ic(ns[m], m))
                s.step(5)
                assert np.array_equal(b.get_spectral(m), s.get_spectral())
        e = [b.diagnostics(m)[0] for m in range(len(ns))]
        assert all(np.isfinite(e))


def test_launch_shapes_bitwise_identical():
    n = 64
    u = random_field(n, 31)
    ref = None
    with dns_amd.Solver(n, 0.01, 0.004, "mixed") as s:
        assert s.launch_shapes() == 4  # grids up to 512
        for shape in range(4):
            s.set_launch_shape(shape)
            s.load_velocity(u)
            s.step(4)
            got = s.get_spectral()
            if ref is None:
                ref = got
            assert np.array_equal(got, ref), shape


def test_graph_replay_equals_direct_launches():
    n = 48
    u = random_field(n, 32)
    with dns_amd.Solver(n, 0.01, 0.004) as a, dns_amd.Solver(n, 0.01, 0.004) as b:
        a.load_velocity(u)
        b.load_velocity(u)
        b.set_kernel_timing(True)  # timing forces direct launches
        for _ in range(3):
            a.step(3)
            b.step(3)
            assert np.array_equal(a.get_spectral(), b.get_spectral())
        assert sum(cnt for _, cnt in b.kernel_timing().values()) > 0


def test_autotuned_batch_still_matches_single_runs():
    ns = [64, 96, 64, 128]
    nus = [0.01, 0.005, 0.0025, 0.002]
    with dns_amd.Batch(ns, nus, 2e-3, "mixed") as b:
        for m, n in enumerate(ns):
            b.load_velocity(m, _ic(n, m))
        shapes = b.autotune(2)
        assert all(0 <= v < 4 for v in shapes)
        b.step(6)
        for m in (0, 1, 3):
            with dns_amd.Solver(ns[m], nus[m], 2e-3, "mixed") as s:
                s.load_velocity(_ic(ns[m], m))
                s.step(6)
                assert np.array_equal(b.get_spectral(m), s.get_spectral())
