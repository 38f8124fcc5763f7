"""Dissipation-rate validation (dns_amd.dissipation): eps = 2 Z / Re of the
n=256 Taylor-Green runs against the digitised Brachet et al. points the
reference plots them over (tgv.gp:9-11, img/ref.txt).

The curves here come from the reference's own stdout files (excerpted in
tests/golden/tgv_trajectories.json); the GPU runs reproduce those files to
~1e-15 (bench.py --validate applies the same match to its own curves)."""
import json
import os

import numpy as np
import pytest

from dns_amd import dissipation as dis

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
RES = (100, 200, 400, 800, 1600, 3000)


def _golden_curves(factor=2.0, dt_scale=1.0, relabel=None):
    with open(os.path.join(GOLDEN, "tgv_trajectories.json")) as f:
        gold = json.load(f)
    curves = {}
    for r in RES:
        rows = gold[f"256/{r}"]["rows"]
        t = np.array([float(x[1]) for x in rows]) * dt_scale
        z = np.array([float(x[3]) for x in rows])
        re_used = relabel.get(r, r) if relabel else r
        curves[r] = (t, factor / 2.0 * dis.dissipation(z, re_used))
    return curves


@pytest.fixture(scope="module")
def points():
    return dis.load_points(os.path.join(GOLDEN, "brachet_dissipation.json"))


def test_points_fixture(points):
    assert points.shape == (756, 2)
    assert np.all(np.diff(points[:, 0]) >= 0)
    assert 0 < points[:, 0].min() and points[:, 0].max() < 10
    assert 0 < points[:, 1].min() and points[:, 1].max() < dis.FULL_SCALE


def test_read_stdout_format():
    # main.c:267: "% 10ld % .16e % .16Le % .16Le"
    lines = ["% 10ld % .16e % .16e % .16e\n" % (s, s * 0.0025, 0.125 - 1e-6 * s, 0.375 + 1e-5 * s)
             for s in range(0, 41, 10)]
    step, t, e, z = dis.read_stdout(lines + ["\n"])
    assert step.tolist() == [0, 10, 20, 30, 40]
    assert t[2] == 0.05 and e[0] == 0.125 and z[1] == 0.375 + 1e-4


def test_dissipation_closed_form():
    # Taylor-Green at t=0: Z = 0.375 (main.c:257-266 on the TGV field), so
    # eps(0) = 0.75 / Re
    assert dis.dissipation([0.375], 100)[0] == 0.0075


def test_golden_runs_match_brachet(points):
    res = dis.match(points, _golden_curves())
    s = res["summary"]
    assert s["points"] == 756
    assert s["max_frac_full_scale"] < 0.025, s      # every point within 2.5% of full scale
    assert s["median"] < 1e-4, s
    assert min(s["per_curve"].values()) >= 80, s    # the points spread over all six curves


@pytest.mark.parametrize("mistake", ["dropped factor 2", "wrong time step", "swapped Re"])
def test_planted_mistakes_fail(points, mistake):
    if mistake == "dropped factor 2":
        curves = _golden_curves(factor=1.0)
    elif mistake == "wrong time step":
        curves = _golden_curves(dt_scale=2.0)
    else:
        curves = _golden_curves(relabel={100: 3000, 3000: 100, 200: 1600, 1600: 200})
    s = dis.match(points, curves)["summary"]
    assert s["max_frac_full_scale"] > 0.05 or min(s["per_curve"].values()) < 80, s


def test_match_rejects_bad_curves(points):
    with pytest.raises(ValueError):
        dis.match(points, {})
    with pytest.raises(ValueError):
        dis.match(points, {100: (np.array([1.0, 0.5]), np.array([0.1, 0.2]))})


def test_cli(tmp_path, capsys, points):
    with open(os.path.join(GOLDEN, "tgv_trajectories.json")) as f:
        gold = json.load(f)
    args = ["--points", os.path.join(GOLDEN, "brachet_dissipation.json")]
    for r in RES:
        p = tmp_path / f"{r:04d}"
        with open(p, "w") as f:
            for s, t, e, z in gold[f"256/{r}"]["rows"]:
                f.write("% 10ld % .16e % .16e % .16e\n" % (int(s), float(t), float(e), float(z)))
        args.append(f"{r}={p}")
    assert dis.main(args) == 0
    out = json.loads(capsys.readouterr().out)
    assert out == dis.match(points, _golden_curves())["summary"]
