"""The reference's Taylor-Green sweep at n=128 on the GPU (run.sh:1-10: six
runs, Re = 100..3000, nu = 1/Re as "%.16e", dt = 0.0025, t = 10, i.e. 4000
RK4 steps), executed as one FP64 batch.

* The reference's own stdout files (directory 0128, all 401 lines per file
  in tests/golden/tgv_trajectories.json): energy and enstrophy within 1e-12
  relative on every line through step 400 (t = 1), for every Re; the step
  where each run first leaves 1e-12 is printed.  Beyond that the 0128 files
  are not comparable: they also leave the reference's own 0256 files, which
  agree with them to the last printed digit through step 300 (2.7e-14 apart
  at step 400, 2e-5 in energy at step 4000 for Re = 100), while this
  solver's n=128 run -- GPU and CPU oracle alike -- prints the 0256 files'
  step-400 line exactly.  Oracle variants with a dealias cutoff anywhere
  from |k| < 32 to |k| < 43.3 (box or sphere) or with denormals flushed to
  zero all print the same step-400 line as the unmodified oracle, so the
  departure is neither a truncation nor an underflow effect of the 128^3
  grid; its cause in the 0128 runs is not known.
* The dissipation rate eps = 2 Z / Re of the six runs (tgv.gp:9-11) matched
  to the digitised Brachet et al. points (img/ref.txt): every curve gets
  points, and the largest distance stays under 10% of the plot's 0.02 full
  scale (Re 1600 and 3000 are under-resolved on 128^3; the reference plots
  its 0256 runs, which `bench.py --validate` reproduces within 4.7e-15 and
  which match the points within 2.1%)."""
import os

import numpy as np
import pytest

import dns_amd
from conftest import ROOT, load_golden, tgv_field
from dns_amd import dissipation as dis

pytestmark = pytest.mark.gpu

RES = (100, 200, 400, 800, 1600, 3000)
N, DT, STEPS = 128, 0.0025, 4000
WINDOW = 400  # steps over which the reference's n=128 files are comparable


def test_tgv_sweep_n128_matches_reference_files():
    gold = load_golden()
    keys = [f"{N}/{r}" for r in RES]
    nus = [float(gold[k]["nu"]) for k in keys]
    rows = [{int(x[0]): (float(x[2]), float(x[3])) for x in gold[k]["rows"]} for k in keys]
    assert all(len(r) == STEPS // 10 + 1 for r in rows)
    u = tgv_field(N)[:3]
    dev = {r: [0.0, 0.0, None] for r in RES}  # max rel dev E, Z, first step above 1e-12
    curves = {r: ([], []) for r in RES}
    with dns_amd.Batch([N] * len(RES), nus, DT, "fp64") as b:
        for m in range(len(RES)):
            b.load_velocity(m, u)
        step = 0
        while True:
            for m, r in enumerate(RES):
                e, z = b.diagnostics(m)
                ge, gz = rows[m][step]
                de, dz = abs(e - ge) / ge, abs(z - gz) / gz
                d = dev[r]
                d[0], d[1] = max(d[0], de), max(d[1], dz)
                if d[2] is None and max(de, dz) > 1e-12:
                    d[2] = step
                # t as the reference accumulates it (main.c:421): the step count times dt
                curves[r][0].append(step * DT)
                curves[r][1].append(2 * z / r)
            if step == STEPS:
                break
            b.step(10)
            step += 10
            if step % 1000 == 0:
                print(f"sweep n=128: step {step}", flush=True)
    for r in RES:
        print(f"Re {r}: max rel dev E {dev[r][0]:.2e} Z {dev[r][1]:.2e}, first step above 1e-12: {dev[r][2]}",
              flush=True)
    # every line through step 400 (t = 1) within 1e-12 of the reference's files
    assert all(d[2] is None or d[2] > WINDOW for d in dev.values()), dev
    pts = dis.load_points(os.path.join(ROOT, "tests", "golden", "brachet_dissipation.json"))
    s = dis.match(pts, {r: (np.array(c[0]), np.array(c[1])) for r, c in curves.items()})["summary"]
    print(f"Brachet match: {s}", flush=True)
    assert all(v > 0 for v in s["per_curve"].values())
    assert s["max_frac_full_scale"] < 0.1
