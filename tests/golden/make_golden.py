"""Extract golden-trajectory fixtures from the reference's own data files.

The reference ships the stdout of its solver for the Taylor-Green vortex
(n=128 in directory 0128, n=256 in 0256; Re in {100,...,3000}, nu=1/Re,
dt=0.0025, T=10, i.e. 4000 RK4 steps, one line every 10 steps; recipe in
run.sh:1-10 and hal.sh:1-16).  This script keeps, as data (step, t, energy,
enstrophy), every line of the n=128 files and, of the n=256 files, the
lines of steps 0..400, every 100th step after that and the final line --
no reference source is copied.

It also copies the reference's digitised dissipation-rate points (img/ref.txt:
(t, eps) pairs read off the Taylor-Green curves of Brachet et al. 1983 for
Re = 100..3000, plotted by tgv.gp against the runs' 2 Z / Re) into
brachet_dissipation.json, sorted by time.

Run in the development container (the reference is not on the GPU box):
    python tests/golden/make_golden.py /root/reference
"""
import json
import os
import sys


def main(root):
    out = {}
    for n in ("0128", "0256"):
        for re_ in ("0100", "0200", "0400", "0800", "1600", "3000"):
            rows = []
            with open(os.path.join(root, n, re_)) as f:
                for line in f:
                    step, t, e, z = line.split()
                    step = int(step)
                    if n == "0128" or step <= 400 or step % 100 == 0:
                        rows.append([step, t, e, z])
            out[f"{int(n)}/{int(re_)}"] = {
                "n": int(n), "Re": int(re_), "nu": "%.16e" % (1.0 / int(re_)),
                "dt": 0.0025, "T": 10.0, "rows": rows}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tgv_trajectories.json")
    with open(path, "w") as f:
        f.write("{\n")
        for ki, (k, v) in enumerate(out.items()):
            f.write('"%s": {"n": %d, "Re": %d, "nu": "%s", "dt": %r, "T": %r, "rows": [\n'
                    % (k, v["n"], v["Re"], v["nu"], v["dt"], v["T"]))
            f.write(",\n".join(json.dumps(r) for r in v["rows"]))
            f.write("]}%s\n" % ("," if ki < len(out) - 1 else ""))
        f.write("}\n")
    print(path, sum(len(v["rows"]) for v in out.values()), "rows")
    pts = []
    with open(os.path.join(root, "img", "ref.txt")) as f:
        for line in f:
            if line.strip():
                t, e = line.split()
                pts.append([t, e])
    pts.sort(key=lambda r: float(r[0]))
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "brachet_dissipation.json")
    with open(path, "w") as f:
        f.write('{"source": "img/ref.txt", "columns": ["t", "eps"], "points": [\n')
        f.write(",\n".join(json.dumps(r) for r in pts))
        f.write("]}\n")
    print(path, len(pts), "points")


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
