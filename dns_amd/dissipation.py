"""Dissipation-rate validation of Taylor-Green runs against Brachet et al.

The reference checks its solver by plotting, for six n=256 Taylor-Green
runs (Re = 100, 200, 400, 800, 1600, 3000; nu = 1/Re, dt = 0.0025, t = 10),
the energy dissipation rate

    eps(t) = 2 Z(t) / Re        (tgv.gp:11: column 2 against 2*$4/r)

where Z is the enstrophy column of the solver's stdout, over points read
off the curves of Brachet et al. (1983) (img/ref.txt, plotted by
tgv.gp:10).  The comparison there is visual; here it is a number.  The
digitised points carry no Reynolds-number label, so each point is matched
to the curve nearest to it at its own time (curves interpolated linearly
between printed lines); a run set agrees when every point lies within a
small fraction of the plot's full scale (0.02) of some curve and the points
spread over all curves.

    python -m dns_amd.dissipation --points brachet.json 100=out.0100 200=out.0200 ...

reads `dns` stdout files (format `% 10ld % .16e % .16Le % .16Le`,
main.c:267) and prints the summary as JSON.
"""
import argparse
import json
import sys

import numpy as np

FULL_SCALE = 0.02  # y range of the reference's plot (tgv.gp:9)


def read_stdout(lines):
    """(step, t, E, Z) arrays from `dns` stdout lines (an iterable of str)."""
    rows = [ln.split() for ln in lines if ln.strip()]
    a = np.array([[float(x) for x in r[:4]] for r in rows], dtype=np.float64).reshape(-1, 4)
    return a[:, 0].astype(np.int64), a[:, 1], a[:, 2], a[:, 3]


def dissipation(z, re):
    """eps = 2 Z / Re (tgv.gp:11)."""
    return 2.0 * np.asarray(z, dtype=np.float64) / float(re)


def match(points, curves):
    """Match digitised (t, eps) points to the nearest of `curves`.

    points: array (m, 2); curves: {Re: (t array ascending, eps array)}.
    Returns {"dist": distances (m,), "re": assigned Re (m,), "summary": {...}}
    with the summary's max / median / 90th percentile distance (absolute and
    as a fraction of FULL_SCALE) and the number of points per curve.  Points
    outside a curve's time span are compared with its end values."""
    pts = np.asarray(points, dtype=np.float64).reshape(-1, 2)
    keys = sorted(curves)
    if not keys:
        raise ValueError("no curves")
    d = np.empty((len(keys), len(pts)))
    for c, r in enumerate(keys):
        t, e = (np.asarray(x, dtype=np.float64) for x in curves[r])
        if t.ndim != 1 or t.shape != e.shape or len(t) < 2 or np.any(np.diff(t) <= 0):
            raise ValueError(f"curve {r}: times must be a strictly increasing 1-D array matching eps")
        d[c] = np.abs(np.interp(pts[:, 0], t, e) - pts[:, 1])
    best = np.argmin(d, axis=0)
    dist = d[best, np.arange(len(pts))]
    assigned = np.array([keys[b] for b in best])
    summary = {
        "points": int(len(pts)),
        "max": float(dist.max()) if len(pts) else 0.0,
        "median": float(np.median(dist)) if len(pts) else 0.0,
        "p90": float(np.percentile(dist, 90)) if len(pts) else 0.0,
        "max_frac_full_scale": float(dist.max() / FULL_SCALE) if len(pts) else 0.0,
        "per_curve": {str(r): int(np.sum(assigned == r)) for r in keys},
    }
    return {"dist": dist, "re": assigned, "summary": summary}


def load_points(path):
    """(t, eps) points of a JSON file {"points": [[t, eps], ...]} (numbers or
    numeric strings)."""
    with open(path) as f:
        doc = json.load(f)
    return np.array([[float(a), float(b)] for a, b in doc["points"]], dtype=np.float64)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--points", required=True, help="JSON file of digitised (t, eps) points")
    ap.add_argument("runs", nargs="+", help="RE=PATH of a dns stdout file")
    a = ap.parse_args(argv)
    curves = {}
    for spec in a.runs:
        re_s, _, path = spec.partition("=")
        if not path:
            ap.error(f"expected RE=PATH, got {spec!r}")
        with open(path) as f:
            _, t, _, z = read_stdout(f)
        re = float(re_s)
        curves[int(re) if re.is_integer() else re] = (t, dissipation(z, re))
    res = match(load_points(a.points), curves)
    print(json.dumps(res["summary"]))
    return 0


if __name__ == "__main__":
    sys.exit(main())
