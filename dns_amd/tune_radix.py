"""Radix-schedule tuning of the FFT kernels, done offline per grid size.

A schedule (fft_engine.hpp, DNS_SCHED) fixes the elements per thread E and
the radix of every Stockham stage; it is a compile-time constant of every
kernel, so each candidate is a separate build.  The candidates of a size n
are the orderings of every radix multiset (radices from {2, 3, 4, 8}
dividing the table's E, product n) with the fewest stages: each stage is
one exchange through LDS.  Each candidate is compiled into a library of that
one size, with the table's schedule replaced through -DDNS_SCHED_N,
-DDNS_SCHED_E and -DDNS_SCHED_R, and timed on the GPU against the table's
schedule (FP64 and mixed Taylor-Green steps, default launch shape), each
in several fresh processes taking turns, since a library's speed varies by
a few percent between processes with where its buffers land in device
memory.  Every candidate must also reproduce the table schedule's result to
rounding (energy and enstrophy after the timed steps within 1e-12
relative; the spectral state within 1e-11 of its max norm up to n=256).
The chosen schedules are entered in the table by hand.

    python -m dns_amd.tune_radix build 128 256     # on the CPU: dns_amd/_tune/
    python -m dns_amd.tune_radix run 128 256       # on the GPU: one JSON line per candidate
"""
import argparse
import concurrent.futures as cf
import ctypes
import itertools
import json
import math
import os
import re
import shutil
import sys
import time

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
TUNE = os.path.join(PKG, "_tune")


def table():
    """{n: (E, radices)} of the DNS_SCHED table."""
    with open(os.path.join(PKG, "csrc", "fft_engine.hpp")) as f:
        src = f.read()
    out = {}
    for m in re.finditer(r"^DNS_SCHED\((\d+), (\d+), ([\d, ]+)\)", src, re.M):
        out[int(m.group(1))] = (int(m.group(2)), tuple(int(x) for x in m.group(3).split(",")))
    return out


def candidates(n):
    """Schedules with the table's E and the fewest stages; the table's own
    schedule first."""
    e, cur = table()[n]
    radices = [r for r in (8, 4, 3, 2) if e % r == 0]
    best = []
    for stages in range(1, 12):
        sets = [c for c in itertools.combinations_with_replacement(radices, stages) if math.prod(c) == n]
        if sets:
            for c in sets:
                best.extend(sorted(set(itertools.permutations(c)), reverse=True))
            break
    best = [cur] + [c for c in best if c != cur]
    return e, best


def name(n, rad):
    return f"n{n}_" + "-".join(map(str, rad))


def build(sizes, jobs=8):
    from dns_amd import build as B
    os.makedirs(TUNE, exist_ok=True)
    work = []
    for n in sizes:
        e, cands = candidates(n)
        # the other translation units, built once per size against a size
        # list holding n alone
        common = os.path.join(TUNE, f"n{n}_common")
        shutil.rmtree(common, ignore_errors=True)
        src = os.path.join(common, "tree", "csrc")
        shutil.copytree(os.path.join(PKG, "csrc"), src)
        # engine.hpp includes ../../include/dns.h relative to itself:
        # common/include
        os.makedirs(os.path.join(common, "include"), exist_ok=True)
        shutil.copy(os.path.join(ROOT, "include", "dns.h"), os.path.join(common, "include", "dns.h"))
        with open(os.path.join(src, "grid_sizes.h"), "w") as f:
            f.write(f"#pragma once\n#define DNS_GRID_SIZES(X) X({n})\n")
        for fn in sorted(os.listdir(src)):
            if fn.endswith(".hip") and fn != "inst.hip":
                work.append((os.path.join(src, fn), os.path.join(common, fn[:-4] + ".o"), []))
        for rad in cands:
            d = os.path.join(TUNE, name(n, rad))
            os.makedirs(d, exist_ok=True)
            flags = [f"-DDNS_N={n}", f"-DDNS_SCHED_N={n}", f"-DDNS_SCHED_E={e}",
                     "-DDNS_SCHED_R=" + ",".join(map(str, rad))]
            work.append((os.path.join(src, "inst.hip"), os.path.join(d, "inst.o"), flags))

    def cc(w):
        s, o, flags = w
        B._run([B.HIPCC, *B.HIP_FLAGS, *flags, "-c", s, "-o", o])

    with cf.ThreadPoolExecutor(jobs) as ex:
        list(ex.map(cc, work))
    for n in sizes:
        common = os.path.join(TUNE, f"n{n}_common")
        objs = [os.path.join(common, f) for f in sorted(os.listdir(common)) if f.endswith(".o")]
        for rad in candidates(n)[1]:
            d = os.path.join(TUNE, name(n, rad))
            B._run([B.HIPCC, f"--offload-arch={B.ARCH}", "-shared", "-fPIC", "-o",
                    os.path.join(d, "libdns_hip.so"), os.path.join(d, "inst.o"), *objs])
            os.remove(os.path.join(d, "inst.o"))
        for f in objs:
            os.remove(f)
        shutil.rmtree(os.path.join(common, "tree"))
        shutil.rmtree(os.path.join(common, "include"))
        print("built", n, len(candidates(n)[1]), "candidates", flush=True)


def _tgv(n):
    import numpy as np
    x = 2 * np.pi * np.arange(n) / n
    s, c = np.sin(x), np.cos(x)
    u = np.zeros((3, n, n, n))
    u[0] = s[:, None, None] * c[None, :, None] * c[None, None, :]
    u[1] = -c[:, None, None] * s[None, :, None] * c[None, None, :]
    return u


def _time(path, n, prec, u, steps, want_state):
    """One candidate library in this process: steps/s (best of three timed
    runs of `steps` steps, default launch shape), energy and enstrophy after
    them, and the spectral state if asked."""
    import numpy as np
    L = ctypes.CDLL(path)
    P, I, D, Lg = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_long
    L.dns_create.argtypes = [ctypes.POINTER(P), I, D, D, I, I]
    L.dns_load_velocity.argtypes = [P, P]
    L.dns_step.argtypes = [P, Lg]
    L.dns_diagnostics.argtypes = [P, ctypes.POINTER(D), ctypes.POINTER(D)]
    L.dns_get_spectral.argtypes = [P, P]
    L.dns_destroy.argtypes = [P]
    ctx = P()
    nu, dt = (1 / 3000, 5e-4) if n == 512 else (1 / 1600, 1e-3)
    assert L.dns_create(ctypes.byref(ctx), n, nu, dt, prec, 0) == 0
    try:
        assert L.dns_load_velocity(ctx, u.ctypes.data) == 0
        e, z = D(), D()
        assert L.dns_step(ctx, 3) == 0
        L.dns_diagnostics(ctx, ctypes.byref(e), ctypes.byref(z))
        best = 1e30
        for _ in range(3):
            t0 = time.perf_counter()
            assert L.dns_step(ctx, steps) == 0
            assert L.dns_diagnostics(ctx, ctypes.byref(e), ctypes.byref(z)) == 0
            best = min(best, time.perf_counter() - t0)
        state = None
        if want_state:
            nf = n // 2 + 1
            state = np.empty((3, n, n, nf, 2))
            assert L.dns_get_spectral(ctx, state.ctypes.data) == 0
        return {"steps_per_s": steps / best, "E": e.value, "Z": z.value}, state
    finally:
        L.dns_destroy(ctx)


def _one(path, n, prec, steps, state_out):
    """Child process: time one candidate library, print its result line."""
    import numpy as np
    r, st = _time(path, n, prec, _tgv(n), steps, bool(state_out))
    if st is not None:
        np.save(state_out, st)
    print("RESULT " + json.dumps(r), flush=True)


def run(sizes, out, repeats=3, precisions=("fp64", "mixed")):
    """Times every built candidate of each size.  The speed of one library
    varies by a few percent from process to process with where its buffers
    land in device memory, so each candidate runs in `repeats` fresh
    processes, the candidates taking turns (rotated order per repeat), and
    the median counts.  The first repeat also checks each candidate against
    the table's schedule: energy and enstrophy after the timed steps (FP64
    within 1e-12, mixed 1e-5) and, up to n=256, the spectral state (1e-11,
    mixed 1e-4, of its max norm)."""
    import statistics
    import subprocess
    import tempfile
    import numpy as np
    res = []
    tmp = tempfile.mkdtemp(prefix="tune_radix_")
    for n in sizes:
        steps = max(6, int(2e6 / n ** 2))
        e, cands = candidates(n)
        for pname in precisions:
            prec = {"fp64": 0, "mixed": 1}[pname]
            runs = {c: [] for c in cands}
            check = {}
            for rep in range(repeats):
                order = cands[rep % len(cands):] + cands[:rep % len(cands)]
                for rad in order:
                    path = os.path.join(TUNE, name(n, rad), "libdns_hip.so")
                    st_out = os.path.join(tmp, name(n, rad) + ".npy") if rep == 0 and n <= 256 else ""
                    p = subprocess.run([sys.executable, "-m", "dns_amd.tune_radix", "one", path, str(n), str(prec),
                                        str(steps), st_out or "-"], capture_output=True, text=True, timeout=900,
                                       cwd=ROOT)
                    line = [l for l in p.stdout.splitlines() if l.startswith("RESULT ")]
                    if p.returncode or not line:
                        raise RuntimeError(f"candidate {name(n, rad)} {pname} failed:\n{p.stderr[-2000:]}")
                    r = json.loads(line[0][7:])
                    runs[rad].append(r["steps_per_s"])
                    if rep == 0:
                        check[rad] = (r, st_out)
                    print(f"n={n} {pname} {'-'.join(map(str, rad))} rep {rep}: {r['steps_per_s']:.2f} steps/s",
                          flush=True)
            r0, st0 = check[cands[0]]
            ref_state = np.load(st0) if st0 else None
            med0 = statistics.median(runs[cands[0]])
            for rad in cands:
                r, st = check[rad]
                med = statistics.median(runs[rad])
                row = {"n": n, "precision": pname, "E_per_thread": e, "radices": list(rad), "steps": steps,
                       "table": rad == cands[0], "steps_per_s": runs[rad], "median": med,
                       "speedup_vs_table": med / med0,
                       "dE": abs(r["E"] - r0["E"]) / r0["E"], "dZ": abs(r["Z"] - r0["Z"]) / r0["Z"]}
                if ref_state is not None:
                    row["dstate"] = float(np.max(np.abs(np.load(st) - ref_state)) / np.max(np.abs(ref_state)))
                tol, stol = (1e-12, 1e-11) if prec == 0 else (1e-5, 1e-4)
                row["ok"] = row["dE"] <= tol and row["dZ"] <= tol and row.get("dstate", 0.0) <= stol
                print(json.dumps(row), flush=True)
                res.append(row)
            for rad in cands:
                if check[rad][1]:
                    os.remove(check[rad][1])
    if out:
        with open(out, "w") as f:
            json.dump(res, f, indent=1)


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    if argv is None:
        argv = sys.argv[1:]
    if argv and argv[0] == "one":  # child process of `run`
        path, n, prec, steps, st = argv[1:6]
        return _one(path, int(n), int(prec), int(steps), "" if st == "-" else st)
    ap.add_argument("what", choices=["list", "build", "run"])
    ap.add_argument("sizes", type=int, nargs="+")
    ap.add_argument("-o", "--out", default=None)
    ap.add_argument("--repeats", type=int, default=3, help="fresh processes per candidate (run)")
    ap.add_argument("--precisions", default="fp64,mixed", help="comma-separated: fp64, mixed (run)")
    a = ap.parse_args(argv)
    if a.what == "list":
        for n in a.sizes:
            e, c = candidates(n)
            print(n, e, len(c), c)
    elif a.what == "build":
        build(a.sizes)
    else:
        run(a.sizes, a.out, a.repeats, tuple(a.precisions.split(",")))


if __name__ == "__main__":
    sys.exit(main())
