"""Benchmark: RK4 steps/s of the dns step on one MI355X.

Workload (one "step" = one RK4 time step of main.c:334-422 over the whole
grid): Taylor-Green vortex n=512, nu=1/3000, dt=5e-4, inputs resident in
HBM before the timed region.  FP64 is the headline `value`; the mixed form
is timed the same way and reported with its energy error against the FP64
run after the same number of steps.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 512]

Prints one JSON line.  `value` times K steps replayed from the library's
captured step graph; a second run of K steps launches the kernels directly
with HIP events recorded around each on the library's own stream, which
gives `roofline` for the kernel with the largest share of the step time:
algorithmic bytes per launch (dns_amd/roofline.py) divided by its average
launch duration.  `cpu_baseline` times the
CPU oracle (this project's OpenMP C restatement of main.c) for a bounded
sample of the same workload on this host.
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "RK4 steps/s at 512³ (FP64 vs mixed FP32 @ energy err≤1e-5), % HBM peak"


def tgv(n):
    """Taylor-Green vortex as tgv.py:14-27 of the reference builds it."""
    dx = 2 * math.pi / n
    i = np.arange(n)
    s, c = np.sin(dx * i), np.cos(dx * i)
    u = np.empty((3, n, n, n))
    u[0] = s[:, None, None] * c[None, :, None] * c[None, None, :]
    u[1] = (-c)[:, None, None] * s[None, :, None] * c[None, None, :]
    u[2] = 0.0
    return u


def gpu_run(dns_amd, n, nu, dt, precision, u, warmup, steps):
    from dns_amd import roofline
    s = dns_amd.Solver(n, nu, dt, precision)
    s.load_velocity(u)
    shape = s.autotune(2)  # fastest of the bitwise-identical launch shapes
    if warmup:
        s.step(warmup)
    s.diagnostics()
    # timed region: steps replayed from the captured step graph
    t0 = time.perf_counter()
    s.step(steps)
    e, z = s.diagnostics()  # synchronises the library's stream
    wall = time.perf_counter() - t0
    # same number of steps launched directly with HIP events around every
    # kernel on the library's stream: per-kernel durations for the roofline
    s.set_kernel_timing(True)
    t1 = time.perf_counter()
    s.step(steps)
    s.diagnostics()
    wall_ev = time.perf_counter() - t1
    timing = s.kernel_timing()
    s.set_kernel_timing(False)
    per_pass = roofline.summarize(timing, n, precision, steps)
    kern_ms = sum(d["ms"] for d in per_pass.values())
    s.close()
    return {"steps_per_s": steps / wall, "ms_per_step": 1e3 * wall / steps, "energy": e, "enstrophy": z,
            "launch_shape": shape,
            "kernel_ms_per_step": kern_ms / steps, "ms_per_step_with_events": 1e3 * wall_ev / steps,
            "passes": per_pass}


def mixed_error(dns_amd, n, nu, dt, u, steps, every=10):
    """Largest relative deviation of the energy and enstrophy of each mixed
    form (stage increments summed in float: "mixed"; in double:
    "mixed_acc64") from the FP64 run over every diagnostics line (every 10
    steps, as dns prints them) of a `steps`-step run, and the max-norm
    deviation of the final physical velocity relative to max |u|."""
    hist, fin = {}, {}
    for prec in ("fp64", "mixed", "mixed_acc64"):
        with dns_amd.Solver(n, nu, dt, prec) as s:
            s.load_velocity(u)
            lines = [s.diagnostics()]
            for _ in range(steps // every):
                s.step(every)
                lines.append(s.diagnostics())
            hist[prec] = np.array(lines)
            fin[prec] = s.get_stale_velocity()
    a = hist["fp64"]

    def dev(p):
        b = hist[p]
        return {"max_rel_energy": float(np.max(np.abs(b[:, 0] - a[:, 0]) / a[:, 0])),
                "max_rel_enstrophy": float(np.max(np.abs(b[:, 1] - a[:, 1]) / a[:, 1])),
                "final_velocity_max_norm": float(np.max(np.abs(fin[p] - fin["fp64"])) / np.max(np.abs(fin["fp64"])))}
    out = {"steps": steps, "lines": len(a), **dev("mixed")}
    out["mixed_acc64"] = dev("mixed_acc64")
    return out


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(n, nu, dt, u, steps):
    """Times the CPU oracle's RK4 step (oracle/_build/liboracle.so) on all
    host cores, threads bound close (OMP_PROC_BIND, read by the OpenMP
    runtime when the library loads)."""
    os.environ.setdefault("OMP_PROC_BIND", "close")
    path = os.path.join(ROOT, "oracle", "_build", "liboracle.so")
    L = ctypes.CDLL(path)
    L.orc_create.restype = ctypes.c_void_p
    L.orc_create.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double]
    L.orc_load.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    L.orc_step.argtypes = [ctypes.c_void_p, ctypes.c_long]
    L.orc_destroy.argtypes = [ctypes.c_void_p]
    p = L.orc_create(n, nu, dt)
    L.orc_load(p, u.ctypes.data)
    t0 = time.perf_counter()
    L.orc_step(p, steps)
    wall = time.perf_counter() - t0
    L.orc_destroy(p)
    return {"value": steps / wall, "unit": "steps/s", "cores": int(L.orc_threads()), "kind": "port",
            "cpu_model": cpu_model(), "omp_proc_bind": os.environ.get("OMP_PROC_BIND"),
            "host_cpus": os.cpu_count(),
            "sample": f"{steps} RK4 step(s) of the same n={n} Taylor-Green workload, FP64, "
                      f"{wall:.1f} s wall, OpenMP C restatement of main.c with its own FFT "
                      f"(FFTW is not installed)"}


def ensemble(a):
    """24 simulations, n in {64,96,128,192} x Re in {100,...,3000}, half
    Taylor-Green and half isotropic, mixed precision, one graph per step."""
    import dns_amd
    res = [100, 200, 400, 800, 1600, 3000]
    ns, nus = [], []
    for n in (64, 96, 128, 192):
        for r in res:
            ns.append(n)
            nus.append(1.0 / r)
    dt = 1e-3
    b = dns_amd.Batch(ns, nus, dt, "mixed")
    for m, n in enumerate(ns):
        b.load_velocity(m, tgv(n) if m % 2 == 0 else iso_field(n, 20261202 + m))
    shapes = b.autotune(3)
    b.step(a.warmup)
    b.diagnostics(0)
    t0 = time.perf_counter()
    b.step(a.steps)
    e = [b.diagnostics(m) for m in range(len(ns))]
    wall = time.perf_counter() - t0
    pts = sum(n ** 3 for n in ns)
    line = {"metric": "ensemble RK4 steps/s (24 members, n in {64,96,128,192}, mixed precision)",
            "value": a.steps / wall, "unit": "ensemble steps/s", "n_gpus": 1, "steps": a.steps,
            "warmup": a.warmup, "ms_per_step": 1e3 * wall / a.steps, "higher_is_better": True,
            "member_steps_per_s": len(ns) * a.steps / wall, "grid_points": pts,
            "points_per_s": pts * a.steps / wall, "dtype": "f32+f64", "data": "synthetic",
            "config": {"workload": "batched ensemble, one captured graph per step", "members": len(ns),
                       "n": sorted(set(ns)), "Re": res, "dt": dt},
            "launch_shapes": {n: shapes[ns.index(n)] for n in sorted(set(ns))},
            "finite": all(np.isfinite(x[0]) for x in e)}
    b.close()
    print(json.dumps(line), flush=True)
    return 0


def _oracle_final_enstrophy(tmp, name, nu, dt, nsteps, threads):
    """Final enstrophy of the CPU oracle's command line (dns_cpu) run of
    `nsteps` steps on the raw file tmp/name."""
    import subprocess
    exe = os.path.join(ROOT, "oracle", "_build", "dns_cpu")
    env = dict(os.environ, OMP_NUM_THREADS=str(threads))
    # -t between steps nsteps-1 and nsteps: the run stops right after step nsteps
    r = subprocess.run([exe, "-i", name, "-n", repr(nu), "-s", repr(dt), "-t", repr((nsteps - 0.5) * dt)],
                       cwd=tmp, capture_output=True, text=True, env=env, check=True)
    last = r.stdout.split("\n")[-2].split()
    assert int(last[0]) == nsteps, last
    return float(last[3])


def adjoint(a):
    """Gradient of the final enstrophy, n=128 isotropic (k0=4, E=0.5),
    nu=1/800, dt=2e-3, 200 steps, checkpoint every 20; checked against
    central differences (eps=1e-6) of the GPU FP64 forward along 8 seeded
    divergence-free directions, and along the first --adj-cpu-dirs of them
    against central differences of the FP64 CPU oracle (dns_cpu, four
    processes of a quarter of the host's threads, run while the GPU works)."""
    import concurrent.futures as cf
    import tempfile
    import dns_amd
    n, nu, dt, nsteps, every, eps = 128, 1.0 / 800, 2e-3, a.adj_steps, 20, 1e-6
    u0 = iso_field(n, 20261202)
    dirs = [iso_field(n, 777 + k) for k in range(a.adj_dirs)]
    tmp = tempfile.TemporaryDirectory()
    pool = cf.ThreadPoolExecutor(max(1, 2 * a.adj_cpu_dirs))
    threads = max(1, (os.cpu_count() or 4) // max(1, 2 * a.adj_cpu_dirs))
    jobs = {}
    for k in range(min(a.adj_cpu_dirs, a.adj_dirs)):
        for sgn in (1, -1):
            name = f"d{k}{sgn:+d}.raw"
            raw = np.zeros((4, n, n, n))
            raw[:3] = u0 + sgn * eps * dirs[k]
            raw.tofile(os.path.join(tmp.name, name))
            jobs[(k, sgn)] = pool.submit(_oracle_final_enstrophy, tmp.name, name, nu, dt, nsteps, threads)
    s = dns_amd.Solver(n, nu, dt, "fp64")
    if a.adj_direct:
        s.set_kernel_timing(True)  # direct launches only (no graph capture)
    s.adjoint_forward(u0, 2, 1)  # workspace and graphs warm
    s.adjoint_backward()
    t0 = time.perf_counter()
    J = s.adjoint_forward(u0, nsteps, every)
    t1 = time.perf_counter()
    g = s.adjoint_backward()
    t2 = time.perf_counter()
    errs, cpu_errs, gpu_pm = [], [], {}
    for k, d in enumerate(dirs):
        for sgn in (1, -1):
            s.load_velocity(u0 + sgn * eps * d)
            s.step(nsteps)
            gpu_pm[(k, sgn)] = s.diagnostics()[1]
        fd = (gpu_pm[(k, 1)] - gpu_pm[(k, -1)]) / (2 * eps)
        ad = float(np.sum(g * d))
        errs.append(abs(ad - fd) / abs(fd))
        print(f"adjoint: direction {k}: adjoint {ad:.12e} GPU fd {fd:.12e} rel {errs[-1]:.2e}", file=sys.stderr,
              flush=True)
    s.close()
    fwd_dev = []
    for k in range(min(a.adj_cpu_dirs, a.adj_dirs)):
        cpu = {sgn: jobs[(k, sgn)].result() for sgn in (1, -1)}
        fd = (cpu[1] - cpu[-1]) / (2 * eps)
        ad = float(np.sum(g * dirs[k]))
        cpu_errs.append(abs(ad - fd) / abs(fd))
        fwd_dev += [abs(gpu_pm[(k, sgn)] - cpu[sgn]) / abs(cpu[sgn]) for sgn in (1, -1)]
        print(f"adjoint: direction {k}: adjoint {ad:.12e} oracle fd {fd:.12e} rel {cpu_errs[-1]:.2e}",
              file=sys.stderr, flush=True)
    pool.shutdown()
    tmp.cleanup()
    line = {"metric": "adjoint gradient of final enstrophy vs central differences",
            "value": max(errs + cpu_errs), "unit": "max relative error", "higher_is_better": False, "n_gpus": 1,
            "dtype": "f64", "data": "synthetic", "J": J, "forward_s": t1 - t0, "backward_s": t2 - t1,
            "backward_over_forward": (t2 - t1) / (t1 - t0),
            "rel_errors_gpu_fd": errs, "rel_errors_oracle_fd": cpu_errs,
            "final_enstrophy_gpu_vs_oracle_max_rel": max(fwd_dev) if fwd_dev else None,
            "config": {"workload": "gradient of final enstrophy w.r.t. initial velocity", "n": n, "nu": nu, "dt": dt,
                       "steps": nsteps, "checkpoint_every": every, "eps": eps, "directions": a.adj_dirs,
                       "oracle_directions": min(a.adj_cpu_dirs, a.adj_dirs)}}
    print(json.dumps(line), flush=True)
    return 0


def dump_overhead(a):
    """Wall time of the drop-in command line with and without -d (snapshot
    and XDMF files every 10 steps, written while the next steps run) on the
    Taylor-Green field, n = --n, --steps steps; files go to --dump-dir and
    are deleted afterwards.  Times are the command line's own -v step-loop
    figure (diagnostics and snapshots included, reading excluded)."""
    import re
    import shutil
    import subprocess
    import tempfile
    n, steps = a.n, a.steps
    nu, dt = (1.0 / 3000, 5e-4) if n == 512 else (1.0 / 1600, 1e-3)
    exe = os.path.join(ROOT, "dns_amd", "bin", "dns")
    d = tempfile.mkdtemp(dir=a.dump_dir)
    res = {}
    try:
        raw = os.path.join(d, "in.raw")
        f = np.zeros((4, n, n, n))
        f[:3] = tgv(n)
        f.tofile(raw)
        del f
        args = ["-v", "-i", raw, "-n", repr(nu), "-s", repr(dt), "-t", repr((steps - 0.5) * dt)]
        for mode in ("plain", "dump", "plain2", "dump2"):
            out = os.path.join(d, mode)
            os.mkdir(out)
            print(f"dump: {mode} n={n} steps={steps}", file=sys.stderr, flush=True)
            extra = ["-d"] if mode.startswith("dump") else []
            r = subprocess.run([exe, *extra, *args], cwd=out, capture_output=True, text=True, check=True)
            m = re.search(r"(\d+) steps in ([0-9.]+) s", r.stderr)
            files = sorted(os.listdir(out))
            res[mode] = {"steps": int(m.group(1)), "wall_s": float(m.group(2)), "files": len(files),
                         "bytes": sum(os.path.getsize(os.path.join(out, x)) for x in files)}
            shutil.rmtree(out)
    finally:
        shutil.rmtree(d, ignore_errors=True)
    plain = min(res["plain"]["wall_s"], res["plain2"]["wall_s"])
    dump = min(res["dump"]["wall_s"], res["dump2"]["wall_s"])
    line = {"metric": "dns -d wall time over dns without snapshots", "value": dump / plain, "unit": "ratio",
            "higher_is_better": False, "n_gpus": 1, "dtype": "f64", "data": "synthetic",
            "plain_s": plain, "dump_s": dump, "runs": res,
            "config": {"workload": f"Taylor-Green n={n}, {steps} steps, snapshot every 10 steps", "n": n,
                       "dump_dir": a.dump_dir}}
    print(json.dumps(line), flush=True)
    return 0


def validate(a):
    """Taylor-Green n=256, Re in {100,...,3000}, dt=0.0025, 4000 steps (t=10),
    FP64, run as one batch and compared line by line with the reference's
    own trajectories (directory 0256 of the reference, excerpted in
    tests/golden/tgv_trajectories.json); also the dissipation rate
    eps = 2 Z / Re (tgv.gp:11), its peak, and the match of the six eps
    curves with the digitised Brachet et al. points the reference plots
    them over (img/ref.txt, in tests/golden/brachet_dissipation.json)."""
    import dns_amd
    from dns_amd import dissipation as dis
    with open(os.path.join(ROOT, "tests", "golden", "tgv_trajectories.json")) as f:
        gold = json.load(f)
    res = [100, 200, 400, 800, 1600, 3000]
    n, dt = 256, 0.0025
    keys = [f"256/{r}" for r in res]
    nus = [float(gold[k]["nu"]) for k in keys]
    rows = [{int(r[0]): (float(r[2]), float(r[3])) for r in gold[k]["rows"]} for k in keys]
    b = dns_amd.Batch([n] * len(res), nus, dt, "fp64")
    u = tgv(n)
    for m in range(len(res)):
        b.load_velocity(m, u)
    out = {r: {"max_rel_E": 0.0, "max_rel_Z": 0.0, "first_step_over_1e-12": None, "peak_eps": 0.0,
               "peak_eps_t": 0.0, "gold_peak_eps": 0.0} for r in res}
    curves = {r: ([], []) for r in res}
    t0 = time.perf_counter()
    step = 0
    while step <= a.val_steps:
        for m, r in enumerate(res):
            e, z = b.diagnostics(m)
            curves[r][0].append(step * dt)
            curves[r][1].append(2 * z / r)
            o = out[r]
            if step in rows[m]:
                ge, gz = rows[m][step]
                de, dz = abs(e - ge) / ge, abs(z - gz) / gz
                o["max_rel_E"] = max(o["max_rel_E"], de)
                o["max_rel_Z"] = max(o["max_rel_Z"], dz)
                if o["first_step_over_1e-12"] is None and max(de, dz) > 1e-12:
                    o["first_step_over_1e-12"] = step
                if 2 * gz / r > o["gold_peak_eps"]:
                    o["gold_peak_eps"] = 2 * gz / r
            if 2 * z / r > o["peak_eps"]:
                o["peak_eps"], o["peak_eps_t"] = 2 * z / r, step * dt
        if step == a.val_steps:
            break
        b.step(10)
        step += 10
        if step % 500 == 0:
            print(f"validate: step {step}", file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    b.close()
    brachet = dis.match(dis.load_points(os.path.join(ROOT, "tests", "golden", "brachet_dissipation.json")),
                        {r: (np.array(c[0]), np.array(c[1])) for r, c in curves.items()})["summary"]
    line = {"metric": "Taylor-Green n=256 trajectories vs the reference's golden runs", "value":
            max(max(o["max_rel_E"], o["max_rel_Z"]) for o in out.values()), "unit": "max relative deviation",
            "higher_is_better": False, "n_gpus": 1, "dtype": "f64", "steps": a.val_steps, "wall_s": wall,
            "per_Re": out, "brachet_eps_match": brachet,
            "config": {"workload": "TGV n=256, dt=0.0025, nu=1/Re", "Re": res}}
    print(json.dumps(line), flush=True)
    return 0


def iso_field(n, seed=20261202):
    """Seeded isotropic field of dns-ic (k0=4, kinetic energy 0.5)."""
    import subprocess
    import tempfile
    exe = os.path.join(ROOT, "dns_amd", "bin", "dns-ic")
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "iso.raw")
        subprocess.run([exe, "-t", "iso", "-N", str(n), "-S", str(seed), "-o", path], check=True,
                       capture_output=True)
        return np.fromfile(path, dtype="<f8").reshape(4, n, n, n)[:3].copy()


def configs(a):
    """The long runs of the measurement plan, FP64 and mixed side by side:
    Taylor-Green and the seeded isotropic field at n=256 (nu=6.25e-4,
    dt=1e-3, 2000 steps) and Taylor-Green at n=512 (nu=1/3000, dt=5e-4,
    500 steps).  Reported per run: steps/s of each precision (graph
    replay, diagnostics every 10 steps as the command line does), the
    largest relative deviation of the mixed energy and enstrophy from the
    FP64 run over all diagnostics lines, and the max-norm deviation of the
    final physical velocity relative to max |u|."""
    import dns_amd
    runs = [("tgv", 256, 6.25e-4, 1e-3, 2000), ("iso", 256, 6.25e-4, 1e-3, 2000),
            ("tgv", 512, 1 / 3000, 5e-4, 500)]
    out = []
    for ic, n, nu, dt, steps in runs:
        u = tgv(n) if ic == "tgv" else iso_field(n)
        res = {"ic": ic, "n": n, "nu": nu, "dt": dt, "steps": steps}
        hist, finals = {}, {}
        for prec in ("fp64", "mixed"):
            print(f"configs: {ic} n={n} {prec}", file=sys.stderr, flush=True)
            with dns_amd.Solver(n, nu, dt, prec) as s:
                s.load_velocity(u)
                s.step(2)
                s.load_velocity(u)  # graph and launch shapes warm, state reset
                lines = [s.diagnostics()]
                t0 = time.perf_counter()
                for _ in range(steps // 10):
                    s.step(10)
                    lines.append(s.diagnostics())
                wall = time.perf_counter() - t0
                hist[prec] = lines
                finals[prec] = s.get_stale_velocity()
                res[f"{prec}_steps_per_s"] = steps / wall
        de = max(abs(m[0] - f[0]) / f[0] for f, m in zip(hist["fp64"], hist["mixed"]))
        dz = max(abs(m[1] - f[1]) / f[1] for f, m in zip(hist["fp64"], hist["mixed"]))
        ua, ub = finals["fp64"], finals["mixed"]
        res.update({"mixed_max_rel_E": de, "mixed_max_rel_Z": dz,
                    "mixed_max_norm_u": float(np.abs(ub - ua).max() / np.abs(ua).max()),
                    "E_final_fp64": hist["fp64"][-1][0], "Z_final_fp64": hist["fp64"][-1][1],
                    "speedup": res["mixed_steps_per_s"] / res["fp64_steps_per_s"]})
        out.append(res)
        print(f"configs: {json.dumps(res)}", file=sys.stderr, flush=True)
    worst = max(r["mixed_max_rel_E"] for r in out)
    line = {"metric": "long runs, FP64 and mixed: max relative energy deviation of mixed", "value": worst,
            "unit": "relative", "higher_is_better": False, "n_gpus": 1, "dtype": "f64 / f32+f64",
            "data": "synthetic", "runs": out}
    print(json.dumps(line), flush=True)
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--cpu-steps", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--ensemble", action="store_true", help="batched ensemble workload instead")
    ap.add_argument("--adjoint", action="store_true", help="adjoint gradient check instead")
    ap.add_argument("--adj-steps", type=int, default=200)
    ap.add_argument("--adj-dirs", type=int, default=8)
    ap.add_argument("--validate", action="store_true", help="n=256 golden-trajectory validation instead")
    ap.add_argument("--configs", action="store_true", help="long FP64/mixed runs at n=256 and 512 instead")
    ap.add_argument("--val-steps", type=int, default=4000)
    ap.add_argument("--err-steps", type=int, default=500,
                    help="steps of the FP64-vs-mixed error run (0: skip)")
    ap.add_argument("--adj-cpu-dirs", type=int, default=2,
                    help="directions also differenced with the CPU oracle")
    ap.add_argument("--adj-direct", action="store_true",
                    help="adjoint: launch every kernel directly, no graph capture (for tracing profilers)")
    ap.add_argument("--dump", action="store_true", help="wall time of dns -d against dns instead")
    ap.add_argument("--dump-dir", default="/tmp", help="directory the snapshots are written to")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU: runs on several GPUs are not supported (ensemble members "
              "would shard with no communication; one large grid would need x/y slabs and an "
              "all-to-all per transform)", file=sys.stderr)
        return 2
    if a.ensemble:
        return ensemble(a)
    if a.adjoint:
        return adjoint(a)
    if a.validate:
        return validate(a)
    if a.configs:
        return configs(a)
    if a.dump:
        return dump_overhead(a)
    import dns_amd
    from dns_amd import roofline

    n = a.n
    nu, dt = (1.0 / 3000, 5e-4) if n == 512 else (1.0 / 1600, 1e-3)
    u = tgv(n)
    print(f"bench: n={n} fp64 warmup={a.warmup} steps={a.steps}", file=sys.stderr, flush=True)
    r64 = gpu_run(dns_amd, n, nu, dt, "fp64", u, a.warmup, a.steps)
    mixed = None
    if not a.no_mixed:
        print("bench: mixed", file=sys.stderr, flush=True)
        rmx = gpu_run(dns_amd, n, nu, dt, "mixed", u, a.warmup, a.steps)
        mixed = {"steps_per_s": rmx["steps_per_s"], "ms_per_step": rmx["ms_per_step"],
                 "launch_shape": rmx["launch_shape"],
                 "step_bytes": roofline.step_bytes_fused(n, "mixed"),
                 "speedup_vs_fp64": rmx["steps_per_s"] / r64["steps_per_s"],
                 "passes": {k: {"avg_ms": v["avg_ms"], "GBs": v["GBs"], "frac": v["frac"]}
                            for k, v in rmx["passes"].items()}}
        # the same form with the stage increments summed in double
        print("bench: mixed_acc64", file=sys.stderr, flush=True)
        r2 = gpu_run(dns_amd, n, nu, dt, "mixed_acc64", u, a.warmup, a.steps)
        mixed["acc64"] = {"steps_per_s": r2["steps_per_s"], "ms_per_step": r2["ms_per_step"],
                          "step_bytes": roofline.step_bytes_fused(n, "mixed_acc64"),
                          "speedup_vs_fp64": r2["steps_per_s"] / r64["steps_per_s"]}
        if a.err_steps:
            print(f"bench: mixed forms vs fp64 over {a.err_steps} steps", file=sys.stderr, flush=True)
            err = mixed_error(dns_amd, n, nu, dt, u, a.err_steps)
            mixed["acc64"]["error_vs_fp64"] = err.pop("mixed_acc64")
            mixed["error_vs_fp64"] = err
    dom_name, dom = max(r64["passes"].items(), key=lambda kv: kv[1]["ms"])
    # HBM traffic per launch measured by rocprofv3 FETCH_SIZE/WRITE_SIZE runs
    # (calibrated on a copy kernel), committed with the profiles
    traffic = None
    tpath = os.path.join(ROOT, "profiles", f"traffic_n{n}_fp64.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            traffic = json.load(f).get("traffic_bytes_per_launch", {}).get(dom_name)
    step_b = roofline.step_bytes_fused(n, "fp64")
    line = {
        "metric": METRIC, "value": r64["steps_per_s"], "unit": "steps/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": r64["ms_per_step"],
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"Taylor-Green vortex n={n}, nu={nu:.6g}, dt={dt:g}, one RK4 step (FP64)",
                   "n": n, "nu": nu, "dt": dt, "precision": "fp64", "launch_shape": r64["launch_shape"]},
        "roofline": {"bound": "hbm", "kernel": dom_name, "achieved": dom["GBs"], "peak": roofline.PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": dom["frac"], "traffic": traffic,
                     "bytes_per_launch": dom["bytes_per_launch"], "avg_launch_ms": dom["avg_ms"]},
        "step_roofline": {"algorithmic_bytes_per_step": step_b,
                          "achieved_GBs": step_b * r64["steps_per_s"] / 1e9,
                          "frac": step_b * r64["steps_per_s"] / 1e9 / roofline.PEAK_HBM_GBS},
        "kernels": {k: {"avg_ms": v["avg_ms"], "GBs": v["GBs"], "frac": v["frac"],
                        "ms_per_step": v["ms_per_step"]} for k, v in r64["passes"].items()},
        "kernel_ms_per_step": r64["kernel_ms_per_step"],
        "ms_per_step_with_events": r64["ms_per_step_with_events"],
        "mixed": mixed,
        "energy": r64["energy"], "enstrophy": r64["enstrophy"],
    }
    if not a.no_cpu:
        print(f"bench: cpu baseline n={n} steps={a.cpu_steps}", file=sys.stderr, flush=True)
        line["cpu_baseline"] = cpu_baseline(n, nu, dt, u, a.cpu_steps)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
