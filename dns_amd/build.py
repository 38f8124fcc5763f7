"""Build libdns_hip.so (HIP, gfx950), the `dns` and `dns-ic` command-line
tools (host C) and, when asked, the CPU oracle used by the tests.

    python -m dns_amd.build            # library + tools
    python -m dns_amd.build --oracle   # also oracle/_build (test infrastructure)

Everything is built in-tree so the artefacts travel with the source tree.
"""
import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
OBJ = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libdns_hip.so")
BIN = os.path.join(PKG, "bin")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("DNS_OFFLOAD_ARCH", "gfx950")
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # packed FP32 instructions are disabled for all device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]
CC = os.environ.get("CC", "gcc")


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    return out


def grid_sizes():
    """Grid sizes listed in csrc/grid_sizes.h (X(n) entries)."""
    import re
    with open(os.path.join(CSRC, "grid_sizes.h")) as f:
        return [int(m) for m in re.findall(r"X\((\d+)\)", f.read())]


def units():
    """(source, object, extra flags) of every translation unit: inst.hip
    once per grid size, every other .hip source once."""
    out = []
    for f in sorted(os.listdir(CSRC)):
        if not f.endswith(".hip"):
            continue
        src = os.path.join(CSRC, f)
        if f == "inst.hip":
            # largest grids first: they take longest to compile
            for n in sorted(grid_sizes(), reverse=True):
                out.append((src, os.path.join(OBJ, f"inst_{n}.o"), [f"-DDNS_N={n}"]))
        else:
            out.append((src, os.path.join(OBJ, f[:-4] + ".o"), []))
    return out


def build_library(jobs=None, verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    jobs = jobs or min(8, os.cpu_count() or 1)

    def compile_one(u):
        src, obj, extra = u
        _run([HIPCC, *HIP_FLAGS, *extra, "-c", src, "-o", obj])
        return obj

    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(compile_one, units()))
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


def build_tools(verbose=False):
    os.makedirs(BIN, exist_ok=True)
    inc = os.path.join(ROOT, "include")
    cflags = ["-std=gnu99", "-O2", "-Wall", "-Wextra", f"-I{inc}"]
    dns = os.path.join(BIN, "dns")
    _run([CC, *cflags, "-pthread", os.path.join(PKG, "cli", "dns.c"), f"-L{PKG}", "-ldns_hip",
          "-Wl,-rpath,$ORIGIN/..", "-lm", "-o", dns])
    ic = os.path.join(BIN, "dns-ic")
    _run([CC, *cflags, os.path.join(PKG, "cli", "dns_ic.c"), f"-L{PKG}", "-ldns_hip",
          "-Wl,-rpath,$ORIGIN/..", "-lm", "-o", ic])
    if verbose:
        print("built", dns, ic)
    return dns


def build_probe(verbose=False):
    """The HBM calibration probe (copy kernels of known size that scale the
    rocprofv3 FETCH_SIZE / WRITE_SIZE counters); built only on request."""
    os.makedirs(BIN, exist_ok=True)
    probe = os.path.join(BIN, "hbm_probe")
    _run([HIPCC, *HIP_FLAGS, os.path.join(PKG, "probe", "hbm_probe.hip"), "-o", probe])
    if verbose:
        print("built", probe)
    return probe


def build_oracle(verbose=False):
    """The CPU oracle is test infrastructure; built here only so that one
    command prepares a checkout for the test-suite."""
    make = shutil.which("make") or "make"
    _run([make, "-s", "-C", os.path.join(ROOT, "oracle")])
    if verbose:
        print("built oracle/_build")


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--oracle", action="store_true", help="also build the CPU oracle")
    ap.add_argument("--probe", action="store_true", help="also build the HBM calibration probe")
    ap.add_argument("-j", type=int, default=None)
    a = ap.parse_args(argv)
    build_library(a.j, verbose=True)
    build_tools(verbose=True)
    if a.oracle:
        build_oracle(verbose=True)
    if a.probe:
        build_probe(verbose=True)


if __name__ == "__main__":
    sys.exit(main())
