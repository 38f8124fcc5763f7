"""Algorithmic HBM bytes of the step's kernels (the roofline numerator).

Counting rule: every logical array element a kernel must read or write is
counted once, at the precision it is stored in; padding, cache re-reads and
twiddle tables are not counted.  With Uc = n^2 (n/2+1) complex elements per
spectral field and Rr = n^3 reals per physical field:

  pass  stage 1                 stages 2,3                stage 4
  A     r 3Uc(f64) w 3Uc        r 3Uc(SR) w 5Uc           as 2,3
  B     r 3Uc w 3Uc             r 5Uc w 6Uc               as 2,3
  C     r 3Uc+3Rr w 3Zc         r 6Uc w 3Zc               r 6Uc w 3Zc+3Rr
  D     r 3Zc w 3Yc             as 1                      as 1
  E     r 3Yc+3Uc(f64)          r 3Yc+3Uc(SR)+6Uc(f64)    r 3Yc+3Uc(SR)+3Uc(f64)
        w 3Uc(SR)+3Uc(f64)      w 3Uc(SR)+3Uc(f64)        w 3Uc(f64)
        (FP64; the mixed-precision state terms are listed in slot_bytes)

where Zc = n^2 nk and Yc = n (2nk-1) nk count the dealiased part of the
nonlinear term (nk = #{m >= 0 : m < kmax}, kmax = 2/3 (n/2+1)).  From stage 2
on, A hands u, v, w, omega_y, omega_z to B, which forms omega_x from v and w
(5 fields read, 6 written) -- except in FP64 where a y pencil needs more
than 64 threads at the schedule's elements per thread (n=1024): there A
writes all six fields and B reads six.

Partials (A..E outputs) and the stale velocity are in the working precision
R; the step-start state and the accumulator are always double; SR is the
stage state's precision (= R).  A step at n=512 (A fused into the previous
stage's E) moves 171.2 GB in FP64 and 92.1 GB in mixed precision.
"""

PEAK_HBM_GBS = 8000.0  # MI355X HBM3E, published peak
PASSES = ("xinv_curl", "ypass_inv", "zfused", "ypass_fwd", "xfwd_rk", "xfwd_rk_xinv")


def n_box(n):
    """Wavenumbers m >= 0 with m < kmax = 2/3 (n/2+1) (main.c:231-238)."""
    kmax = 2.0 / 3.0 * (n // 2 + 1)
    m = 0
    while m <= n // 2 and m < kmax:
        m += 1
    return m


def elements_per_thread(n):
    """Elements per thread E of n's radix schedule (DNS_SCHED entries of
    csrc/fft_engine.hpp: a pencil takes n / E threads)."""
    import os
    import re
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "csrc", "fft_engine.hpp")
    with open(path) as f:
        table = {int(a): int(b) for a, b in re.findall(r"DNS_SCHED\((\d+),\s*(\d+)", f.read())}
    return table[n]


def curlx_in_y(n, precision="fp64"):
    """Whether the inverse y pass forms omega_x (see the table above): not
    where FP64 y pencils take twice the schedule's elements per thread (a
    pencil over more than 64 threads, when n divides by 2E)."""
    e = elements_per_thread(n)
    wide = precision == "fp64" and n // e > 64 and n % (2 * e) == 0
    return not wide


def slot_bytes(n, precision="fp64"):
    """{slot name: algorithmic bytes of one launch}; slot = pass.s<stage>.

    The forward transform of the nonlinear term is needed only inside the
    dealiased box (|kx|,|ky|,kz < kmax); C writes n*n*nk, D reads n*n*nk and
    writes n*(2nk-1)*nk, E reads n*(2nk-1)*nk complex per field."""
    uc = n * n * (n // 2 + 1)
    rr = n ** 3
    nk = n_box(n)
    zc = n * n * nk                # after the z pass, kz < kmax
    yc = n * (2 * nk - 1) * nk     # after the y pass, also |ky| < kmax
    r = 8 if precision == "fp64" else 4      # real scalar in the working precision
    cr, csr, c64 = 2 * r, 2 * r, 16
    out = {}
    ax = 5 if curlx_in_y(n, precision) else 6     # fields A writes from stage 2 on
    for st in range(4):
        a = 3 * uc * c64 + 3 * uc * cr if st == 0 else 3 * uc * csr + ax * uc * cr
        b = (3 + 3 if st == 0 else ax + 6) * uc * cr
        if st == 0:
            c = 3 * uc * cr + 3 * rr * r + 3 * zc * cr
        elif st < 3:
            c = 6 * uc * cr + 3 * zc * cr
        else:
            c = 6 * uc * cr + 3 * zc * cr + 3 * rr * r
        d = 3 * zc * cr + 3 * yc * cr
        if precision == "fp64":
            if st == 0:
                e = 3 * yc * cr + 3 * uc * c64 + 3 * uc * csr + 3 * uc * c64
            elif st < 3:
                e = 3 * yc * cr + 3 * uc * csr + 6 * uc * c64 + 3 * uc * csr + 3 * uc * c64
            else:
                e = 3 * yc * cr + 3 * uc * csr + 3 * uc * c64 + 3 * uc * c64
        else:  # float copy of Un and running sum S (float; double in mixed_acc64); Un, Un+1 double
            c32 = 8
            cs = c64 if precision == "mixed_acc64" else c32
            if st == 0:    # r Un copy; w Ur(SR), S
                e = 3 * yc * cr + 3 * uc * c32 + 3 * uc * csr + 3 * uc * cs
            elif st < 3:   # r Ur, Un copy, S; w Ur, S
                e = 3 * yc * cr + 3 * uc * csr + 3 * uc * c32 + 3 * uc * cs + 3 * uc * csr + 3 * uc * cs
            else:          # r Ur, Un(f64), S; w Un+1(f64) and its float copy
                e = 3 * yc * cr + 3 * uc * csr + 3 * uc * c64 + 3 * uc * cs + 3 * uc * c64 + 3 * uc * c32
        # E fused with the next stage's A: A's writes without its read
        nxt_writes = 3 * uc * cr if st == 3 else ax * uc * cr
        for p, v in zip(PASSES, (a, b, c, d, e, e + nxt_writes)):
            out[f"{p}.s{st + 1}"] = v
    return out


def step_bytes_fused(n, precision="fp64"):
    """Bytes of one step as dns_step runs it in steady state: pass A is fused
    into the previous stage's E (the standalone A runs only after the state
    is set)."""
    sb = slot_bytes(n, precision)
    return sum(v for k, v in sb.items() if k.split(".")[0] in ("ypass_inv", "zfused", "ypass_fwd", "xfwd_rk_xinv"))


def step_bytes(n, precision="fp64"):
    return sum(slot_bytes(n, precision).values())


def summarize(timing, n, precision, steps):
    """Per pass: launches, average ms, algorithmic GB/s and fraction of peak,
    from {slot: (total ms, launches)} accumulated over `steps` steps."""
    sb = slot_bytes(n, precision)
    per_pass = {}
    for slot, (ms, cnt) in timing.items():
        if cnt == 0:
            continue
        p = slot.split(".")[0]
        d = per_pass.setdefault(p, {"ms": 0.0, "launches": 0, "bytes": 0})
        d["ms"] += ms
        d["launches"] += cnt
        d["bytes"] += sb[slot] * cnt
    for p, d in per_pass.items():
        d["avg_ms"] = d["ms"] / d["launches"]
        d["GBs"] = d["bytes"] / (d["ms"] * 1e-3) / 1e9
        d["frac"] = d["GBs"] / PEAK_HBM_GBS
        d["ms_per_step"] = d["ms"] / steps
        d["bytes_per_launch"] = d["bytes"] / d["launches"]
    return per_pass


def adjoint_launch_bytes(n):
    """Algorithmic bytes of the fused backward stage's kernels (FP64), per
    launch, keyed by role; a backward step launches xa once and yb, z, yd,
    xe four times each (roles z.r3 / z.r12 / z.r0 and xe.first / xe.next /
    xe.last by stage; yd moves six fields at stages 3..1, three at stage 0),
    after one recorded forward step (record: the forward step's bytes plus
    the 24 recorded physical fields).
      xa   r lamU 3U;                       w lamV 3U, lamN partials 3Yc
      yb   r 3Yc (box rows);                w 3Zc (all rows, kz < kmax)
      z    r 3Zc + u, omega 6R (+ lamS 3R at stage 3);
           w lam_omega, lam_u 6U (stage 0: 3U + lamS' 3R)
      yd   r/w 6U (stage 0: 3U)
      xe   r F 6U (stage 0: 3U), lamV 3U, lamUn 3U (+ lamU 3U between);
           w lamUn 3U (+ lamV 3U, lamN partials 3Yc when fused with xa)"""
    uc = n * n * (n // 2 + 1) * 16
    rr = n ** 3 * 8
    nk = n_box(n)
    zc = n * n * nk * 16
    yc = n * (2 * nk - 1) * nk * 16
    return {
        "xa": 3 * uc + 3 * uc + 3 * yc,
        "yb": 3 * yc + 3 * zc,
        "z.r3": 3 * zc + 6 * rr + 3 * rr + 6 * uc,
        "z.r12": 3 * zc + 6 * rr + 6 * uc,
        "z.r0": 3 * zc + 6 * rr + 3 * uc + 3 * rr,
        "yd6": 12 * uc,
        "yd3": 6 * uc,
        "xe.first": 6 * uc + 3 * uc + 3 * uc + 3 * uc + 3 * uc + 3 * yc,
        "xe.next": 6 * uc + 3 * uc + 3 * uc + 3 * uc + 3 * uc + 3 * uc + 3 * yc,
        "xe.last": 3 * uc + 3 * uc + 3 * uc + 3 * uc,
        "record": step_bytes_fused(n, "fp64") + 24 * rr,
    }


def adjoint_step_bytes(n):
    """Bytes of one backward step (recorded forward step and four fused
    backward stages)."""
    b = adjoint_launch_bytes(n)
    return (b["record"] + b["xa"] + 4 * b["yb"] + b["z.r3"] + 2 * b["z.r12"] + b["z.r0"] + 3 * b["yd6"] + b["yd3"]
            + b["xe.first"] + 2 * b["xe.next"] + b["xe.last"])
