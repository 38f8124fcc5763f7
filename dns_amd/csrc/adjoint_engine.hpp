// Discrete adjoint of the RK4 step (FP64): the checkpointed forward and
// backward sweeps for the gradient of the final enstrophy with respect to the
// initial velocity, and the adjoint of one step.  Included by engine.hpp; the
// algebra is documented in adjoint.hpp, the fused backward passes in
// adjoint_fused.hpp.
//
// A backward step needs, per stage r, the physical velocity u_r and vorticity
// omega_r of the forward step; the forward recomputation records them
// (24 physical fields per step: k_zfused with a record pointer).  A segment of
// `every` steps is regenerated from its checkpoint once, recording every step,
// when the recordings fit in device memory; otherwise the segment's states are
// kept and each step is recomputed with recording right before its backward
// step.
#pragma once
#include "adjoint.hpp"
#include "adjoint_fused.hpp"

namespace dnsk {

// Device workspace of the adjoint, allocated on first use.
struct AdjWork {
  size_t C3 = 0, R3 = 0, REC = 0;  // bytes of 3 spectral / 3 physical fields / one step's recording
  double2 *lamU = nullptr, *lamUn = nullptr, *lamV = nullptr, *tmpN = nullptr, *scr = nullptr;
  double *lamS = nullptr, *lamSout = nullptr, *lu = nullptr;
  std::vector<void *> ck;    // checkpoints: (spectral 3, physical 3) per segment start
  std::vector<void *> recs;  // recordings of the steps of one segment (or one step)
  std::vector<void *> seg;   // step states of a segment (only without a recording per step)
  void *one = nullptr;       // (spectral 3, physical 3) of dns_adjoint_step
  long nsteps = 0, every = 0;
  std::vector<void *> all;
  ~AdjWork() {
    for (void *p : all) (void)hipFree(p);
  }
  int get(void **p, size_t bytes) {
    DNS_HIP(hipMalloc(p, bytes));
    DNS_HIP(hipMemset(*p, 0, bytes));
    all.push_back(*p);
    return DNS_OK;
  }
  // like get, but an allocation failure is not an error: *p stays null
  void try_get(void **p, size_t bytes) {
    *p = nullptr;
    if (hipMalloc(p, bytes) != hipSuccess) {
      (void)hipGetLastError();
      *p = nullptr;
      return;
    }
    all.push_back(*p);
  }
};

template <int N> struct Adjoint {
  static constexpr int E = Sched<N>::E, TPP = N / E;
  static constexpr int T = 8 * TPP > 512 ? 512 / TPP : 8;
  static constexpr int PPW = z_pencils_per_block<N>();
  static constexpr int NF = N / 2 + 1;
  static dim3 gx() { return dim3((NF + T - 1) / T, N); }

  static int workspace(dns_ctx *c, AdjWork *&w) {
    if (c->adj) {
      w = (AdjWork *)c->adj;
      return DNS_OK;
    }
    w = new AdjWork();
    c->adj = w;
    w->C3 = 3 * sizeof(double2) * c->g.fs;
    w->R3 = 3 * sizeof(double) * c->g.rs;
    w->REC = 8 * w->R3;  // 4 stages x (u, omega)
    int rc = 0;
    double2 **cs[] = {&w->lamU, &w->lamUn, &w->lamV, &w->tmpN, &w->scr};
    for (auto p : cs)
      if (!rc) rc = w->get((void **)p, w->C3);
    double **rs[] = {&w->lamS, &w->lamSout, &w->lu};
    for (auto p : rs)
      if (!rc) rc = w->get((void **)p, w->R3);
    return rc;
  }

  // phys[f] = scale * C2R(spec[f]), f < 3; scr: 3 spectral fields of scratch.
  static int c2r3(dns_ctx *c, const double2 *spec, double *phys, double scale, double2 *scr) {
    hipStream_t s = c->stream;
    const cx<double> *tw = (const cx<double> *)c->tw64;
    k_xplain<N, double, T, true, double, true><<<gx(), T * TPP, 0, s>>>(spec, scr, 3, c->g, tw);
    DNS_LAUNCHED();
    k_ypass<N, double, T, true><<<gx(), T * TPP, 0, s>>>(scr, 3, c->g, tw);
    DNS_LAUNCHED();
    for (int f = 0; f < 3; f++) {
      k_zc2r<N, double, PPW><<<N * N / PPW, PPW * TPP, 0, s>>>(scr + f * c->g.fs, phys + f * c->g.rs, c->g, scale,
                                                              tw);
      DNS_LAUNCHED();
    }
    return DNS_OK;
  }

  static int copy(dns_ctx *c, void *dst, const void *src, size_t bytes) {
    DNS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->stream));
    return DNS_OK;
  }

  // One forward step from the context's state, recording u_r, omega_r.
  static int record(dns_ctx *c, void *rec) {
    c->rec_phys = (double *)rec;
    int rc = c->impl->step(c, 1);
    c->rec_phys = nullptr;
    return rc;
  }

  // Backward stages of one step from its recording: (lamU, lamS) = adjoints of
  // (U_{n+1}, u_stale') in, adjoints of (U_n, u_stale) out.
  static int backward_stages(dns_ctx *c, AdjWork *w, const double *rec) {
    static const double a_rk[4] = {1 / 6.0, 1 / 3.0, 1 / 3.0, 1 / 6.0};  // main.c:15
    static const double b_rk[4] = {0.5, 0.5, 1.0, 0.0};                 // main.c:16
    hipStream_t s = c->stream;
    const Geo g = c->g;
    const double dt = c->dt[0], nu = c->nu[0], kmax = c->kmax, sc = c->invn3;
    const double2 *tw = (const double2 *)c->tw64;
    double2 *W = (double2 *)c->W;
    const dim3 gxa = gx(), bxa(T * TPP);
    const dim3 gyt((g.nk + T - 1) / T, N), gyf = gx(), by(T * TPP);
    const dim3 gz(N * N / PPW), bz(PPW * TPP);
    auto stage = [&](int r) { return AdjStage{a_rk[r], b_rk[r], dt, nu, kmax}; };
    c->w_ready = 0;  // W holds the adjoint's partials from here on
    k_adj_xa<N, T><<<gxa, bxa, 0, s>>>(w->lamU, w->lamV, W, g, stage(3), tw);
    DNS_LAUNCHED();
    for (int r = 3; r >= 0; r--) {
      const double *rr = rec + (size_t)r * 6 * g.rs;
      k_ypass<N, double, T, true, 1, 0, true, E, false, true><<<gyt, by, 0, s>>>(W, 3, g, tw);
      DNS_LAUNCHED();
      if (r == 3)
        k_adj_z<N, PPW, false, true><<<gz, bz, 0, s>>>(W, rr, w->lamS, w->lamSout, g, sc, tw);
      else if (r > 0)
        k_adj_z<N, PPW, false, false><<<gz, bz, 0, s>>>(W, rr, w->lamS, w->lamSout, g, sc, tw);
      else
        k_adj_z<N, PPW, true, false><<<gz, bz, 0, s>>>(W, rr, w->lamS, w->lamSout, g, sc, tw);
      DNS_LAUNCHED();
      k_ypass<N, double, T, false, 1, 0, false, E, false, true><<<gyf, by, 0, s>>>(W, r > 0 ? 6 : 3, g, tw);
      DNS_LAUNCHED();
      switch (r) {
        case 3: k_adj_xe<N, T, true, true, true><<<gxa, bxa, 0, s>>>(W, w->lamU, w->lamUn, w->lamV, g, stage(2), tw); break;
        case 2: k_adj_xe<N, T, true, false, true><<<gxa, bxa, 0, s>>>(W, w->lamU, w->lamUn, w->lamV, g, stage(1), tw); break;
        case 1: k_adj_xe<N, T, true, false, true><<<gxa, bxa, 0, s>>>(W, w->lamU, w->lamUn, w->lamV, g, stage(0), tw); break;
        default: k_adj_xe<N, T, false, false, false><<<gxa, bxa, 0, s>>>(W, w->lamU, w->lamUn, w->lamV, g, stage(0), tw); break;
      }
      DNS_LAUNCHED();
    }
    std::swap(w->lamU, w->lamUn);
    std::swap(w->lamS, w->lamSout);
    return DNS_OK;
  }

  static void *state_alloc(AdjWork *w) {
    void *p = nullptr;
    if (w->get(&p, w->C3 + w->R3)) return nullptr;
    return p;
  }

  static int save_state(dns_ctx *c, AdjWork *w, void *dst) {
    int rc = copy(c, dst, c->uhn, w->C3);
    return rc ? rc : copy(c, (char *)dst + w->C3, c->us, w->R3);
  }

  static int restore_state(dns_ctx *c, AdjWork *w, const void *src) {
    int rc = copy(c, c->uhn, src, w->C3);
    if (!rc) rc = copy(c, c->us, (const char *)src + w->C3, w->R3);
    c->w_ready = 0;
    return rc;
  }

  // Forward sweep from u0 with checkpoints every `every` steps; J = final
  // enstrophy sum; seeds lamU = dJ/dU_final, lamS = 0.  Also sizes the
  // recordings of a segment: `every` of them if they fit, else one, with the
  // segment's states kept instead.
  static int forward(dns_ctx *c, const double *u0, long nsteps, long every, double *J) {
    (void)J;
    AdjWork *w;
    int rc = workspace(c, w);
    if (rc) return rc;
    if (every < 1) every = 1;
    const long nck = (nsteps + every - 1) / every;
    while ((long)w->ck.size() < nck) {
      void *p = state_alloc(w);
      if (!p) return DNS_ENOMEM;
      w->ck.push_back(p);
    }
    const long want = std::min(every, nsteps);
    while ((long)w->recs.size() < want) {
      void *p = nullptr;
      w->try_get(&p, w->REC);
      if (!p) break;
      w->recs.push_back(p);
    }
    if (w->recs.empty()) return DNS_ENOMEM;
    if ((long)w->recs.size() < want) {  // fallback: states of the segment, one recording
      while ((long)w->seg.size() < every) {
        void *p = state_alloc(w);
        if (!p) return DNS_ENOMEM;
        w->seg.push_back(p);
      }
    }
    w->nsteps = nsteps;
    w->every = every;
    if ((rc = c->impl->load(c, 0, u0))) return rc;
    for (long i = 0; i < nsteps; i++) {
      if (i % every == 0 && (rc = save_state(c, w, w->ck[i / every]))) return rc;
      if ((rc = c->impl->step(c, 1))) return rc;
    }
    hipStream_t s = c->stream;
    k_enstrophy_grad<<<N * N, 128, 0, s>>>((const double2 *)c->uhn, w->lamU, N, c->g, c->invn3 * c->invn3);
    DNS_LAUNCHED();
    DNS_HIP(hipMemsetAsync(w->lamS, 0, w->R3, s));
    DNS_HIP(hipStreamSynchronize(s));
    return DNS_OK;
  }

  // Backward sweep over the checkpointed trajectory; grad = seed * dJ/du0.
  static int backward(dns_ctx *c, double seed, double *grad) {
    AdjWork *w = (AdjWork *)c->adj;
    if (!w || w->nsteps <= 0) return DNS_EINVAL;
    hipStream_t s = c->stream;
    int rc;
    const long nsteps = w->nsteps, every = w->every;
    const bool per_segment = (long)w->recs.size() >= std::min(every, nsteps);
    for (long s0 = ((nsteps - 1) / every) * every; s0 >= 0; s0 -= every) {
      const long s1 = std::min(nsteps, s0 + every);
      if ((rc = restore_state(c, w, w->ck[s0 / every]))) return rc;
      if (per_segment) {
        // regenerate the segment once, recording every step
        for (long i = s0; i < s1; i++)
          if ((rc = record(c, w->recs[i - s0]))) return rc;
        for (long i = s1 - 1; i >= s0; i--)
          if ((rc = backward_stages(c, w, (const double *)w->recs[i - s0]))) return rc;
      } else {
        for (long i = s0; i < s1; i++) {
          if ((rc = save_state(c, w, w->seg[i - s0]))) return rc;
          if (i + 1 < s1 && (rc = c->impl->step(c, 1))) return rc;
        }
        for (long i = s1 - 1; i >= s0; i--) {
          if ((rc = restore_state(c, w, w->seg[i - s0])) || (rc = record(c, w->recs[0]))) return rc;
          if ((rc = backward_stages(c, w, (const double *)w->recs[0]))) return rc;
        }
      }
    }
    // u0 enters as U_0 = R2C(u0) and u_stale = u0: grad = C2R(W^-1 lamU) + lamS
    k_hweight_inv<<<N * N, 128, 0, s>>>(w->lamU, w->tmpN, N, c->g);
    DNS_LAUNCHED();
    if ((rc = c2r3(c, w->tmpN, w->lu, seed, w->scr))) return rc;
    k_axpy_scaled<<<4096, 256, 0, s>>>(w->lamS, w->lu, seed, 3 * c->g.rs);
    DNS_LAUNCHED();
    DNS_HIP(hipMemcpyAsync(grad, w->lu, w->R3, hipMemcpyDefault, s));
    DNS_HIP(hipStreamSynchronize(s));
    w->nsteps = 0;  // the sweep consumed the trajectory
    c->w_ready = 0;
    return DNS_OK;
  }
};

}  // namespace dnsk
