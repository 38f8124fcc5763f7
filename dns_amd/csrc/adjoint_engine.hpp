// Discrete adjoint of the RK4 step (FP64): transforms, one backward step,
// and the checkpointed forward/backward sweeps for the gradient of the
// final enstrophy with respect to the initial velocity.  Included by
// engine.hpp; the algebra is documented in adjoint.hpp.
#pragma once
#include "adjoint.hpp"

namespace dnsk {

// Device workspace of the adjoint, allocated on first use.
struct AdjWork {
  size_t C3 = 0, R3 = 0;  // bytes of 3 spectral / 3 physical fields
  double2 *S[4] = {};     // stage states U_0 (= U_n) .. U_3
  double *Us = nullptr;   // stale velocity at the step start
  double2 *lamU = nullptr, *lamNext = nullptr, *lamUr = nullptr, *lamUn = nullptr, *tmpN = nullptr,
          *tmpC = nullptr, *Fu = nullptr, *Fw = nullptr, *scr = nullptr;
  double *lamS = nullptr, *lamSout = nullptr, *lamN = nullptr, *u = nullptr, *w = nullptr, *lu = nullptr,
         *lw = nullptr;
  std::vector<void *> ck;   // checkpoints: (spectral 3, physical 3) per segment start
  std::vector<void *> seg;  // states of the steps of one segment
  long nsteps = 0, every = 0;
  std::vector<void *> all;
  ~AdjWork() {
    for (void *p : all) hipFree(p);
  }
  int get(void **p, size_t bytes) {
    DNS_HIP(hipMalloc(p, bytes));
    DNS_HIP(hipMemset(*p, 0, bytes));
    all.push_back(*p);
    return DNS_OK;
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
    int rc = 0;
    for (int r = 0; r < 4 && !rc; r++) rc = w->get((void **)&w->S[r], w->C3);
    double2 **cs[] = {&w->lamU, &w->lamNext, &w->lamUr, &w->lamUn, &w->tmpN, &w->tmpC, &w->Fu, &w->Fw, &w->scr};
    for (auto p : cs)
      if (!rc) rc = w->get((void **)p, w->C3);
    double **rs[] = {&w->Us, &w->lamS, &w->lamSout, &w->lamN, &w->u, &w->w, &w->lu, &w->lw};
    for (auto p : rs)
      if (!rc) rc = w->get((void **)p, w->R3);
    return rc;
  }

  // Every spectral array of the adjoint is in the state layout.
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

  // spec[f] = R2C(phys[f]), f < 3; scr: 3 spectral fields of scratch.
  static int r2c3(dns_ctx *c, const double *phys, double2 *spec, double2 *scr) {
    hipStream_t s = c->stream;
    const cx<double> *tw = (const cx<double> *)c->tw64;
    k_zr2c<N, double, PPW><<<N * N / PPW, PPW * TPP, 0, s>>>(phys, scr, 3, c->g, tw);
    DNS_LAUNCHED();
    k_ypass<N, double, T, false><<<gx(), T * TPP, 0, s>>>(scr, 3, c->g, tw);
    DNS_LAUNCHED();
    k_xplain<N, double, T, false, double, false, true><<<gx(), T * TPP, 0, s>>>(scr, spec, 3, c->g, tw);
    DNS_LAUNCHED();
    return DNS_OK;
  }

  static int copy(dns_ctx *c, void *dst, const void *src, size_t bytes) {
    DNS_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->stream));
    return DNS_OK;
  }

  // Set the context state to (U, us) and run one forward step recording the
  // stage states U_1..U_3 in w->S[1..3] (w->S[0] = U).
  static int record_step(dns_ctx *c, AdjWork *w, const double2 *U, const double *us) {
    int rc = copy(c, c->uhn, U, w->C3);
    if (!rc) rc = copy(c, c->us, us, w->R3);
    if (!rc) rc = copy(c, w->S[0], U, w->C3);
    if (rc) return rc;
    c->w_ready = 0;
    c->rec[0] = w->S[1];
    c->rec[1] = w->S[2];
    c->rec[2] = w->S[3];
    rc = c->impl->step(c, 1);
    c->rec[0] = c->rec[1] = c->rec[2] = nullptr;
    return rc;
  }

  // One backward step: (w->lamU, w->lamS) = adjoints of (U_{n+1}, u_stale')
  // in, adjoints of (U_n, u_stale) out; the step starts from (U, us).
  static int backward_step(dns_ctx *c, AdjWork *w, const double2 *U, const double *us) {
    static const double a_rk[4] = {1 / 6.0, 1 / 3.0, 1 / 3.0, 1 / 6.0};  // main.c:15
    static const double b_rk[4] = {0.5, 0.5, 1.0, 0.0};                 // main.c:16
    hipStream_t s = c->stream;
    const Geo g = c->g;
    const double sc = c->invn3, nu = c->nu[0], dt = c->dt[0];
    int rc = record_step(c, w, U, us);
    if (!rc) rc = copy(c, w->lamUn, w->lamU, w->C3);
    if (rc) return rc;
    const dim3 grow(N * N), brow(128);
    const int gp = 4096, bp = 256;
    for (int r = 3; r >= 0; r--) {
      k_adj_spec_pre<<<grow, brow, 0, s>>>(w->lamU, r < 3 ? w->lamNext : nullptr, w->tmpN, w->lamUr, N, g,
                                          a_rk[r], b_rk[r], dt, nu, c->kmax);
      DNS_LAUNCHED();
      if ((rc = c2r3(c, w->tmpN, w->lamN, 1.0, w->scr))) return rc;
      const double *ur = us;
      if (r > 0) {
        if ((rc = c2r3(c, w->S[r], w->u, sc, w->scr))) return rc;
        ur = w->u;
      }
      k_curl_spec<<<grow, brow, 0, s>>>(w->S[r], w->tmpC, N, g);
      DNS_LAUNCHED();
      if ((rc = c2r3(c, w->tmpC, w->w, sc, w->scr))) return rc;
      k_adj_cross<<<gp, bp, 0, s>>>(w->lamN, ur, w->w, r == 3 ? w->lamS : nullptr, w->lu, w->lw, g.rs,
                                    r == 0 ? 1.0 : sc, sc);
      DNS_LAUNCHED();
      if ((rc = r2c3(c, w->lw, w->Fw, w->scr))) return rc;
      if (r > 0 && (rc = r2c3(c, w->lu, w->Fu, w->scr))) return rc;
      k_adj_spec_post<<<grow, brow, 0, s>>>(w->Fw, r > 0 ? w->Fu : nullptr, w->lamUr, w->lamUn, N, g);
      DNS_LAUNCHED();
      if (r == 0 && (rc = copy(c, w->lamSout, w->lu, w->R3))) return rc;
      std::swap(w->lamNext, w->lamUr);
    }
    std::swap(w->lamU, w->lamUn);
    std::swap(w->lamS, w->lamSout);
    return DNS_OK;
  }

  static void *state_alloc(dns_ctx *c, AdjWork *w) {
    void *p = nullptr;
    if (w->get(&p, w->C3 + w->R3)) return nullptr;
    (void)c;
    return p;
  }

  static int save_state(dns_ctx *c, AdjWork *w, void *dst) {
    int rc = copy(c, dst, c->uhn, w->C3);
    return rc ? rc : copy(c, (char *)dst + w->C3, c->us, w->R3);
  }

  // Forward sweep from u0 with checkpoints every `every` steps; J = final
  // enstrophy sum; seeds lamU = dJ/dU_final, lamS = 0.
  static int forward(dns_ctx *c, const double *u0, long nsteps, long every, double *J) {
    AdjWork *w;
    int rc = workspace(c, w);
    if (rc) return rc;
    if (every < 1) every = 1;
    const long nck = (nsteps + every - 1) / every;
    while ((long)w->ck.size() < nck) {
      void *p = state_alloc(c, w);
      if (!p) return DNS_ENOMEM;
      w->ck.push_back(p);
    }
    while ((long)w->seg.size() < every) {
      void *p = state_alloc(c, w);
      if (!p) return DNS_ENOMEM;
      w->seg.push_back(p);
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
    for (long s0 = ((nsteps - 1) / every) * every; s0 >= 0; s0 -= every) {
      const long s1 = std::min(nsteps, s0 + every);
      // regenerate the states of steps s0..s1-1 from the checkpoint
      void *ck = w->ck[s0 / every];
      if ((rc = copy(c, c->uhn, ck, w->C3)) || (rc = copy(c, c->us, (char *)ck + w->C3, w->R3))) return rc;
      c->w_ready = 0;
      for (long i = s0; i < s1; i++) {
        if ((rc = save_state(c, w, w->seg[i - s0]))) return rc;
        if (i + 1 < s1 && (rc = c->impl->step(c, 1))) return rc;
      }
      for (long i = s1 - 1; i >= s0; i--) {
        void *st = w->seg[i - s0];
        if ((rc = backward_step(c, w, (const double2 *)st, (const double *)((char *)st + w->C3)))) return rc;
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
