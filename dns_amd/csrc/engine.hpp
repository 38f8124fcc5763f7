// Context and per-grid-size engine of libdns_hip.
//
// A context holds M >= 1 simulations ("members") of the same grid size n
// on one stream; every buffer below is repeated per member (member stride
// in Geo) and every step kernel covers all members through blockIdx.z.
// The single-simulation API uses M = 1; a batch (dns_batch_*) owns one
// context per distinct n.
//
// Device memory plan per member (U = one spectral field, n*n*nfp complex;
// R = one physical field, n^3 reals):
//   uhn  3U double  spectral state at the start of the step (Û_n); stage 4
//                   writes Û_{n+1} here
//   uhr  3U SR      stage state Û_r (the argument of stages 2..4)
//   acc  3U double  RK4 accumulator (Û1 of main.c)
//   W    6U R       FFT partials: u,v,w in fields 0..2, omega in 3..5;
//                   N = u x omega in 0..2 after the fused z pass; each
//                   x-plane holds one spare row (Geo::wps, Geo::wfs).  Also
//                   the scratch of the initial transform and of snapshots
//                   (plain layout, fields fs apart).
//   us   3R R       stale physical velocity (main.c keeps U,V,W from the
//                   previous stage 4 for stage 1, main.c:344-353)
//   phat 1U double  stage-4 pressure, allocated on first capture request
// FP64: R = SR = double.  Mixed: R = SR = float.
#pragma once
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>
#include "../../include/dns.h"
#include "kernels.hpp"

namespace dnsk {

void set_error(const char *what, hipError_t e);
void set_error_text(const char *what);

#define DNS_HIP(x)                                                   \
  do {                                                               \
    hipError_t e_ = (x);                                             \
    if (e_ != hipSuccess) {                                          \
      ::dnsk::set_error(#x, e_);                                     \
      return e_ == hipErrorOutOfMemory ? DNS_ENOMEM : DNS_EHIP;      \
    }                                                                \
  } while (0)

#define DNS_LAUNCHED() DNS_HIP(hipGetLastError())

struct Impl;

}  // namespace dnsk

// Kernel timing slots: DNS_NPASS * stage + pass; passes 0..4 = A..E of
// kernels.hpp, 5 = E fused with the next stage's A.
enum { DNS_NPASS = 6, DNS_NSLOT = 4 * DNS_NPASS };

struct dns_ctx {
  int n, nf, nfp, prec, device, members;
  double kmax, invn3;
  std::vector<double> nu, dt;  // per member
  hipStream_t stream;
  dnsk::Geo g;
  void *uhn, *uhr, *acc, *W, *us, *phat;
  void *tw64, *tw32;
  double2 *nudt;  // device copy of (nu, dt) per member
  double *part, *dres;
  int capture, have_phat;
  // W holds the omega partials of the next step's stage 1 (written by the
  // fused E+A of the last stage 4); cleared whenever W or the state change
  int w_ready;
  long long bytes;
  dnsk::Impl *impl;
  // when set, step() records the physical velocity and vorticity of every
  // stage of a one-step call: stage r into rec_phys + 6 r n^3, fields
  // (u, v, w, omega_x, omega_y, omega_z) (the adjoint's recomputation, FP64)
  double *rec_phys;
  void *adj;  // adjoint workspace (dnsk::AdjWork), allocated on first use
  int variant;            // launch shape (see Engine::step_launch)
  int use_graph;          // replay steps from a captured graph
  // graphs of GRAPH_STEPS steps and of one step for launch shape `gvariant`
  hipGraphExec_t gexec, gexec1;
  int gvariant;
  // HIP-event timing of the step's kernels on `stream` (off by default)
  int timing;
  std::vector<hipEvent_t> ev_pool;
  struct TimedLaunch {
    int slot;
    hipEvent_t a, b;
  };
  std::vector<TimedLaunch> ev_pending;
  double slot_ms[DNS_NSLOT];
  long slot_n[DNS_NSLOT];
  // asynchronous snapshots: device staging [4][n][n][n] doubles, its pinned
  // host copy, the copy stream and the events that order them (allocated by
  // the first dns_snapshot_begin)
  double *snap_dev, *snap_host;
  hipStream_t snap_stream;
  hipEvent_t snap_ready, snap_done;
  int snap_pending;
};

namespace dnsk {

struct Impl {
  virtual ~Impl() {}
  virtual size_t sr_size() const = 0;  // bytes per complex element of uhr
  virtual size_t acc_bytes(long fs) const = 0;
  virtual size_t r_size() const = 0;   // bytes per real element of W/us
  virtual int make_twiddles(dns_ctx *c) = 0;
  virtual int load(dns_ctx *c, int m, const double *uvw) = 0;
  // Pass A of stage 1 for every member unless the partials are ready.
  virtual int prepare(dns_ctx *c) = 0;
  // nsteps steps of all members; with `ready` the partials of stage 1 are
  // taken as present (graph capture after prepare()).
  virtual int step(dns_ctx *c, long nsteps) = 0;
  virtual int variants() const = 0;
  virtual int set_variant(dns_ctx *c, int v) = 0;
  virtual int snapshot(dns_ctx *c, int m, double *uvwp) = 0;
  // c2r(U,V,W,P)/n^3 of member 0 into `dst` (device [4][n][n][n] doubles)
  // on the context's stream, asynchronously
  virtual int snapshot_device(dns_ctx *c, double *dst) = 0;
  virtual int rfft3(dns_ctx *c, const double *in, double *out) = 0;
  virtual int irfft3(dns_ctx *c, const double *in, double *out) = 0;
  virtual int get_us(dns_ctx *c, int m, double *out) = 0;
  virtual int set_us(dns_ctx *c, int m, const double *in) = 0;
  // After Un of member m (m < 0: all members) was set from outside: refresh
  // the float copy of Un that mixed precision keeps.
  virtual int state_changed(dns_ctx *c, int m) = 0;
  // spectral state of member m <-> host [3][n][n][n/2+1] complex (x-major)
  virtual int get_spectral(dns_ctx *c, int m, double *out) = 0;
  virtual int set_spectral(dns_ctx *c, int m, const double *in) = 0;
  // discrete adjoint (FP64 contexts, member 0)
  virtual int adj_forward(dns_ctx *c, const double *u0, long nsteps, long every) = 0;
  virtual int adj_backward(dns_ctx *c, double seed, double *grad) = 0;
  virtual int adj_step(dns_ctx *c, const double *uh, const double *us, const double *lam_uh,
                       const double *lam_us, double *out_uh, double *out_us) = 0;
};

// Event bracketing of one launch when timing is on.
inline hipEvent_t ev_get(dns_ctx *c) {
  if (!c->ev_pool.empty()) {
    hipEvent_t e = c->ev_pool.back();
    c->ev_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}
inline void t_begin(dns_ctx *c, int slot) {
  if (!c->timing) return;
  dns_ctx::TimedLaunch t{slot, ev_get(c), ev_get(c)};
  if (t.a) (void)hipEventRecord(t.a, c->stream);
  c->ev_pending.push_back(t);
}
inline void t_end(dns_ctx *c) {
  if (!c->timing || c->ev_pending.empty()) return;
  hipEvent_t b = c->ev_pending.back().b;
  if (b) (void)hipEventRecord(b, c->stream);
}
// roctx ranges (rocprofv3 --marker-trace) around every step and RK stage
// the host launches, when the environment sets DNS_ROCTX: the roctx
// library is opened at run time (roctx_api in dns_api.hip), so the library
// has no link dependency on it.  Under a captured graph the ranges mark the
// capture, not the replays.
struct RoctxApi {
  int (*push)(const char *) = nullptr;
  int (*pop)() = nullptr;
};
const RoctxApi &roctx_api();
struct RoctxRange {
  bool on;
  explicit RoctxRange(const char *name) : on(roctx_api().push != nullptr) {
    if (on) roctx_api().push(name);
  }
  ~RoctxRange() {
    if (on) roctx_api().pop();
  }
  RoctxRange(const RoctxRange &) = delete;
  RoctxRange &operator=(const RoctxRange &) = delete;
};

#define DNS_TIMED(c, slot, launch) \
  do {                             \
    t_begin(c, slot);              \
    launch;                        \
    t_end(c);                      \
    DNS_LAUNCHED();                \
  } while (0)

template <typename TI, typename TO>
__global__ void k_convert(const TI *__restrict__ in, TO *__restrict__ out, long count) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < count; i += (long)gridDim.x * blockDim.x)
    out[i] = (TO)in[i];
}

// Forward twiddles of the radix schedule of N: per stage s, entries
// k*(r-1) + m-1 = exp(-2 pi i m k / (Ns r)), computed in long double.
template <int N> std::vector<long double> twiddles_ld() {
  std::vector<long double> t(2 * tw_size<N>() + 2, 0.0L);
  const long double PI = 3.141592653589793238462643383279502884L;
  int off = 0, Ns = 1;
  for (int s = 0; s < Sched<N>::NS; s++) {
    const int r = Sched<N>::r[s];
    for (int k = 0; k < Ns; k++)
      for (int m = 1; m < r; m++) {
        const long double a = 2.0L * PI * (long double)(m * k) / (long double)(Ns * r);
        t[2 * (off + k * (r - 1) + m - 1)] = cosl(a);
        t[2 * (off + k * (r - 1) + m - 1) + 1] = -sinl(a);
      }
    off += Ns * (r - 1);
    Ns *= r;
  }
  return t;
}

}  // namespace dnsk
#include "adjoint_engine.hpp"
namespace dnsk {

template <int N, typename R, typename SR, bool S64 = false> struct Engine final : Impl {
  static constexpr int E = Sched<N>::E, TPP = N / E;
  // Column tiles of the strided passes: 128-byte row segments (8 double /
  // 16 float columns).  Blocks holding double state or the fused x pass
  // have at most 512 threads (256 registers per thread, no spills at
  // n=1024); the float x pass gets its 16 columns in 512 threads with 16
  // elements per thread where the schedule allows (powers of two from 256
  // on); float y passes use up to 1024 threads.
  static constexpr int T_D = 8 * TPP > 512 ? 512 / TPP : 8;
  // y passes: double grids whose pencils need more than 64 threads (n=1024)
  // take 16 elements per thread, keeping 8-column tiles in 512 threads
  // (n=1024: inverse 25.0 -> 22.5 ms, forward 7.7 -> 5.7 ms).
  static constexpr int E_Y = (sizeof(R) == 8 && TPP > 64 && N % (2 * E) == 0) ? 2 * E : E;
  static constexpr int TPP_Y = N / E_Y;
  // 8 double / 16 float columns (float y passes use up to 1024 threads;
  // narrower float tiles over-fetch: 64-byte segments read 1.8x their bytes)
  static constexpr int T_Y = sizeof(R) == 8 ? (8 * TPP_Y > 512 ? 512 / TPP_Y : 8)
                                            : (16 * TPP_Y > 1024 ? 1024 / TPP_Y : 16);
  static constexpr int T_YF = T_Y;
  // fields loaded ahead: inverse y pass two in float (1.47 -> 1.38 ms at
  // n=512), one in double (a second costs occupancy), none with 16-element
  // threads (register budget); the forward y pass one in double only
  static constexpr int PIPE_Y = sizeof(R) == 8 ? (E_Y > E ? 0 : 1) : 2;
  static constexpr int PIPE_YF = sizeof(R) == 8 ? (E_Y > E ? 0 : 1) : 0;
  // the six-field inverse pass (curlX formed in registers) holds one more
  // array and loads no field ahead: in float two ahead spill at 1024
  // threads and one leaves a small spill (1.14 against 1.13 ms at n=512);
  // in double one ahead costs 192 against 160 VGPRs (2.51 against 2.47 ms)
  static constexpr int PIPE_YC = 0;
  // curlX formed by the inverse y pass, except where that pass runs 16
  // double elements per thread (n=1024): there its extra array spills
  // (348 bytes; 16.1 -> 21.5 ms), so the x pass transforms and stores curlX
  static constexpr bool CX_IN_Y = !(sizeof(R) == 8 && E_Y > E);
  // plain transforms of the working precision (rfft3 / irfft3)
  static constexpr int T_P = sizeof(R) == 8 ? T_D : (16 * TPP > 1024 ? 1024 / TPP : 16);
  static constexpr int E_X = (sizeof(R) == 4 && E == 8 && N >= 256) ? 16 : E;
  static constexpr int TPP_X = N / E_X;
  // Epilogue batching of the x pass, stages 1-3 (EPI_X) and 4 (EPI_X3): the
  // float epilogue of the mixed state leaves room for four positions per
  // load batch with 16-element threads (n=512: 1.94 -> 1.86 ms in stage 1,
  // 2.66 -> 2.64 in stages 2-3, 2.99 -> 2.95 in stage 4; prefetching the
  // next batch instead costs more).  The double epilogue keeps the default.
  static constexpr int EPI_X = (sizeof(SR) == 4 && E_X == 16) ? 4 : 0;
  static constexpr int EPI_X3 = EPI_X;
  static constexpr int T_X = sizeof(R) == 8 ? T_D : (16 * TPP_X > 512 ? 512 / TPP_X : 16);
  // z pencils per block: at most one wave (or one pencil when a pencil needs
  // more than a wave), so the pass's barriers never couple independent
  // pencils across waves.
  static constexpr int PPW = z_pencils_per_block<N>();
  static constexpr int NF = N / 2 + 1;
  // recording kernels exist for FP64 only (the adjoint's precision)
  static constexpr bool REC_OK = sizeof(R) == 8;
  // minimum waves per SIMD the z pass's registers must allow.  Float with 8
  // elements per thread and a pencil in at most one wave (n = 32..512): 4
  // (at most 128 VGPRs instead of 136-152, spilling at most 28 bytes per
  // lane), 0.982 -> 0.911 ms per launch at n=512 and 0.128 -> 0.119 ms at
  // n=256 in alternating same-box runs.  Not with 12 or 6 elements per
  // thread (the radix-3 grids), where that bound spills 84-616 and up to 72
  // bytes per lane.  FP64 keeps 1 (n=1024: 2 spills 92-124 bytes per lane
  // in stages 2-4 and measured no faster, 20.4 against 20.0-20.2 ms).
  static constexpr int MINW_Z = (sizeof(R) == 4 && TPP <= 64 && E == 8) ? 4 : 1;

  size_t sr_size() const override { return sizeof(cx<SR>); }
  // bytes per member of the accumulator region: 3 double fields, or in
  // mixed precision the float copy of Un (3 float fields) and the running
  // sum of the increments (3 float fields, or 3 double fields with S64)
  size_t acc_bytes(long fs) const override {
    return sizeof(SR) == 4 ? (3 * sizeof(float2) + 3 * (S64 ? sizeof(double2) : sizeof(float2))) * fs
                           : 3 * sizeof(double2) * fs;
  }
  size_t r_size() const override { return sizeof(R); }

  static const cx<R> *twR(dns_ctx *c) { return (const cx<R> *)(sizeof(R) == 8 ? c->tw64 : c->tw32); }
  static const cx<double> *twD(dns_ctx *c) { return (const cx<double> *)c->tw64; }

  dim3 grid_x(int T, int m = 1) const { return dim3((NF + T - 1) / T, N, m); }
  // grid of the x-pass kernels (k_xinv_curl, k_xfwd_rk_xinv): one-dimensional
  // where x_tile pairs the tiles of a line
  template <int T> dim3 grid_xp(int m = 1) const {
    if constexpr (x_tiles_paired<T, R>()) return dim3(paired_blocks((NF + T - 1) / T, N), 1, m);
    else return grid_x(T, m);
  }

  // Member m's slice of the W scratch, as bytes.
  static char *wslice(dns_ctx *c, int m) { return (char *)c->W + (size_t)m * c->g.mW * sizeof(cx<R>); }

  int make_twiddles(dns_ctx *c) override {
    auto t = twiddles_ld<N>();
    std::vector<double> d(t.size());
    std::vector<float> f(t.size());
    for (size_t i = 0; i < t.size(); i++) {
      d[i] = (double)t[i];
      f[i] = (float)t[i];  // single rounding from the long double value
    }
    DNS_HIP(hipMalloc(&c->tw64, d.size() * sizeof(double)));
    DNS_HIP(hipMemcpy(c->tw64, d.data(), d.size() * sizeof(double), hipMemcpyHostToDevice));
    DNS_HIP(hipMalloc(&c->tw32, f.size() * sizeof(float)));
    DNS_HIP(hipMemcpy(c->tw32, f.data(), f.size() * sizeof(float), hipMemcpyHostToDevice));
    c->bytes += (long long)(d.size() * sizeof(double) + f.size() * sizeof(float));
    return DNS_OK;
  }

  int state_changed(dns_ctx *c, int m) override {
    if constexpr (sizeof(SR) == 4) {
      const long m0 = m < 0 ? 0 : m, m1 = m < 0 ? c->members : m + 1;
      for (long k = m0; k < m1; k++) {
        const double *un = (const double *)((double2 *)c->uhn + k * c->g.mS);
        float *cp = (float *)((float2 *)c->acc + k * (S64 ? 3 : 2) * c->g.mS);
        k_convert<double, float><<<4096, 256, 0, c->stream>>>(un, cp, 2 * c->g.mS);
        DNS_LAUNCHED();
      }
    }
    return DNS_OK;
  }

  int load(dns_ctx *c, int m, const double *uvw) override {
    c->w_ready = 0;
    hipStream_t s = c->stream;
    double *tmp = (double *)wslice(c, m);  // 3 physical double fields fit in a member's W
    const size_t rb = sizeof(double) * c->g.rs;
    DNS_HIP(hipMemcpyAsync(tmp, uvw, 3 * rb, hipMemcpyDefault, s));
    R *us = (R *)c->us + m * c->g.mR;
    if (sizeof(R) == 8) {
      DNS_HIP(hipMemcpyAsync(us, tmp, 3 * rb, hipMemcpyDeviceToDevice, s));
    } else {
      k_convert<double, R><<<4096, 256, 0, s>>>(tmp, us, 3 * c->g.rs);
      DNS_LAUNCHED();
    }
    // partials in the member's uhn slice, result in its W slice (the
    // physical input is dead by then), then copied into uhn
    double2 *uh = (double2 *)c->uhn + m * c->g.mS;
    double2 *ws = (double2 *)wslice(c, m);
    k_zr2c<N, double, PPW><<<N * N / PPW, PPW * TPP, 0, s>>>(tmp, uh, 3, c->g, twD(c));
    DNS_LAUNCHED();
    k_ypass<N, double, T_D, false><<<grid_x(T_D), T_D * TPP, 0, s>>>(uh, 3, c->g, twD(c));
    DNS_LAUNCHED();
    k_xplain<N, double, T_D, false, double, false, true><<<grid_x(T_D), T_D * TPP, 0, s>>>(uh, ws, 3, c->g, twD(c));
    DNS_LAUNCHED();
    DNS_HIP(hipMemcpyAsync(uh, ws, 3 * sizeof(double2) * c->g.fs, hipMemcpyDeviceToDevice, s));
    int rc = state_changed(c, m);
    if (rc) return rc;
    DNS_HIP(hipStreamSynchronize(s));
    return DNS_OK;
  }

  int get_spectral(dns_ctx *c, int m, double *out) override {
    c->w_ready = 0;
    hipStream_t s = c->stream;
    double2 *ws = (double2 *)wslice(c, m);
    k_state_rows<double2, false><<<N * N, 256, 0, s>>>((const double2 *)c->uhn + m * c->g.mS, ws, N, c->g, 3);
    DNS_LAUNCHED();
    DNS_HIP(hipMemcpy2DAsync(out, sizeof(double2) * NF, ws, sizeof(double2) * c->g.nfp, sizeof(double2) * NF,
                             (size_t)3 * N * N, hipMemcpyDeviceToHost, s));
    DNS_HIP(hipStreamSynchronize(s));
    return DNS_OK;
  }

  int set_spectral(dns_ctx *c, int m, const double *in) override {
    c->w_ready = 0;
    hipStream_t s = c->stream;
    double2 *ws = (double2 *)wslice(c, m);
    const size_t pitch = sizeof(double2) * c->g.nfp;
    DNS_HIP(hipMemset2DAsync(ws, pitch, 0, pitch, (size_t)3 * N * N, s));
    DNS_HIP(hipMemcpy2DAsync(ws, pitch, in, sizeof(double2) * NF, sizeof(double2) * NF, (size_t)3 * N * N,
                             hipMemcpyHostToDevice, s));
    k_state_rows<double2, true><<<N * N, 256, 0, s>>>(ws, (double2 *)c->uhn + m * c->g.mS, N, c->g, 3);
    DNS_LAUNCHED();
    int rc = state_changed(c, m);
    if (rc) return rc;
    DNS_HIP(hipStreamSynchronize(s));
    return DNS_OK;
  }

  int prepare(dns_ctx *c) override {
    if (c->w_ready) return DNS_OK;
    const dim3 gx = grid_xp<T_X>(c->members), bx(T_X * TPP_X);
    DNS_TIMED(c, 0, (k_xinv_curl<N, R, T_X, false, double, 1, E_X><<<gx, bx, 0, c->stream>>>(
                        (const double2 *)c->uhn, (cx<R> *)c->W, c->g, twR(c))));
    c->w_ready = 1;
    return DNS_OK;
  }

  int set_variant(dns_ctx *c, int v) override {
    if (v < 0 || v >= variants()) return DNS_EINVAL;
    c->variant = v;
    return DNS_OK;
  }

  // Launch shapes selectable at run time (all give identical arithmetic):
  // bit 0 halves the strided passes' column tile, bit 1 doubles the z pass's
  // pencils per block.  Only grids up to 512 carry the alternatives.
  static constexpr bool TUNABLE = N <= 512;
  static constexpr int T_YH = T_Y >= 4 ? T_Y / 2 : T_Y;
  static constexpr int T_YFH = T_YF >= 4 ? T_YF / 2 : T_YF;
  static constexpr int T_XH = T_X >= 4 ? T_X / 2 : T_X;
  // two waves of pencils; where a wave's pencils do not fill it exactly
  // (64 % TPP != 0) a second set would straddle the wave boundary, which the
  // lane exchanges of the Hermitian loads do not allow: one set only
  static constexpr int PPW2 = (TPP > 64 || 64 % TPP == 0) && (N * N) % (2 * PPW) == 0 ? 2 * PPW : PPW;
  int variants() const override { return TUNABLE ? 4 : 1; }

  int step(dns_ctx *c, long nsteps) override {
    int rc = prepare(c);
    if (rc) return rc;
    if (c->use_graph && !c->timing && !c->capture && !c->rec_phys) return step_graph(c, nsteps);
    return step_launch(c, nsteps);
  }

  int step_launch(dns_ctx *c, long nsteps) {
    if constexpr (TUNABLE) {
      switch (c->variant) {
        case 1: return step_body<T_YH, T_YFH, T_XH, PPW>(c, nsteps);
        case 2: return step_body<T_Y, T_YF, T_X, PPW2>(c, nsteps);
        case 3: return step_body<T_YH, T_YFH, T_XH, PPW2>(c, nsteps);
        default: break;
      }
    }
    return step_body<T_Y, T_YF, T_X, PPW>(c, nsteps);
  }

  // Steps replayed from captured graphs (per launch shape): one of
  // GRAPH_STEPS steps for the bulk of a call, one of a single step for the
  // remainder, so a call of many steps costs one graph launch per
  // GRAPH_STEPS steps.
  static constexpr long GRAPH_STEPS = 10;

  int capture(dns_ctx *c, long k, hipGraphExec_t *out) {
    hipGraph_t graph = nullptr;
    DNS_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rc = step_launch(c, k);
    hipError_t e = hipStreamEndCapture(c->stream, &graph);
    if (rc || e != hipSuccess) {
      if (graph) (void)hipGraphDestroy(graph);
      if (rc) return rc;
      DNS_HIP(e);
    }
    e = hipGraphInstantiate(out, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    DNS_HIP(e);
    return DNS_OK;
  }

  int step_graph(dns_ctx *c, long nsteps) {
    if (c->gvariant != c->variant) {
      if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
      if (c->gexec1) (void)hipGraphExecDestroy(c->gexec1);
      c->gexec = c->gexec1 = nullptr;
      c->gvariant = c->variant;
    }
    const long bulk = nsteps / GRAPH_STEPS, rest = nsteps % GRAPH_STEPS;
    int rc;
    if (bulk && !c->gexec && (rc = capture(c, GRAPH_STEPS, &c->gexec))) return rc;
    if (rest && !c->gexec1 && (rc = capture(c, 1, &c->gexec1))) return rc;
    for (long i = 0; i < bulk; i++) DNS_HIP(hipGraphLaunch(c->gexec, c->stream));
    for (long i = 0; i < rest; i++) DNS_HIP(hipGraphLaunch(c->gexec1, c->stream));
    c->w_ready = 1;
    return DNS_OK;
  }

  // TY: inverse y-pass tile; TYF: forward y-pass tile; TX: x-pass tile
  // (E_X elements per thread); PP: z pencils per block.
  template <int TY, int TYF, int TX, int PP> int step_body(dns_ctx *c, long nsteps) {
    hipStream_t s = c->stream;
    const Geo g = c->g;
    const int M = c->members;
    cx<R> *W = (cx<R> *)c->W;
    cx<SR> *uhr = (cx<SR> *)c->uhr;
    double2 *uhn = (double2 *)c->uhn, *acc = (double2 *)c->acc;
    const cx<R> *tw = twR(c);
    const R invn3 = (R)c->invn3;
    static const double a_rk[4] = {1 / 6.0, 1 / 3.0, 1 / 3.0, 1 / 6.0};  // main.c:15
    static const double b_rk[4] = {0.5, 0.5, 1.0, 0.0};                 // main.c:16
    // y passes: paired tiles where a tile row is narrower than a line
    constexpr bool PY = x_tiles_paired<TY, R>(), PYF = x_tiles_paired<TYF, R>();
    const dim3 gy = PY ? dim3(paired_blocks((NF + TY - 1) / TY, N), 1, M) : grid_x(TY, M), by(TY * TPP_Y);
    const dim3 gx = grid_xp<TX>(M), bx(TX * TPP_X);
    const dim3 gz(N * N / PP, 1, M), bz(PP * TPP);
    // the forward y pass covers the dealiased kz columns only
    const dim3 gy_trunc = PYF ? dim3(paired_blocks((g.nk + TYF - 1) / TYF, N), 1, M)
                              : dim3((g.nk + TYF - 1) / TYF, N, M),
               byf(TYF * TPP_Y);
    static const char *const stage_names[4] = {"rk stage 1", "rk stage 2", "rk stage 3", "rk stage 4"};
    for (long st = 0; st < nsteps; st++) {
      const RoctxRange step_range("dns step");
      double2 *ph = (c->capture && st == nsteps - 1) ? (double2 *)c->phat : nullptr;
      for (int rk = 0; rk < 4; rk++) {
        const RoctxRange stage_range(stage_names[rk]);
        RKParams q{c->nudt, c->kmax, a_rk[rk], b_rk[rk]};
        const int sl = DNS_NPASS * rk;
        // pass A of this stage ran fused into the previous stage's E
        // the adjoint's recording of this stage's u and omega (FP64, M = 1)
        double *rp = (sizeof(R) == 8 && c->rec_phys) ? c->rec_phys + (size_t)rk * 6 * g.rs : nullptr;
        if (rk == 0) {
          DNS_TIMED(c, sl + 1, (k_ypass<N, R, TY, true, 1, PIPE_Y, false, E_Y, false, true, PY><<<gy, by, 0, s>>>(W + 3 * g.wfs, 3, g, tw, g.mW)));
          if (rp)
            DNS_TIMED(c, sl + 2, (k_zfused<N, R, PP, 0, R, MINW_Z, REC_OK><<<gz, bz, 0, s>>>(W, (R *)c->us, g, invn3, tw, rp)));
          else
            DNS_TIMED(c, sl + 2, (k_zfused<N, R, PP, 0, R, MINW_Z><<<gz, bz, 0, s>>>(W, (R *)c->us, g, invn3, tw)));
        } else {
          if constexpr (CX_IN_Y)
            DNS_TIMED(c, sl + 1, (k_ypass<N, R, TY, true, 1, PIPE_YC, false, E_Y, true, true, PY><<<gy, by, 0, s>>>(W, 6, g, tw, g.mW)));
          else
            DNS_TIMED(c, sl + 1, (k_ypass<N, R, TY, true, 1, PIPE_Y, false, E_Y, false, true, PY><<<gy, by, 0, s>>>(W, 6, g, tw, g.mW)));
          if (rp && rk < 3)
            DNS_TIMED(c, sl + 2, (k_zfused<N, R, PP, 1, R, MINW_Z, REC_OK><<<gz, bz, 0, s>>>(W, (R *)c->us, g, invn3, tw, rp)));
          else if (rp)
            DNS_TIMED(c, sl + 2, (k_zfused<N, R, PP, 2, R, MINW_Z, REC_OK><<<gz, bz, 0, s>>>(W, (R *)c->us, g, invn3, tw, rp)));
          else if (rk < 3)
            DNS_TIMED(c, sl + 2, (k_zfused<N, R, PP, 1, R, MINW_Z><<<gz, bz, 0, s>>>(W, (R *)c->us, g, invn3, tw)));
          else
            DNS_TIMED(c, sl + 2, (k_zfused<N, R, PP, 2, R, MINW_Z><<<gz, bz, 0, s>>>(W, (R *)c->us, g, invn3, tw)));
        }
        DNS_TIMED(c, sl + 3,
                  (k_ypass<N, R, TYF, false, 1, PIPE_YF, true, E_Y, false, true, PYF><<<gy_trunc, byf, 0, s>>>(W, 3, g, tw, g.mW)));
        // E fused with the next stage's A (the next step's stage 1 after rk 3)
        switch (rk) {
          case 0:
            DNS_TIMED(c, sl + 5, (k_xfwd_rk_xinv<N, R, TX, 0, SR, 1, EPI_X, E_X, !CX_IN_Y, S64><<<gx, bx, 0, s>>>(W, uhr, uhn, acc, nullptr, g, q, tw)));
            break;
          case 1:
            DNS_TIMED(c, sl + 5, (k_xfwd_rk_xinv<N, R, TX, 1, SR, 1, EPI_X, E_X, !CX_IN_Y, S64><<<gx, bx, 0, s>>>(W, uhr, uhn, acc, nullptr, g, q, tw)));
            break;
          case 2:
            DNS_TIMED(c, sl + 5, (k_xfwd_rk_xinv<N, R, TX, 2, SR, 1, EPI_X, E_X, !CX_IN_Y, S64><<<gx, bx, 0, s>>>(W, uhr, uhn, acc, nullptr, g, q, tw)));
            break;
          default:
            DNS_TIMED(c, sl + 5, (k_xfwd_rk_xinv<N, R, TX, 3, SR, 1, EPI_X3, E_X, !CX_IN_Y, S64><<<gx, bx, 0, s>>>(W, uhr, uhn, acc, ph, g, q, tw)));
            break;
        }
      }
      if (ph) c->have_phat = 1;
    }
    c->w_ready = 1;
    return DNS_OK;
  }

  // c2r (double) of one spectral double field into a physical double field
  // times `scale`, one member.
  // c2r of one spectral state field (state layout) into `phys`, x scale.
  int inv3_double(dns_ctx *c, const double2 *spec, double2 *tmpc, double *phys, double scale) {
    hipStream_t s = c->stream;
    k_xplain<N, double, T_D, true, double, true><<<grid_x(T_D), T_D * TPP, 0, s>>>(spec, tmpc, 1, c->g, twD(c));
    DNS_LAUNCHED();
    k_ypass<N, double, T_D, true><<<grid_x(T_D), T_D * TPP, 0, s>>>(tmpc, 1, c->g, twD(c));
    DNS_LAUNCHED();
    k_zc2r<N, double, PPW><<<N * N / PPW, PPW * TPP, 0, s>>>(tmpc, phys, c->g, scale, twD(c));
    DNS_LAUNCHED();
    return DNS_OK;
  }

  int snapshot(dns_ctx *c, int m, double *uvwp) override {
    c->w_ready = 0;
    hipStream_t s = c->stream;
    double2 *tmpc = (double2 *)wslice(c, m);
    double *tmpr = (double *)((char *)tmpc + sizeof(double2) * c->g.fs);
    const size_t rb = sizeof(double) * c->g.rs;
    for (int f = 0; f < 4; f++) {
      if (f == 3 && !c->have_phat) {
        memset(uvwp + 3 * c->g.rs, 0, rb);
        continue;
      }
      const double2 *src = f < 3 ? (const double2 *)c->uhn + m * c->g.mS + f * c->g.fs
                                 : (const double2 *)c->phat + m * c->g.fs;
      int rc = inv3_double(c, src, tmpc, tmpr, c->invn3);
      if (rc) return rc;
      DNS_HIP(hipMemcpyAsync(uvwp + f * c->g.rs, tmpr, rb, hipMemcpyDeviceToHost, s));
      DNS_HIP(hipStreamSynchronize(s));
    }
    return DNS_OK;
  }

  int snapshot_device(dns_ctx *c, double *dst) override {
    c->w_ready = 0;
    hipStream_t s = c->stream;
    double2 *tmpc = (double2 *)wslice(c, 0);
    const size_t rb = sizeof(double) * c->g.rs;
    for (int f = 0; f < 4; f++) {
      double *out = dst + f * c->g.rs;
      if (f == 3 && !c->have_phat) {
        DNS_HIP(hipMemsetAsync(out, 0, rb, s));
        continue;
      }
      const double2 *src = f < 3 ? (const double2 *)c->uhn + f * c->g.fs : (const double2 *)c->phat;
      int rc = inv3_double(c, src, tmpc, out, c->invn3);
      if (rc) return rc;
    }
    return DNS_OK;
  }

  int rfft3(dns_ctx *c, const double *in, double *out) override {
    c->w_ready = 0;
    hipStream_t s = c->stream;
    const Geo g = c->g;
    double *phys = (double *)c->W;
    cx<R> *spec = (cx<R> *)((char *)c->W + sizeof(double) * g.rs);
    DNS_HIP(hipMemcpyAsync(phys, in, sizeof(double) * g.rs, hipMemcpyHostToDevice, s));
    k_zr2c<N, R, PPW><<<N * N / PPW, PPW * TPP, 0, s>>>(phys, spec, 1, g, twR(c));
    DNS_LAUNCHED();
    k_ypass<N, R, T_P, false><<<grid_x(T_P), T_P * TPP, 0, s>>>(spec, 1, g, twR(c));
    DNS_LAUNCHED();
    k_xplain<N, R, T_P, false, R><<<grid_x(T_P), T_P * TPP, 0, s>>>(spec, spec, 1, g, twR(c));
    DNS_LAUNCHED();
    double2 *specd = (double2 *)spec;
    if (sizeof(R) != 8) {  // widen into the region after the float spectrum
      specd = (double2 *)((char *)spec + sizeof(cx<R>) * g.fs);
      k_convert<R, double><<<4096, 256, 0, s>>>((const R *)spec, (double *)specd, 2 * g.fs);
      DNS_LAUNCHED();
    }
    DNS_HIP(hipMemcpy2DAsync(out, sizeof(double2) * NF, specd, sizeof(double2) * g.nfp, sizeof(double2) * NF,
                             (size_t)N * N, hipMemcpyDeviceToHost, s));
    DNS_HIP(hipStreamSynchronize(s));
    return DNS_OK;
  }

  int irfft3(dns_ctx *c, const double *in, double *out) override {
    c->w_ready = 0;
    hipStream_t s = c->stream;
    const Geo g = c->g;
    // W: [spectral double input | complex R partial | physical double output]
    double2 *specd = (double2 *)c->W;
    cx<R> *tmpc = (cx<R> *)((char *)c->W + sizeof(double2) * g.fs);
    double *phys = (double *)((char *)tmpc + sizeof(cx<R>) * g.fs);
    DNS_HIP(hipMemset2DAsync(specd, sizeof(double2) * g.nfp, 0, sizeof(double2) * g.nfp, (size_t)N * N, s));
    DNS_HIP(hipMemcpy2DAsync(specd, sizeof(double2) * g.nfp, in, sizeof(double2) * NF, sizeof(double2) * NF,
                             (size_t)N * N, hipMemcpyHostToDevice, s));
    k_xplain<N, R, T_P, true, double><<<grid_x(T_P), T_P * TPP, 0, s>>>(specd, tmpc, 1, g, twR(c));
    DNS_LAUNCHED();
    k_ypass<N, R, T_P, true><<<grid_x(T_P), T_P * TPP, 0, s>>>(tmpc, 1, g, twR(c));
    DNS_LAUNCHED();
    k_zc2r<N, R, PPW><<<N * N / PPW, PPW * TPP, 0, s>>>(tmpc, phys, g, (R)1, twR(c));
    DNS_LAUNCHED();
    DNS_HIP(hipMemcpyAsync(out, phys, sizeof(double) * g.rs, hipMemcpyDeviceToHost, s));
    DNS_HIP(hipStreamSynchronize(s));
    return DNS_OK;
  }

  int get_us(dns_ctx *c, int m, double *out) override {
    hipStream_t s = c->stream;
    const size_t cnt = 3 * (size_t)c->g.rs;
    const R *us = (const R *)c->us + m * c->g.mR;
    const double *src = (const double *)us;
    if (sizeof(R) != 8) {
      c->w_ready = 0;
      double *tmp = (double *)wslice(c, m);
      k_convert<R, double><<<4096, 256, 0, s>>>(us, tmp, (long)cnt);
      DNS_LAUNCHED();
      src = tmp;
    }
    DNS_HIP(hipMemcpyAsync(out, src, cnt * sizeof(double), hipMemcpyDeviceToHost, s));
    DNS_HIP(hipStreamSynchronize(s));
    return DNS_OK;
  }

  int adj_forward(dns_ctx *c, const double *u0, long nsteps, long every) override;
  int adj_backward(dns_ctx *c, double seed, double *grad) override;
  int adj_step(dns_ctx *c, const double *uh, const double *us, const double *lam_uh, const double *lam_us,
               double *out_uh, double *out_us) override;

  int set_us(dns_ctx *c, int m, const double *in) override {
    c->w_ready = 0;
    hipStream_t s = c->stream;
    const size_t cnt = 3 * (size_t)c->g.rs;
    R *us = (R *)c->us + m * c->g.mR;
    if (sizeof(R) == 8) {
      DNS_HIP(hipMemcpyAsync(us, in, cnt * sizeof(double), hipMemcpyHostToDevice, s));
    } else {
      double *tmp = (double *)wslice(c, m);
      DNS_HIP(hipMemcpyAsync(tmp, in, cnt * sizeof(double), hipMemcpyHostToDevice, s));
      k_convert<double, R><<<4096, 256, 0, s>>>(tmp, us, (long)cnt);
      DNS_LAUNCHED();
    }
    DNS_HIP(hipStreamSynchronize(s));
    return DNS_OK;
  }
};

template <int N, typename R, typename SR, bool S64>
int Engine<N, R, SR, S64>::adj_forward(dns_ctx *c, const double *u0, long nsteps, long every) {
  if constexpr (sizeof(R) != 8) {
    return DNS_EINVAL;
  } else {
    return Adjoint<N>::forward(c, u0, nsteps, every, nullptr);
  }
}

template <int N, typename R, typename SR, bool S64>
int Engine<N, R, SR, S64>::adj_backward(dns_ctx *c, double seed, double *grad) {
  if constexpr (sizeof(R) != 8) {
    return DNS_EINVAL;
  } else {
    return Adjoint<N>::backward(c, seed, grad);
  }
}

template <int N, typename R, typename SR, bool S64>
int Engine<N, R, SR, S64>::adj_step(dns_ctx *c, const double *uh, const double *us, const double *lam_uh,
                                    const double *lam_us, double *out_uh, double *out_us) {
  if constexpr (sizeof(R) != 8) {
    return DNS_EINVAL;
  } else {
    AdjWork *w;
    int rc = Adjoint<N>::workspace(c, w);
    if (rc) return rc;
    if (w->recs.empty()) {
      void *p = nullptr;
      if ((rc = w->get(&p, w->REC))) return rc;
      w->recs.push_back(p);
    }
    hipStream_t s = c->stream;
    const size_t pitch = sizeof(double2) * c->nfp, width = sizeof(double2) * NF, rows = (size_t)3 * N * N;
    // host spectra are x-major; tmpN takes them into the state layout
    DNS_HIP(hipMemset2DAsync(w->tmpN, pitch, 0, pitch, rows, s));
    DNS_HIP(hipMemcpy2DAsync(w->tmpN, pitch, uh, width, width, rows, hipMemcpyDefault, s));
    k_state_rows<double2, true><<<N * N, 256, 0, s>>>(w->tmpN, (double2 *)c->uhn, N, c->g, 3);
    DNS_LAUNCHED();
    DNS_HIP(hipMemcpyAsync(c->us, us, w->R3, hipMemcpyDefault, s));
    DNS_HIP(hipMemset2DAsync(w->tmpN, pitch, 0, pitch, rows, s));
    DNS_HIP(hipMemcpy2DAsync(w->tmpN, pitch, lam_uh, width, width, rows, hipMemcpyDefault, s));
    k_state_rows<double2, true><<<N * N, 256, 0, s>>>(w->tmpN, w->lamU, N, c->g, 3);
    DNS_LAUNCHED();
    DNS_HIP(hipMemcpyAsync(w->lamS, lam_us, w->R3, hipMemcpyDefault, s));
    c->w_ready = 0;
    if ((rc = Adjoint<N>::record(c, w->recs[0]))) return rc;
    if ((rc = Adjoint<N>::backward_stages(c, w, (const double *)w->recs[0]))) return rc;
    k_state_rows<double2, false><<<N * N, 256, 0, s>>>(w->lamU, w->tmpN, N, c->g, 3);
    DNS_LAUNCHED();
    DNS_HIP(hipMemcpy2DAsync(out_uh, width, w->tmpN, pitch, width, rows, hipMemcpyDefault, s));
    DNS_HIP(hipMemcpyAsync(out_us, w->lamS, w->R3, hipMemcpyDefault, s));
    DNS_HIP(hipStreamSynchronize(s));
    c->w_ready = 0;
    return DNS_OK;
  }
}

template <int N> Impl *make_engine(int prec) {
  if (prec == DNS_MIXED) return new Engine<N, float, float>();
  if (prec == DNS_MIXED_ACC64) return new Engine<N, float, float, true>();
  return new Engine<N, double, double>();
}

}  // namespace dnsk
