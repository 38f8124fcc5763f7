// C ABI of libdns_hip (declared in include/dns.h).
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <algorithm>
#include <cstdlib>
#include <dlfcn.h>
#include <new>
#include "engine.hpp"
#include "grid_sizes.h"

namespace dnsk {

// roctx entry points, opened once on first use when DNS_ROCTX is set (and
// not "0"); a missing library leaves the ranges off.
const RoctxApi &roctx_api() {
  static const RoctxApi api = [] {
    RoctxApi a;
    const char *e = getenv("DNS_ROCTX");
    if (!e || !*e || strcmp(e, "0") == 0) return a;
    const char *libs[] = {"librocprofiler-sdk-roctx.so.1", "libroctx64.so.4"};
    for (const char *l : libs) {
      void *h = dlopen(l, RTLD_NOW | RTLD_LOCAL);
      if (!h) continue;
      a.push = (int (*)(const char *))dlsym(h, "roctxRangePushA");
      a.pop = (int (*)())dlsym(h, "roctxRangePop");
      if (a.push && a.pop) break;
      a.push = nullptr;
      a.pop = nullptr;
    }
    return a;
  }();
  return api;
}

static thread_local char g_err[512];

void set_error(const char *what, hipError_t e) {
  snprintf(g_err, sizeof g_err, "%s: %s", what, hipGetErrorString(e));
}

void set_error_text(const char *what) { snprintf(g_err, sizeof g_err, "%s", what); }

// One translation unit per grid size (inst.hip compiled with -DDNS_N=<n>)
// instantiates the engine.
#define DNS_DECLARE(n) Impl *make_engine_##n(int);
DNS_GRID_SIZES(DNS_DECLARE)
#undef DNS_DECLARE

static Impl *make_impl(int n, int prec) {
  switch (n) {
#define DNS_CASE(n) \
  case n: return make_engine_##n(prec);
    DNS_GRID_SIZES(DNS_CASE)
#undef DNS_CASE
    default: return nullptr;
  }
}

}  // namespace dnsk

using namespace dnsk;

static const int DIAG_BLOCKS = 1024;

static int alloc(dns_ctx *c, void **p, size_t bytes) {
  DNS_HIP(hipMalloc(p, bytes));
  DNS_HIP(hipMemsetAsync(*p, 0, bytes, c->stream));
  c->bytes += (long long)bytes;
  return DNS_OK;
}

// A context of `members` simulations of grid n (members >= 1).
static int ctx_create(dns_ctx **out, int n, int members, const double *nu, const double *dt, int precision,
                      int device) {
  *out = nullptr;
  g_err[0] = 0;
  if (precision != DNS_FP64 && precision != DNS_MIXED && precision != DNS_MIXED_ACC64) return DNS_EINVAL;
  if (members < 1) return DNS_EINVAL;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    snprintf(g_err, sizeof g_err, "no HIP device found");
    return DNS_ENODEV;
  }
  if (device < 0 || device >= ndev) return DNS_ENODEV;
  Impl *impl = make_impl(n, precision);
  if (!impl) {
    snprintf(g_err, sizeof g_err, "unsupported grid size n=%d", n);
    return DNS_EINVAL;
  }
  dns_ctx *c = new (std::nothrow) dns_ctx();
  if (!c) {
    delete impl;
    return DNS_ENOMEM;
  }
  c->impl = impl;
  c->n = n;
  c->nf = n / 2 + 1;
  c->nfp = (c->nf + 15) / 16 * 16;
  c->prec = precision;
  c->device = device;
  c->members = members;
  c->nu.assign(nu, nu + members);
  c->dt.assign(dt, dt + members);
  c->kmax = 2. / 3. * (n / 2 + 1);        // main.c:231
  c->invn3 = 1.0 / ((long)n * n * n);     // main.c:188
  c->g.nfp = c->nfp;
  c->g.fs = (long)n * n * c->nfp;
  c->g.rs = (long)n * n * n;
  c->g.nk = 0;  // wavenumbers m >= 0 inside the dealias box, |m| < kmax (main.c:237-238)
  while (c->g.nk <= n / 2 && (double)c->g.nk < c->kmax) c->g.nk++;
  c->g.wps = (long)(n + 1) * c->nfp;  // step partials: one spare row per x-plane
  c->g.wfs = (long)n * c->g.wps;
  c->g.mW = 6 * c->g.wfs;
  c->g.mS = 3 * c->g.fs;
  c->g.mR = 3 * c->g.rs;
  if (hipSetDevice(device) != hipSuccess) {
    delete c->impl;
    delete c;
    return DNS_ENODEV;
  }
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    delete c->impl;
    delete c;
    return DNS_EHIP;
  }
#define DNS_TRY(x)     \
  do {                 \
    int rc_ = (x);     \
    if (rc_) {         \
      dns_destroy(c);  \
      return rc_;      \
    }                  \
  } while (0)
  const size_t M = (size_t)members, C64 = sizeof(double2) * c->g.fs;
  DNS_TRY(alloc(c, &c->uhn, M * 3 * C64));
  DNS_TRY(alloc(c, &c->uhr, M * 3 * impl->sr_size() * c->g.fs));
  DNS_TRY(alloc(c, &c->acc, M * impl->acc_bytes(c->g.fs)));
  DNS_TRY(alloc(c, &c->W, M * 6 * 2 * impl->r_size() * c->g.wfs));
  DNS_TRY(alloc(c, &c->us, M * 3 * impl->r_size() * c->g.rs));
  DNS_TRY(alloc(c, (void **)&c->part, sizeof(double) * 4 * DIAG_BLOCKS));
  DNS_TRY(alloc(c, (void **)&c->dres, sizeof(double) * 4));
  DNS_TRY(alloc(c, (void **)&c->nudt, sizeof(double2) * M));
  std::vector<double2> nd(M);
  for (size_t m = 0; m < M; m++) nd[m] = make_double2(nu[m], dt[m]);
  if (hipMemcpyAsync(c->nudt, nd.data(), sizeof(double2) * M, hipMemcpyHostToDevice, c->stream) != hipSuccess) {
    dns_destroy(c);
    return DNS_EHIP;
  }
  DNS_TRY(impl->make_twiddles(c));
#undef DNS_TRY
  if (hipStreamSynchronize(c->stream) != hipSuccess) {
    dns_destroy(c);
    return DNS_EHIP;
  }
  *out = c;
  return DNS_OK;
}

static int ctx_diagnostics(dns_ctx *c, int m, double *energy, double *enstrophy) {
  DNS_HIP(hipSetDevice(c->device));
  const int nb = (int)((long)c->n * c->n < DIAG_BLOCKS ? (long)c->n * c->n : DIAG_BLOCKS);
  const double2 *uh = (const double2 *)c->uhn + m * c->g.mS;
  k_diag<256><<<nb, 256, 0, c->stream>>>(uh, c->n, c->g, c->part);
  DNS_LAUNCHED();
  k_diag_final<256><<<1, 256, 0, c->stream>>>(c->part, nb, c->dres);
  DNS_LAUNCHED();
  double r[4];
  DNS_HIP(hipMemcpyAsync(r, c->dres, sizeof r, hipMemcpyDeviceToHost, c->stream));
  DNS_HIP(hipStreamSynchronize(c->stream));
  // main.c:265-266: scale by (1/n^3)^2 after the sum
  const double s = c->invn3 * c->invn3;
  *energy = (r[0] + r[1]) * s;
  *enstrophy = (r[2] + r[3]) * s;
  if (!std::isfinite(*energy) || !std::isfinite(*enstrophy)) return DNS_ENONFINITE;
  return DNS_OK;
}

static int ctx_get_spectral(dns_ctx *c, int m, double *uvw_hat) {
  DNS_HIP(hipSetDevice(c->device));
  return c->impl->get_spectral(c, m, uvw_hat);
}

static int ctx_set_spectral(dns_ctx *c, int m, const double *uvw_hat) {
  DNS_HIP(hipSetDevice(c->device));
  return c->impl->set_spectral(c, m, uvw_hat);
}

// Time every launch shape of the context for `steps` steps from its
// current state, keep the fastest, and restore the state.  All shapes give
// bitwise identical results.
static int ctx_autotune(dns_ctx *c, int steps, int *best_out, double *best_ms) {
  DNS_HIP(hipSetDevice(c->device));
  const int nv = c->impl->variants();
  int best = c->variant;
  double bms = 0;
  if (nv > 1) {
    const size_t sb = (size_t)c->members * c->g.mS * sizeof(double2);
    const size_t ub = (size_t)c->members * c->g.mR * c->impl->r_size();
    void *su = nullptr, *sus = nullptr;
    hipEvent_t a = nullptr, b = nullptr;
    int rc = DNS_OK;
    // keeps the first failure; the state is restored whatever fails later
    auto hip = [&](hipError_t e, const char *what) {
      if (e != hipSuccess && rc == DNS_OK) {
        set_error(what, e);
        rc = e == hipErrorOutOfMemory ? DNS_ENOMEM : DNS_EHIP;
      }
    };
    hip(hipMalloc(&su, sb), "hipMalloc");
    if (rc == DNS_OK) hip(hipMalloc(&sus, ub), "hipMalloc");
    if (rc == DNS_OK) hip(hipEventCreate(&a), "hipEventCreate");
    if (rc == DNS_OK) hip(hipEventCreate(&b), "hipEventCreate");
    const bool saved = rc == DNS_OK;
    if (saved) {
      hip(hipMemcpyAsync(su, c->uhn, sb, hipMemcpyDeviceToDevice, c->stream), "hipMemcpyAsync");
      hip(hipMemcpyAsync(sus, c->us, ub, hipMemcpyDeviceToDevice, c->stream), "hipMemcpyAsync");
    }
    for (int v = 0; v < nv && rc == DNS_OK; v++) {
      c->impl->set_variant(c, v);
      rc = c->impl->step(c, 1);
      if (rc == DNS_OK) hip(hipEventRecord(a, c->stream), "hipEventRecord");
      if (rc == DNS_OK) rc = c->impl->step(c, steps);
      if (rc == DNS_OK) hip(hipEventRecord(b, c->stream), "hipEventRecord");
      if (rc == DNS_OK) hip(hipEventSynchronize(b), "hipEventSynchronize");
      float ms = 0;
      if (rc == DNS_OK) hip(hipEventElapsedTime(&ms, a, b), "hipEventElapsedTime");
      if (rc == DNS_OK && (v == 0 || ms < bms)) {
        bms = ms;
        best = v;
      }
    }
    if (saved) {
      hip(hipMemcpyAsync(c->uhn, su, sb, hipMemcpyDeviceToDevice, c->stream), "hipMemcpyAsync");
      hip(hipMemcpyAsync(c->us, sus, ub, hipMemcpyDeviceToDevice, c->stream), "hipMemcpyAsync");
      const int sc = c->impl->state_changed(c, -1);
      if (rc == DNS_OK) rc = sc;
      hip(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    }
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
    if (su) (void)hipFree(su);
    if (sus) (void)hipFree(sus);
    c->w_ready = 0;
    c->have_phat = 0;
    if (rc) return rc;
    c->impl->set_variant(c, best);
  }
  if (best_out) *best_out = best;
  if (best_ms) *best_ms = bms;
  return DNS_OK;
}

// A batch: one context per distinct grid size, members mapped to
// (context, index).  Steps are replayed from one captured graph that runs the
// contexts on concurrent branches.
struct dns_batch {
  int members, device;
  std::vector<dns_ctx *> groups;
  std::vector<int> group_of, index_of;
  hipStream_t main;
  hipEvent_t fork;
  std::vector<hipEvent_t> join;
  // graphs of BATCH_GRAPH_STEPS steps and of one step of all groups
  hipGraphExec_t exec, exec1;
};

// Steps per replayed batch graph (one graph launch per this many steps).
static const long BATCH_GRAPH_STEPS = 10;

extern "C" {

const char *dns_strerror(int code) {
  switch (code) {
    case DNS_OK: return "success";
    case DNS_EINVAL: return "invalid argument";
    case DNS_ENODEV: return "no usable GPU";
    case DNS_ENOMEM: return "out of device memory";
    case DNS_EHIP: return "HIP runtime error";
    case DNS_ENONFINITE: return "non-finite energy or enstrophy";
    default: return "unknown error";
  }
}

const char *dns_last_error(void) { return g_err; }

int dns_create(dns_ctx **out, int n, double nu, double dt, int precision, int device) {
  if (!out) return DNS_EINVAL;
  int rc = ctx_create(out, n, 1, &nu, &dt, precision, device);
  if (rc == DNS_OK) (*out)->use_graph = 1;
  return rc;
}

void dns_destroy(dns_ctx *c) {
  if (!c) return;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  void *ptrs[] = {c->uhn, c->uhr, c->acc, c->W, c->us, c->phat, c->tw64, c->tw32, c->part, c->dres, c->nudt};
  for (void *p : ptrs)
    if (p) (void)hipFree(p);
  for (auto &t : c->ev_pending) {
    c->ev_pool.push_back(t.a);
    c->ev_pool.push_back(t.b);
  }
  for (hipEvent_t e : c->ev_pool)
    if (e) (void)hipEventDestroy(e);
  delete (AdjWork *)c->adj;
  if (c->snap_stream) (void)hipStreamSynchronize(c->snap_stream);
  if (c->snap_dev) (void)hipFree(c->snap_dev);
  if (c->snap_host) (void)hipHostFree(c->snap_host);
  if (c->snap_ready) (void)hipEventDestroy(c->snap_ready);
  if (c->snap_done) (void)hipEventDestroy(c->snap_done);
  if (c->snap_stream) (void)hipStreamDestroy(c->snap_stream);
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  if (c->gexec1) (void)hipGraphExecDestroy(c->gexec1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c->impl;
  delete c;
}

int dns_get_n(const dns_ctx *c) { return c ? c->n : DNS_EINVAL; }

long long dns_device_bytes(const dns_ctx *c) { return c ? c->bytes : 0; }

int dns_load_velocity(dns_ctx *c, const double *uvw) {
  if (!c || !uvw) return DNS_EINVAL;
  DNS_HIP(hipSetDevice(c->device));
  c->have_phat = 0;
  return c->impl->load(c, 0, uvw);
}

int dns_step(dns_ctx *c, long nsteps) {
  if (!c || nsteps < 0) return DNS_EINVAL;
  if (nsteps == 0) return DNS_OK;
  DNS_HIP(hipSetDevice(c->device));
  return c->impl->step(c, nsteps);
}

int dns_set_pressure_capture(dns_ctx *c, int on) {
  if (!c) return DNS_EINVAL;
  DNS_HIP(hipSetDevice(c->device));
  if (on && !c->phat) {
    int rc = alloc(c, &c->phat, sizeof(double2) * c->g.fs);
    if (rc) return rc;
  }
  c->capture = on ? 1 : 0;
  return DNS_OK;
}

int dns_diagnostics(dns_ctx *c, double *energy, double *enstrophy) {
  if (!c || !energy || !enstrophy) return DNS_EINVAL;
  return ctx_diagnostics(c, 0, energy, enstrophy);
}

int dns_snapshot(dns_ctx *c, double *uvwp) {
  if (!c || !uvwp) return DNS_EINVAL;
  DNS_HIP(hipSetDevice(c->device));
  return c->impl->snapshot(c, 0, uvwp);
}

int dns_snapshot_begin(dns_ctx *c) {
  if (!c) return DNS_EINVAL;
  DNS_HIP(hipSetDevice(c->device));
  const size_t bytes = 4 * sizeof(double) * (size_t)c->g.rs;
  if (!c->snap_dev) {
    DNS_HIP(hipStreamCreateWithFlags(&c->snap_stream, hipStreamNonBlocking));
    DNS_HIP(hipEventCreateWithFlags(&c->snap_ready, hipEventDisableTiming));
    DNS_HIP(hipEventCreateWithFlags(&c->snap_done, hipEventDisableTiming));
    DNS_HIP(hipHostMalloc((void **)&c->snap_host, bytes, hipHostMallocDefault));
    void *d = nullptr;
    DNS_HIP(hipMalloc(&d, bytes));
    c->snap_dev = (double *)d;
    c->bytes += (long long)bytes;
  }
  // the staging buffer is reused: the previous copy out of it must be done
  if (c->snap_pending) DNS_HIP(hipStreamWaitEvent(c->stream, c->snap_done, 0));
  int rc = c->impl->snapshot_device(c, c->snap_dev);
  if (rc) return rc;
  DNS_HIP(hipEventRecord(c->snap_ready, c->stream));
  DNS_HIP(hipStreamWaitEvent(c->snap_stream, c->snap_ready, 0));
  DNS_HIP(hipMemcpyAsync(c->snap_host, c->snap_dev, bytes, hipMemcpyDeviceToHost, c->snap_stream));
  DNS_HIP(hipEventRecord(c->snap_done, c->snap_stream));
  c->snap_pending = 1;
  return DNS_OK;
}

int dns_snapshot_wait(dns_ctx *c, const double **uvwp) {
  if (!c || !uvwp || !c->snap_pending) return DNS_EINVAL;
  DNS_HIP(hipSetDevice(c->device));
  DNS_HIP(hipEventSynchronize(c->snap_done));
  *uvwp = c->snap_host;
  return DNS_OK;
}

int dns_get_spectral(dns_ctx *c, double *uvw_hat) {
  if (!c || !uvw_hat) return DNS_EINVAL;
  return ctx_get_spectral(c, 0, uvw_hat);
}

int dns_set_spectral(dns_ctx *c, const double *uvw_hat) {
  if (!c || !uvw_hat) return DNS_EINVAL;
  return ctx_set_spectral(c, 0, uvw_hat);
}

int dns_get_stale_velocity(dns_ctx *c, double *uvw) {
  if (!c || !uvw) return DNS_EINVAL;
  DNS_HIP(hipSetDevice(c->device));
  return c->impl->get_us(c, 0, uvw);
}

int dns_set_stale_velocity(dns_ctx *c, const double *uvw) {
  if (!c || !uvw) return DNS_EINVAL;
  DNS_HIP(hipSetDevice(c->device));
  return c->impl->set_us(c, 0, uvw);
}

int dns_set_kernel_timing(dns_ctx *c, int on) {
  if (!c) return DNS_EINVAL;
  DNS_HIP(hipSetDevice(c->device));
  DNS_HIP(hipStreamSynchronize(c->stream));
  for (auto &t : c->ev_pending) {
    c->ev_pool.push_back(t.a);
    c->ev_pool.push_back(t.b);
  }
  c->ev_pending.clear();
  for (int i = 0; i < DNS_NSLOT; i++) {
    c->slot_ms[i] = 0;
    c->slot_n[i] = 0;
  }
  c->timing = on ? 1 : 0;
  return DNS_OK;
}

int dns_kernel_timing(dns_ctx *c, double *ms, long *launches) {
  if (!c || !ms || !launches) return DNS_EINVAL;
  DNS_HIP(hipSetDevice(c->device));
  DNS_HIP(hipStreamSynchronize(c->stream));
  for (auto &t : c->ev_pending) {
    float e = 0;
    if (t.a && t.b && hipEventElapsedTime(&e, t.a, t.b) == hipSuccess) {
      c->slot_ms[t.slot] += e;
      c->slot_n[t.slot] += 1;
    }
    c->ev_pool.push_back(t.a);
    c->ev_pool.push_back(t.b);
  }
  c->ev_pending.clear();
  for (int i = 0; i < DNS_NSLOT; i++) {
    ms[i] = c->slot_ms[i];
    launches[i] = c->slot_n[i];
  }
  return DNS_OK;
}

const char *dns_kernel_slot_name(int slot) {
  static const char *pass[DNS_NPASS] = {"xinv_curl", "ypass_inv", "zfused", "ypass_fwd", "xfwd_rk", "xfwd_rk_xinv"};
  static char names[DNS_NSLOT][32];
  if (slot < 0 || slot >= DNS_NSLOT) return "";
  if (!names[slot][0])
    snprintf(names[slot], sizeof names[slot], "%s.s%d", pass[slot % DNS_NPASS], slot / DNS_NPASS + 1);
  return names[slot];
}

int dns_adjoint_forward(dns_ctx *c, const double *u0, long nsteps, long every, double *enstrophy) {
  if (!c || !u0 || !enstrophy || nsteps < 1 || every < 1) return DNS_EINVAL;
  if (c->prec != DNS_FP64 || c->members != 1) return DNS_EINVAL;
  DNS_HIP(hipSetDevice(c->device));
  c->have_phat = 0;
  int rc = c->impl->adj_forward(c, u0, nsteps, every);
  if (rc) return rc;
  double e;
  return ctx_diagnostics(c, 0, &e, enstrophy);
}

int dns_adjoint_backward(dns_ctx *c, double seed, double *grad) {
  if (!c || !grad) return DNS_EINVAL;
  if (c->prec != DNS_FP64 || c->members != 1) return DNS_EINVAL;
  DNS_HIP(hipSetDevice(c->device));
  return c->impl->adj_backward(c, seed, grad);
}

int dns_adjoint_step(dns_ctx *c, const double *uvw_hat, const double *uvw_stale, const double *lam_hat,
                     const double *lam_stale, double *lam_hat_out, double *lam_stale_out) {
  if (!c || !uvw_hat || !uvw_stale || !lam_hat || !lam_stale || !lam_hat_out || !lam_stale_out)
    return DNS_EINVAL;
  if (c->prec != DNS_FP64 || c->members != 1) return DNS_EINVAL;
  DNS_HIP(hipSetDevice(c->device));
  return c->impl->adj_step(c, uvw_hat, uvw_stale, lam_hat, lam_stale, lam_hat_out, lam_stale_out);
}

int dns_launch_shapes(dns_ctx *c) { return c ? c->impl->variants() : DNS_EINVAL; }

int dns_set_launch_shape(dns_ctx *c, int shape) {
  if (!c) return DNS_EINVAL;
  return c->impl->set_variant(c, shape);
}

int dns_get_launch_shape(dns_ctx *c) { return c ? c->variant : DNS_EINVAL; }

int dns_autotune(dns_ctx *c, int steps, int *best_shape) {
  if (!c || steps < 1) return DNS_EINVAL;
  return ctx_autotune(c, steps, best_shape, nullptr);
}

int dns_rfft3(dns_ctx *c, const double *real, double *spec) {
  if (!c || !real || !spec) return DNS_EINVAL;
  DNS_HIP(hipSetDevice(c->device));
  return c->impl->rfft3(c, real, spec);
}

int dns_irfft3(dns_ctx *c, const double *spec, double *real) {
  if (!c || !spec || !real) return DNS_EINVAL;
  DNS_HIP(hipSetDevice(c->device));
  return c->impl->irfft3(c, spec, real);
}


// ---------------------------------------------------------------------------
// Batched ensembles.

void dns_batch_destroy(dns_batch *b) {
  if (!b) return;
  (void)hipSetDevice(b->device);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->exec1) (void)hipGraphExecDestroy(b->exec1);
  if (b->main) (void)hipStreamSynchronize(b->main);
  for (dns_ctx *c : b->groups) dns_destroy(c);
  for (hipEvent_t e : b->join) (void)hipEventDestroy(e);
  if (b->fork) (void)hipEventDestroy(b->fork);
  if (b->main) (void)hipStreamDestroy(b->main);
  delete b;
}

int dns_batch_create(dns_batch **out, int members, const int *n, const double *nu, const double *dt,
                     int precision, int device) {
  if (!out || members < 1 || !n || !nu || !dt) return DNS_EINVAL;
  *out = nullptr;
  dns_batch *b = new (std::nothrow) dns_batch();
  if (!b) return DNS_ENOMEM;
  b->members = members;
  b->device = device;
  b->group_of.assign(members, -1);
  b->index_of.assign(members, -1);
  std::vector<int> sizes;
  for (int m = 0; m < members; m++)
    if (std::find(sizes.begin(), sizes.end(), n[m]) == sizes.end()) sizes.push_back(n[m]);
  for (size_t gi = 0; gi < sizes.size(); gi++) {
    std::vector<double> gnu, gdt;
    for (int m = 0; m < members; m++)
      if (n[m] == sizes[gi]) {
        b->group_of[m] = (int)gi;
        b->index_of[m] = (int)gnu.size();
        gnu.push_back(nu[m]);
        gdt.push_back(dt[m]);
      }
    dns_ctx *c = nullptr;
    int rc = ctx_create(&c, sizes[gi], (int)gnu.size(), gnu.data(), gdt.data(), precision, device);
    if (rc) {
      dns_batch_destroy(b);
      return rc;
    }
    b->groups.push_back(c);
  }
  if (hipStreamCreateWithFlags(&b->main, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess) {
    dns_batch_destroy(b);
    return DNS_EHIP;
  }
  b->join.resize(b->groups.size());
  for (auto &e : b->join)
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
      dns_batch_destroy(b);
      return DNS_EHIP;
    }
  *out = b;
  return DNS_OK;
}

static int batch_member(dns_batch *b, int m, dns_ctx **c, int *i) {
  if (!b || m < 0 || m >= b->members) return DNS_EINVAL;
  *c = b->groups[b->group_of[m]];
  *i = b->index_of[m];
  return DNS_OK;
}

int dns_batch_load_velocity(dns_batch *b, int member, const double *uvw) {
  dns_ctx *c;
  int i, rc = batch_member(b, member, &c, &i);
  if (rc || !uvw) return rc ? rc : DNS_EINVAL;
  DNS_HIP(hipSetDevice(c->device));
  return c->impl->load(c, i, uvw);
}

}  // extern "C"

// Capture k steps of every group into one graph, one concurrent branch per
// group (fork/join events on the batch's own stream).
static int batch_capture(dns_batch *b, long k, hipGraphExec_t *out) {
  hipGraph_t graph = nullptr;
  DNS_HIP(hipStreamBeginCapture(b->main, hipStreamCaptureModeGlobal));
  int rc = DNS_OK;
  hipError_t e = hipEventRecord(b->fork, b->main);
  for (size_t gi = 0; gi < b->groups.size() && rc == DNS_OK && e == hipSuccess; gi++) {
    dns_ctx *c = b->groups[gi];
    e = hipStreamWaitEvent(c->stream, b->fork, 0);
    if (e == hipSuccess) rc = c->impl->step(c, k);
    if (rc == DNS_OK && e == hipSuccess) e = hipEventRecord(b->join[gi], c->stream);
    if (rc == DNS_OK && e == hipSuccess) e = hipStreamWaitEvent(b->main, b->join[gi], 0);
  }
  hipError_t e2 = hipStreamEndCapture(b->main, &graph);
  if (rc || e != hipSuccess || e2 != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    if (rc) return rc;
    DNS_HIP(e != hipSuccess ? e : e2);
  }
  e = hipGraphInstantiate(out, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  DNS_HIP(e);
  return DNS_OK;
}

extern "C" {

int dns_batch_step(dns_batch *b, long nsteps) {
  if (!b || nsteps < 0) return DNS_EINVAL;
  if (nsteps == 0) return DNS_OK;
  DNS_HIP(hipSetDevice(b->device));
  // stage-1 partials of every group must be present before replaying
  for (dns_ctx *c : b->groups) {
    if (c->timing) return DNS_EINVAL;
    int rc = c->impl->prepare(c);
    if (rc) return rc;
    DNS_HIP(hipStreamSynchronize(c->stream));
  }
  const long bulk = nsteps / BATCH_GRAPH_STEPS, rest = nsteps % BATCH_GRAPH_STEPS;
  int rc;
  if (bulk && !b->exec && (rc = batch_capture(b, BATCH_GRAPH_STEPS, &b->exec))) return rc;
  if (rest && !b->exec1 && (rc = batch_capture(b, 1, &b->exec1))) return rc;
  for (long s = 0; s < bulk; s++) DNS_HIP(hipGraphLaunch(b->exec, b->main));
  for (long s = 0; s < rest; s++) DNS_HIP(hipGraphLaunch(b->exec1, b->main));
  // the groups' own streams must see the graph's work before later calls
  DNS_HIP(hipEventRecord(b->fork, b->main));
  for (dns_ctx *c : b->groups) DNS_HIP(hipStreamWaitEvent(c->stream, b->fork, 0));
  return DNS_OK;
}

int dns_batch_autotune(dns_batch *b, int steps) {
  if (!b || steps < 1) return DNS_EINVAL;
  for (dns_ctx *c : b->groups) {
    int rc = ctx_autotune(c, steps, nullptr, nullptr);
    if (rc) return rc;
  }
  // recapture with the chosen shapes
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->exec1) (void)hipGraphExecDestroy(b->exec1);
  b->exec = b->exec1 = nullptr;
  return DNS_OK;
}

int dns_batch_launch_shape(dns_batch *b, int member) {
  dns_ctx *c;
  int i, rc = batch_member(b, member, &c, &i);
  return rc ? rc : c->variant;
}

int dns_batch_diagnostics(dns_batch *b, int member, double *energy, double *enstrophy) {
  dns_ctx *c;
  int i, rc = batch_member(b, member, &c, &i);
  if (rc || !energy || !enstrophy) return rc ? rc : DNS_EINVAL;
  return ctx_diagnostics(c, i, energy, enstrophy);
}

int dns_batch_get_spectral(dns_batch *b, int member, double *uvw_hat) {
  dns_ctx *c;
  int i, rc = batch_member(b, member, &c, &i);
  if (rc || !uvw_hat) return rc ? rc : DNS_EINVAL;
  return ctx_get_spectral(c, i, uvw_hat);
}

int dns_batch_get_stale_velocity(dns_batch *b, int member, double *uvw) {
  dns_ctx *c;
  int i, rc = batch_member(b, member, &c, &i);
  if (rc || !uvw) return rc ? rc : DNS_EINVAL;
  DNS_HIP(hipSetDevice(c->device));
  return c->impl->get_us(c, i, uvw);
}

int dns_batch_member_n(dns_batch *b, int member) {
  dns_ctx *c;
  int i, rc = batch_member(b, member, &c, &i);
  return rc ? rc : c->n;
}

long long dns_batch_device_bytes(const dns_batch *b) {
  long long s = 0;
  if (b)
    for (dns_ctx *c : b->groups) s += c->bytes;
  return s;
}

}  // extern "C"

