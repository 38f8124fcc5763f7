// C ABI of libdns_hip (declared in include/dns.h).
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include "engine.hpp"

namespace dnsk {

static thread_local char g_err[512];

void set_error(const char *what, hipError_t e) {
  snprintf(g_err, sizeof g_err, "%s: %s", what, hipGetErrorString(e));
}

// One translation unit per grid size (inst_<n>.hip) instantiates the engine.
Impl *make_engine_16(int);
Impl *make_engine_32(int);
Impl *make_engine_48(int);
Impl *make_engine_64(int);
Impl *make_engine_96(int);
Impl *make_engine_128(int);
Impl *make_engine_192(int);
Impl *make_engine_256(int);
Impl *make_engine_384(int);
Impl *make_engine_512(int);
Impl *make_engine_768(int);
Impl *make_engine_1024(int);

static Impl *make_impl(int n, int prec) {
  switch (n) {
    case 16: return make_engine_16(prec);
    case 32: return make_engine_32(prec);
    case 48: return make_engine_48(prec);
    case 64: return make_engine_64(prec);
    case 96: return make_engine_96(prec);
    case 128: return make_engine_128(prec);
    case 192: return make_engine_192(prec);
    case 256: return make_engine_256(prec);
    case 384: return make_engine_384(prec);
    case 512: return make_engine_512(prec);
    case 768: return make_engine_768(prec);
    case 1024: return make_engine_1024(prec);
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

#define DNS_TRY(x)      \
  do {                  \
    int rc_ = (x);      \
    if (rc_) {          \
      dns_destroy(c);   \
      return rc_;       \
    }                   \
  } while (0)

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
  *out = nullptr;
  g_err[0] = 0;
  if (precision != DNS_FP64 && precision != DNS_MIXED) return DNS_EINVAL;
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
  c->nu = nu;
  c->dt = dt;
  c->kmax = 2. / 3. * (n / 2 + 1);        // main.c:231
  c->invn3 = 1.0 / ((long)n * n * n);     // main.c:188
  c->g.nfp = c->nfp;
  c->g.fs = (long)n * n * c->nfp;
  c->g.rs = (long)n * n * n;
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
  const size_t C64 = sizeof(double2) * c->g.fs;
  DNS_TRY(alloc(c, &c->uhn, 3 * C64));
  DNS_TRY(alloc(c, &c->uhr, 3 * impl->sr_size() * c->g.fs));
  DNS_TRY(alloc(c, &c->acc, 3 * C64));
  DNS_TRY(alloc(c, &c->W, 6 * 2 * impl->r_size() * c->g.fs));
  DNS_TRY(alloc(c, &c->us, 3 * impl->r_size() * c->g.rs));
  DNS_TRY(alloc(c, (void **)&c->part, sizeof(double) * 4 * DIAG_BLOCKS));
  DNS_TRY(alloc(c, (void **)&c->dres, sizeof(double) * 4));
  DNS_TRY(impl->make_twiddles(c));
  if (hipStreamSynchronize(c->stream) != hipSuccess) {
    dns_destroy(c);
    return DNS_EHIP;
  }
  *out = c;
  return DNS_OK;
}

void dns_destroy(dns_ctx *c) {
  if (!c) return;
  hipSetDevice(c->device);
  if (c->stream) hipStreamSynchronize(c->stream);
  void *ptrs[] = {c->uhn, c->uhr, c->acc, c->W, c->us, c->phat, c->tw64, c->tw32, c->part, c->dres};
  for (void *p : ptrs)
    if (p) hipFree(p);
  for (auto &t : c->ev_pending) {
    c->ev_pool.push_back(t.a);
    c->ev_pool.push_back(t.b);
  }
  for (hipEvent_t e : c->ev_pool)
    if (e) hipEventDestroy(e);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c->impl;
  delete c;
}

int dns_get_n(const dns_ctx *c) { return c ? c->n : DNS_EINVAL; }

long long dns_device_bytes(const dns_ctx *c) { return c ? c->bytes : 0; }

int dns_load_velocity(dns_ctx *c, const double *uvw) {
  if (!c || !uvw) return DNS_EINVAL;
  DNS_HIP(hipSetDevice(c->device));
  c->have_phat = 0;
  return c->impl->load(c, uvw);
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
  DNS_HIP(hipSetDevice(c->device));
  const int nb = (int)((long)c->n * c->n < DIAG_BLOCKS ? (long)c->n * c->n : DIAG_BLOCKS);
  k_diag<256><<<nb, 256, 0, c->stream>>>((const double2 *)c->uhn, c->n, c->g, c->part);
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

int dns_snapshot(dns_ctx *c, double *uvwp) {
  if (!c || !uvwp) return DNS_EINVAL;
  DNS_HIP(hipSetDevice(c->device));
  return c->impl->snapshot(c, uvwp);
}

int dns_get_spectral(dns_ctx *c, double *uvw_hat) {
  if (!c || !uvw_hat) return DNS_EINVAL;
  DNS_HIP(hipSetDevice(c->device));
  DNS_HIP(hipMemcpy2DAsync(uvw_hat, sizeof(double2) * c->nf, c->uhn, sizeof(double2) * c->nfp,
                           sizeof(double2) * c->nf, (size_t)3 * c->n * c->n, hipMemcpyDeviceToHost, c->stream));
  DNS_HIP(hipStreamSynchronize(c->stream));
  return DNS_OK;
}

int dns_set_spectral(dns_ctx *c, const double *uvw_hat) {
  if (!c || !uvw_hat) return DNS_EINVAL;
  DNS_HIP(hipSetDevice(c->device));
  DNS_HIP(hipMemcpy2DAsync(c->uhn, sizeof(double2) * c->nfp, uvw_hat, sizeof(double2) * c->nf,
                           sizeof(double2) * c->nf, (size_t)3 * c->n * c->n, hipMemcpyHostToDevice, c->stream));
  DNS_HIP(hipStreamSynchronize(c->stream));
  return DNS_OK;
}

int dns_get_stale_velocity(dns_ctx *c, double *uvw) {
  if (!c || !uvw) return DNS_EINVAL;
  DNS_HIP(hipSetDevice(c->device));
  return c->impl->get_us(c, uvw);
}

int dns_set_stale_velocity(dns_ctx *c, const double *uvw) {
  if (!c || !uvw) return DNS_EINVAL;
  DNS_HIP(hipSetDevice(c->device));
  return c->impl->set_us(c, uvw);
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
  static const char *pass[5] = {"xinv_curl", "ypass_inv", "zfused", "ypass_fwd", "xfwd_rk"};
  static char names[DNS_NSLOT][32];
  if (slot < 0 || slot >= DNS_NSLOT) return "";
  if (!names[slot][0]) snprintf(names[slot], sizeof names[slot], "%s.s%d", pass[slot % 5], slot / 5 + 1);
  return names[slot];
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

}  // extern "C"
