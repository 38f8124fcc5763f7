// dns_diagnostics_noalloc: the energy / enstrophy reduction of
// dns_diagnostics (main.c:257-266) on caller-owned device data, for callers
// that cannot allocate device memory (code under HIP graph capture, tight
// serving loops).  It allocates nothing: the block partial sums live in a
// file-scope device array of fixed size, `scratch`.
//
// Each GPU holds its own copy of `scratch` from the library's load to the
// end of the process.  In host code the name `scratch` is only a stand-in
// for the device symbol, so the host never passes it to a kernel: the
// address of each device's copy is looked up once per device, with that
// device current (hipGetSymbolAddress), kept in `g_scratch`, and only that
// address is used.  Every call on a GPU shares its copy, which keeps the
// last call's partial sums; calls that overlap on one GPU (different
// streams, or other host threads) race on it and must be serialised by the
// caller.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <mutex>
#include "engine.hpp"

namespace dnsk {

// Blocks of the first reduction pass (the n*n rows are spread over at most
// this many blocks, as dns_diagnostics does); four doubles per block.
static constexpr int NOALLOC_BLOCKS = 1024;
static constexpr int NOALLOC_MAX_N = 1024;

static __device__ __align__(256) double scratch[4 * NOALLOC_BLOCKS];

static constexpr int MAX_DEVICES = 64;
static double *g_scratch[MAX_DEVICES];
static std::mutex g_scratch_mutex;

// Device address of this device's `scratch` (looked up on the first call on
// each device).
static int scratch_of(int dev, double **out) {
  if (dev < 0 || dev >= MAX_DEVICES) return DNS_ENODEV;
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch[dev]) {
    void *p = nullptr;
    DNS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return DNS_EHIP;
    g_scratch[dev] = (double *)p;
  }
  *out = g_scratch[dev];
  return DNS_OK;
}

}  // namespace dnsk

using namespace dnsk;

extern "C" {

int dns_diagnostics_noalloc(const double *uvw_hat, int n, void *stream, double *out) {
  if (!uvw_hat || !out) return DNS_EINVAL;
  if (n < 2 || n % 2 || n > NOALLOC_MAX_N) {
    char msg[160];
    snprintf(msg, sizeof msg, "dns_diagnostics_noalloc: n=%d is not an even size from 2 to %d", n,
             NOALLOC_MAX_N);
    set_error_text(msg);
    return DNS_EINVAL;
  }
  int dev = -1;
  DNS_HIP(hipGetDevice(&dev));
  double *part = nullptr;
  int rc = scratch_of(dev, &part);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  // the host API's x-major layout [3][n][n][n/2+1], unpadded rows
  Geo g{};
  g.nfp = n / 2 + 1;
  g.fs = (long)n * n * g.nfp;
  const int nb = (int)((long)n * n < NOALLOC_BLOCKS ? (long)n * n : NOALLOC_BLOCKS);
  const double invn3 = 1.0 / ((long)n * n * n);  // main.c:188
  k_diag<256, true><<<nb, 256, 0, s>>>((const double2 *)uvw_hat, n, g, part);
  DNS_LAUNCHED();
  k_diag_final<256, true><<<1, 256, 0, s>>>(part, nb, out, invn3 * invn3);
  DNS_LAUNCHED();
  return DNS_OK;
}

}  // extern "C"
