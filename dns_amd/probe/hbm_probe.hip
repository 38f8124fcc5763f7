// hbm_probe: streaming-bandwidth ceilings of the GPU and a copy kernel of
// known size, used to read the passes' measured HBM bandwidth against what
// a plain stream achieves and to calibrate rocprofv3 FETCH_SIZE/WRITE_SIZE.
//
//   hbm_probe            table of read-only / write-only / copy kernels
//   hbm_probe --calib    one copy of 1 GiB (reads 1 GiB, writes 1 GiB)
//   hbm_probe --planes   pencil-tile copies in the layout of the step's
//                        partials at n=512 (x pencils with and without a
//                        spare row per x-plane, y pencils)
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>
#include <cstring>

#define CK(x)                                                                  \
  do {                                                                         \
    hipError_t e_ = (x);                                                       \
    if (e_ != hipSuccess) {                                                    \
      fprintf(stderr, "hbm_probe: %s: %s\n", #x, hipGetErrorString(e_));     \
      exit(1);                                                                 \
    }                                                                          \
  } while (0)

typedef double v2d __attribute__((ext_vector_type(2)));

__global__ void k_copy(const v2d *__restrict__ a, v2d *__restrict__ b, long n) {
  const long stride = (long)gridDim.x * blockDim.x;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += stride) b[i] = a[i];
}
__global__ void k_copy_nt(const v2d *__restrict__ a, v2d *__restrict__ b, long n) {
  const long stride = (long)gridDim.x * blockDim.x;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += stride)
    __builtin_nontemporal_store(__builtin_nontemporal_load(a + i), b + i);
}
__global__ void k_read(const v2d *__restrict__ a, double *out, long n) {
  const long stride = (long)gridDim.x * blockDim.x;
  double s = 0;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += stride) {
    v2d v = a[i];
    s += v.x + v.y;
  }
  if (s == 1234.5678) out[0] = s;  // keeps the loads; never true for zeros
}
__global__ void k_fill(v2d *__restrict__ b, long n) {
  const long stride = (long)gridDim.x * blockDim.x;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += stride) b[i] = v2d{1.0, 2.0};
}

// Copy of 8-column tiles (k) along the pencils of one axis of a complex
// [512][rows][272] layout, 8 elements per thread as in the passes: XPEN
// walks the slowest index (plane stride ps), otherwise the middle one (row
// stride 272).  Grid (34, 512), 512 threads.
template <bool XPEN>
__global__ __launch_bounds__(512) void k_tile_copy(const v2d *__restrict__ a, v2d *__restrict__ b, long ps) {
  constexpr int N = 512, NFP = 272, T = 8, E = 8, TPP = N / E;
  const int p = threadIdx.x % T, t = threadIdx.x / T;
  const long k = (long)blockIdx.x * T + p, fixed = blockIdx.y;
  v2d v[E];
#pragma unroll
  for (int c = 0; c < E; c++) {
    const long m = t + TPP * c;
    v[c] = a[XPEN ? m * ps + fixed * NFP + k : fixed * ps + m * NFP + k];
  }
#pragma unroll
  for (int c = 0; c < E; c++) {
    const long m = t + TPP * c;
    __builtin_nontemporal_store(v[c], b + (XPEN ? m * ps + fixed * NFP + k : fixed * ps + m * NFP + k));
  }
}

static hipStream_t st;

template <class F> static void timeit(const char *name, double bytes, F f) {
  hipEvent_t a, b;
  CK(hipEventCreate(&a));
  CK(hipEventCreate(&b));
  f();
  CK(hipGetLastError());
  CK(hipStreamSynchronize(st));
  const int reps = 10;
  CK(hipEventRecord(a, st));
  for (int r = 0; r < reps; r++) f();
  CK(hipEventRecord(b, st));
  CK(hipEventSynchronize(b));
  CK(hipGetLastError());
  float ms;
  CK(hipEventElapsedTime(&ms, a, b));
  ms /= reps;
  printf("| %-28s | %8.3f | %8.1f | %5.1f |\n", name, ms, bytes / ms / 1e6, bytes / ms / 1e6 / 80.0);
  fflush(stdout);
}

int main(int argc, char **argv) {
  const bool calib = argc > 1 && strcmp(argv[1], "--calib") == 0;
  const bool planes = argc > 1 && strcmp(argv[1], "--planes") == 0;
  CK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  const long n = calib ? (1L << 26) : (1L << 28);  // 1 GiB or 4 GiB of 16-byte elements
  v2d *a, *b;
  double *o;
  CK(hipMalloc(&a, n * 16));
  CK(hipMalloc(&b, n * 16));
  CK(hipMalloc(&o, 64));
  CK(hipMemset(a, 0, n * 16));
  CK(hipMemset(b, 0, n * 16));
  if (calib) {
    k_copy<<<32768, 256, 0, st>>>(a, b, n);
    CK(hipGetLastError());
    CK(hipStreamSynchronize(st));
    printf("calibration copy: read %ld bytes, wrote %ld bytes\n", n * 16, n * 16);
    return 0;
  }
  if (planes) {  // largest index (511*513+511)*272+271 < n
    printf("| tile copy, [512][rows][272] complex double | ms | GB/s | %% of 8 TB/s |\n|---|---|---|---|\n");
    const double tb = 2.0 * 16 * 512 * 512 * 272;
    const dim3 grid(272 / 8, 512);
    timeit("x pencils, 512 rows a plane", tb, [&] { k_tile_copy<true><<<grid, 512, 0, st>>>(a, b, 512L * 272); });
    timeit("x pencils, 513 rows a plane", tb, [&] { k_tile_copy<true><<<grid, 512, 0, st>>>(a, b, 513L * 272); });
    timeit("y pencils", tb, [&] { k_tile_copy<false><<<grid, 512, 0, st>>>(a, b, 512L * 272); });
    return 0;
  }
  printf("| kernel (4 GiB arrays) | ms | GB/s | %% of 8 TB/s |\n|---|---|---|---|\n");
  const double rb = 16.0 * n, cb = 32.0 * n;
  for (int grid : {2048, 32768}) {
    char nm[64];
    snprintf(nm, sizeof nm, "read-only, grid %d", grid);
    timeit(nm, rb, [&] { k_read<<<grid, 256, 0, st>>>(a, o, n); });
    snprintf(nm, sizeof nm, "write-only, grid %d", grid);
    timeit(nm, rb, [&] { k_fill<<<grid, 256, 0, st>>>(b, n); });
    snprintf(nm, sizeof nm, "copy, grid %d", grid);
    timeit(nm, cb, [&] { k_copy<<<grid, 256, 0, st>>>(a, b, n); });
    snprintf(nm, sizeof nm, "copy non-temporal, grid %d", grid);
    timeit(nm, cb, [&] { k_copy_nt<<<grid, 256, 0, st>>>(a, b, n); });
  }
  return 0;
}
