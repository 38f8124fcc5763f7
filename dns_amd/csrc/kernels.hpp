// HIP kernels of the RK4 step (gfx950).  All arrays are device-resident.
//
// Layouts (i = x, j = y, k = z fastest):
//   spectral partials (W): complex [n][n+1][nfp], element (i,j,k) at
//                     i*wps + j*nfp + k, wps = (n+1)*nfp (x slowest, as
//                     main.c:236,358; one spare row per x-plane),
//                     k < nf = n/2+1 valid, nfp = nf rounded up to 16
//                     (rows start on 128-byte boundaries);
//   spectral state (uhn, uhr, acc, phat): complex rows of nfp, row
//                     j*n+i (y slowest), element k at srow(row, k): only
//                     the x pass touches the state during a step, and in
//                     this layout its x-pencils have the short row stride of
//                     a y pass (float x pass 2.72 -> 2.39 ms at n=512), and
//                     rows i, i+1 are interleaved so that a tile of 4
//                     double columns covers whole 128-byte lines;
//   physical field  : real [n][n][n], (i*n+j)*n+k.
// A "pencil" is the set of elements along one axis with the other two
// indices fixed.  x and y pencils are strided: a block owns T pencils with
// consecutive k (a T-wide column tile), so every wave load/store touches
// whole T*sizeof(complex)-byte row segments.  z pencils are contiguous rows.
//
// Per RK stage the host launches (reference lines in main.c):
//   A  k_xinv_curl : curl prologue (354-362) + inverse FFT along x
//   B  k_ypass     : inverse FFT along y (in place)
//   C  k_zfused    : c2r along z of u and omega, 1/n^3 scaling (348-352,
//                    366-371), u x omega (372-377), r2c along z of N
//   D  k_ypass     : forward FFT along y
//   E  forward FFT along x + dealias*dt (381-386), projection and viscous
//      term (387-399), RK update (400-413): k_xfwd_rk_xinv runs it fused
//      with the next stage's A
#pragma once
#include "fft_engine.hpp"
#include <type_traits>

namespace dnsk {

struct Geo {
  int nfp;   // padded row length of spectral fields (complex elements)
  long fs;   // spectral field stride (complex elements) = n*n*nfp
  long rs;   // physical field stride (reals) = n^3
  int nk;    // wavenumbers m >= 0 with m < kmax: the dealiased box is
             // |kx|,|ky| < kmax <=> index in [0,nk) or [n-nk+1,n), kz in [0,nk)
  // Member strides (elements) when one context holds M simulations of the
  // same n: blockIdx.z selects the member.  W-type partials, spectral state
  // (uhn/uhr/acc) and stale physical velocity.
  long mW, mS, mR;
  // The step's partials W use their own strides: an i-plane of n rows plus
  // one spare row (wps = (n+1)*nfp), so that the x pass's pencils, which
  // step from plane to plane, do not keep hitting the same HBM channels
  // (a copy with the unpadded 2.2 MB plane stride of n=512 runs at 4.7 TB/s,
  // with one spare row 5.1 TB/s); wfs = n*wps per field.  Scratch uses of
  // the W buffer (transforms of the state, snapshots) keep the fs layout.
  long wfs, wps;
};

// z pencils per block of the contiguous-axis passes: as many as fill one
// wave (one pencil when a pencil needs a wave or more), reduced to a divisor
// of the n*n rows so that every block is full.
template <int N> constexpr int z_pencils_per_block() {
  constexpr int TPP = N / Sched<N>::E;
  int p = TPP >= 64 ? 1 : 64 / TPP;
  while ((N * N) % p) p--;
  return p;
}

// Column tile (bx) and row (by) of a strided-pass block.  Tiles whose rows
// are narrower than a 128-byte line (T complex elements) split each line
// between tiles 2q and 2q+1 of a row; such grids are launched one-
// dimensional (paired_blocks) and block p takes member p%16 >= 8 of pair
// slot 8*(p/16) + p%8 (slot = row * pairs per row + q), so the two tiles of a
// line run 8 blocks apart, which the hardware dispatches to the same XCD
// (consecutive workgroups go to the eight XCDs in turn): the second tile
// then finds the line in that XCD's L2 instead of fetching it again.  Pairs
// are formed within each row, also when a row has an odd number of tiles.
// ntx = column tiles per row, nty = rows.  Returns false for the blocks past
// the last tile (the launch is rounded up to whole groups of 16): they exit
// before their first barrier.
template <int T, typename R> constexpr bool x_tiles_paired() { return T * sizeof(cx<R>) < 128; }
__host__ __device__ constexpr long paired_blocks(long ntx, long nty) { return ((ntx + 1) / 2 * nty + 7) / 8 * 16; }
template <int T, typename R> __device__ __forceinline__ bool x_tile(int ntx, int nty, int &bx, int &by) {
  if constexpr (x_tiles_paired<T, R>()) {
    const int p = blockIdx.x, np = (ntx + 1) / 2;
    const int slot = 8 * (p / 16) + p % 8;
    by = slot / np;
    bx = 2 * (slot % np) + ((p % 16) >= 8 ? 1 : 0);
    return by < nty && bx < ntx;
  } else {
    bx = blockIdx.x;
    by = blockIdx.y;
    return true;
  }
}

// Index i of an x or y axis lies inside the dealiased box.
__device__ __forceinline__ bool in_box(int i, int n, int nk) { return i < nk || i > n - nk; }

struct RKParams {
  const double2 *nudt;  // per member (viscosity, time step)
  double kmax;
  double a, b;          // a[rk], b[rk] of main.c:15-16 for this stage
};

template <typename R> __device__ __forceinline__ cx<R> zero() { return mk<R>((R)0, (R)0); }

template <typename RO, typename RI> __device__ __forceinline__ cx<RO> cvt(cx<RI> a) {
  return mk<RO>((RO)a.x, (RO)a.y);
}

// Signed wavenumber of index i on an x or y axis: i for i < n/2, i - n
// otherwise (main.c:224-230, Nyquist = -n/2).
template <int N> __device__ __forceinline__ int wave(int i) { return i < N / 2 ? i : i - N; }

// Row r of a spectral state field holds (i, j) = (r % n, r / n).
__device__ __forceinline__ void state_row(long r, int n, int &i, int &j) {
  i = (int)(r % n);
  j = (int)(r / n);
}

// Offset of element k of row r in a spectral state field (uhn, uhr, acc,
// phat and the adjoint's state-shaped arrays).  Rows 2q and 2q+1 (x rows
// i, i+1 of one y plane j, n being even) are interleaved in groups of four
// complex elements: row r's group m sits at (r & ~1)*nfp + 8m + 4(r & 1).
// A 128-byte line of double elements thus holds columns 4m..4m+3 of two x
// rows (of float elements, 4m..4m+7), so an x-pass tile of 4 double
// columns (n=1024: a pencil takes 128 threads, a 512-thread block holds 4)
// reads and writes whole lines, where with one row per line each line was
// split between two neighbouring tiles' blocks: x pass 39.5 -> 30.9 ms at
// n=1024.  Wider tiles gain too: at n=512 (same-box runs taking turns) the
// FP64 x pass 3.91 -> 3.87 ms, the float one 2.36 -> 2.31 ms.  nfp is a
// multiple of 4.
__device__ __forceinline__ long srow(long r, int k, const Geo &g) {
  return (r & ~1L) * g.nfp + ((k >> 2) << 3) + ((r & 1) << 2) + (k & 3);
}

// Converts `nfields` padded spectral fields between the x-major layout of
// the host API (row i*n + j, element k at (i*n + j)*nfp + k) and the state
// layout (row j*n + i, srow): TO_STATE, dst (state) <- src (x-major);
// otherwise dst (x-major) <- src (state).  Grid (n*n) blocks, block r =
// state row j*n + i.
template <typename T, bool TO_STATE>
__global__ __launch_bounds__(256) void k_state_rows(const T *__restrict__ src, T *__restrict__ dst, int n, Geo g,
                                                    int nfields) {
  const long r = blockIdx.x, j = r / n, i = r % n;
  const long x = (i * n + j) * g.nfp;
  for (int f = 0; f < nfields; f++)
    for (int k = threadIdx.x; k < g.nfp; k += blockDim.x) {
      if (TO_STATE) dst[f * g.fs + srow(r, k, g)] = src[f * g.fs + x + k];
      else dst[f * g.fs + x + k] = src[f * g.fs + srow(r, k, g)];
    }
}

// ---------------------------------------------------------------------------
// B/D: FFT along y of `nfields` consecutive spectral fields, in place.
// Grid (ceil(nf/T), n): blockIdx.y = i.
// MINW: minimum resident waves per SIMD the register allocation must allow.
// PIPE: number of fields loaded ahead (0: none; 1: field f+1 is loaded
// while field f is transformed; 2: fields f+1 and f+2).
// TRUNC: the grid covers ceil(nk/T) column tiles and columns k >= g.nk are
// left alone.  Forward (the nonlinear term): only rows inside the dealiased
// box are written.  Inverse (the adjoint's lamN): only rows inside the box
// are read (the others are zero) and all rows are written.
// CURLX (inverse transform of u, v, w and omega, nfields = 6): field 3 is not
// read; curlX = I*(ky w - kz v) (main.c:361, same operation order) is formed
// from fields 1 and 2 as loaded, then transformed and stored as field 3.
// In (x, ky, kz) space ky and kz are per-element constants, so this equals
// the inverse x transform of the spectral curlX.
// PAIR: one-dimensional launch of paired tiles (x_tile; paired_blocks).
template <int N, typename R, int T, bool INV, int MINW = 1, int PIPE = 0, bool TRUNC = false,
          int E = Sched<N>::E, bool CURLX = false, bool WP = false, bool PAIR = false>
__global__ __launch_bounds__(T *(N / E), MINW) void k_ypass(cx<R> *__restrict__ data, int nfields, Geo g,
                                                           const cx<R> *__restrict__ tw, long mstride = 0) {
  constexpr int TPP = N / E;
  // WP: the step's partials (field stride wfs, plane stride wps); otherwise
  // the plain [n][n][nfp] layout of fields fs apart
  const long fstr = WP ? g.wfs : g.fs, pstr = WP ? g.wps : (long)N * g.nfp;
  // non-temporal loads (each element is read once), except in the float
  // forward pass, which they slow down (n=512: 1.47 -> 1.60 ms/step, while
  // the double one gains 2.84 -> 2.75 and the inverse passes 4-5%)
  constexpr bool NTL = !TRUNC || sizeof(R) == 8;
  __shared__ cx<R> lds[T * padlen<N>()];
  __shared__ cx<R> twl[tw_size<N>()];
  int bx = blockIdx.x, plane = blockIdx.y;
  if constexpr (PAIR) {
    if (!x_tile<T, R>(((TRUNC ? g.nk : N / 2 + 1) + T - 1) / T, N, bx, plane)) return;
  }
  tw = stage_twiddles<N, R>(twl, tw);
  const int p = threadIdx.x % T, t = threadIdx.x / T;
  const int k = bx * T + p;
  const bool ok = k < (TRUNC ? g.nk : N / 2 + 1);
  cx<R> *base = data + blockIdx.z * mstride + (long)plane * pstr + k;
  // loaded fields: all, or 0,1,2,4,5 with CURLX
  const int nload = CURLX ? 5 : nfields;
  auto fld = [](int i) { return CURLX && i >= 3 ? i + 1 : i; };
  auto load = [&](cx<R>(&v)[E], int f) {
    const cx<R> *d = base + f * fstr;
#pragma unroll
    for (int c = 0; c < E; c++) {
      const cx<R> *q = d + (long)(t + TPP * c) * g.nfp;
      const bool in = ok && (!(TRUNC && INV) || in_box(t + TPP * c, N, g.nk));
      v[c] = in ? (NTL ? ld_row<T>(q) : *q) : zero<R>();
    }
  };
  auto fft_store = [&](cx<R>(&v)[E], int f) {
    fft_pencil<N, R, INV>(v, t, lds, TileSlot<N, T>{p}, tw);
    if (ok) {
      cx<R> *d = base + f * fstr;
#pragma unroll
      for (int c = 0; c < E; c++)
        if (!TRUNC || INV || in_box(t + TPP * c, N, g.nk)) st<R>(d + (long)(t + TPP * c) * g.nfp, v[c]);
    }
  };
  cx<R> v[E], nx[E], nx2[E];
  cx<R> ox[E];  // CURLX: kz * v of field 1, then curlX
  load(v, 0);
  if (PIPE >= 2 && nload > 1) load(nx, fld(1));
  // not unrolled: an unrolled loop lets loads of later fields move up and
  // raises the register count
#pragma unroll 1
  for (int i = 0; i < nload; i++) {
    const int f = fld(i);
    if (PIPE == 1 && i + 1 < nload) load(nx, fld(i + 1));
    if (PIPE == 2 && i + 2 < nload) load(nx2, fld(i + 2));
    if constexpr (CURLX) {
      const R kz = (R)k;
      if (i == 1) {
#pragma unroll
        for (int c = 0; c < E; c++) ox[c] = mk<R>(kz * v[c].x, kz * v[c].y);
      } else if (i == 2) {
#pragma unroll
        for (int c = 0; c < E; c++) {  // I*(ky w - kz v), I*z = (-Im z, Re z)
          const R ky = (R)wave<N>(t + TPP * c);
          ox[c] = mk<R>(-(ky * v[c].y - ox[c].y), ky * v[c].x - ox[c].x);
        }
      }
    }
    fft_store(v, f);
    if (CURLX && i == 2) fft_store(ox, 3);
    if (i + 1 < nload) {
      if (PIPE) {
#pragma unroll
        for (int c = 0; c < E; c++) {
          v[c] = nx[c];
          if (PIPE == 2) nx[c] = nx2[c];
        }
      } else {
        load(v, fld(i + 1));
      }
    }
  }
}

// Plain FFT along x of `nfields` fields: in (precision RI) -> out (R).
// Used for the initial transform and for snapshots.  SIN / SOUT: the input
// / output is in the state layout; in may equal out only when both layouts
// are the same.  Grid (ceil(nf/T), n): blockIdx.y = j.
template <int N, typename R, int T, bool INV, typename RI, bool SIN = false, bool SOUT = false>
__global__ __launch_bounds__(T *(N / Sched<N>::E)) void k_xplain(const cx<RI> *in, cx<R> *out, int nfields, Geo g,
                                                                  const cx<R> *__restrict__ tw) {
  constexpr int E = Sched<N>::E, TPP = N / E;
  __shared__ cx<R> lds[T * padlen<N>()];
  __shared__ cx<R> twl[tw_size<N>()];
  tw = stage_twiddles<N, R>(twl, tw);
  const int p = threadIdx.x % T, t = threadIdx.x / T;
  const int k = blockIdx.x * T + p;
  const bool ok = k < N / 2 + 1;
  // element at x index i: x-major (i*n + j)*nfp + k, state srow(j*n + i, k)
  const long j = blockIdx.y;
  auto at = [&](bool state, int i) {
    return state ? srow(j * N + i, k, g) : ((long)i * N + j) * g.nfp + k;
  };
  for (int f = 0; f < nfields; f++) {
    cx<R> v[E];
#pragma unroll
    for (int c = 0; c < E; c++) v[c] = ok ? cvt<R, RI>(in[f * g.fs + at(SIN, t + TPP * c)]) : zero<R>();
    fft_pencil<N, R, INV>(v, t, lds, TileSlot<N, T>{p}, tw);
    if (ok) {
#pragma unroll
      for (int c = 0; c < E; c++) st<R>(out + f * g.fs + at(SOUT, t + TPP * c), v[c]);
    }
  }
}

// ---------------------------------------------------------------------------
// A: curl prologue + inverse FFT along x.  The curl follows main.c:359-361:
//   curlZ = I*(kx V - ky U), curlY = I*(kz U - kx W), curlX = I*(ky W - kz V)
// with I*z = (-Im z, Re z).
// WITH_U = false: writes the three omega partials to W fields 3,4,5.
// WITH_U = true : writes u, v, w to fields 0,1,2 and curlY, curlZ to 4,5.
//   curlX has no kx factor, so its inverse x transform is I*(ky w - kz v) of
//   the transformed v and w: the inverse y pass forms it from fields 1 and 2
//   (k_ypass CURLX) and this pass neither transforms nor stores it, unless
//   CX (shapes whose y pass has no registers for it) asks for field 3 too.
// Blocks own x-pencil tiles: grid (ceil(nf/T), n), blockIdx.y = j.  U, V, Wv
// hold the spectral velocity at the thread's positions i = t + TPP*c of
// pencil (j, k).
template <int N, typename R, int T, bool WITH_U, int E, bool CX = false>
__device__ __forceinline__ void curl_inverse_x(cx<R> (&U)[E], cx<R> (&V)[E], cx<R> (&Wv)[E],
                                               cx<R> *__restrict__ W, Geo g, int t, int p, int j, int k, bool ok,
                                               cx<R> *lds, const cx<R> *__restrict__ tw) {
  constexpr int TPP = N / E;
  const long off = (long)j * g.nfp + k, istr = g.wps;  // step partials
  const R ky = (R)wave<N>(j), kz = (R)k;
  const TileSlot<N, T> slot{p};
  auto store = [&](int f, cx<R>(&v)[E]) {
    if (ok) {
#pragma unroll
      for (int c = 0; c < E; c++) st<R>(W + f * g.wfs + off + (t + TPP * c) * istr, v[c]);
    }
  };
  auto irot = [](R a, R b) { return mk<R>(-b, a); };  // I*(a + i b)
  {
    cx<R> v[E];
    if constexpr (!WITH_U || CX) {
#pragma unroll
      for (int c = 0; c < E; c++) v[c] = irot(ky * Wv[c].x - kz * V[c].x, ky * Wv[c].y - kz * V[c].y);  // curlX
      fft_pencil<N, R, true>(v, t, lds, slot, tw);
      store(3, v);
    }
#pragma unroll
    for (int c = 0; c < E; c++) {  // curlY = I*(kz U - kx W)
      const R kx = (R)wave<N>(t + TPP * c);
      v[c] = irot(kz * U[c].x - kx * Wv[c].x, kz * U[c].y - kx * Wv[c].y);
    }
    fft_pencil<N, R, true>(v, t, lds, slot, tw);
    store(4, v);
#pragma unroll
    for (int c = 0; c < E; c++) {  // curlZ = I*(kx V - ky U)
      const R kx = (R)wave<N>(t + TPP * c);
      v[c] = irot(kx * V[c].x - ky * U[c].x, kx * V[c].y - ky * U[c].y);
    }
    fft_pencil<N, R, true>(v, t, lds, slot, tw);
    store(5, v);
    if constexpr (WITH_U) {
      fft_pencil<N, R, true>(U, t, lds, slot, tw);
      store(0, U);
      fft_pencil<N, R, true>(V, t, lds, slot, tw);
      store(1, V);
      fft_pencil<N, R, true>(Wv, t, lds, slot, tw);
      store(2, Wv);
    }
  }
}

template <int N, typename R, int T, bool WITH_U, typename SI, int MINW = 1, int E = Sched<N>::E>
__global__ __launch_bounds__(T *(N / E), MINW) void k_xinv_curl(const cx<SI> *__restrict__ uh,
                                                           cx<R> *__restrict__ W, Geo g,
                                                           const cx<R> *__restrict__ tw) {
  constexpr int TPP = N / E;
  __shared__ cx<R> lds[T * padlen<N>()];
  __shared__ cx<R> twl[tw_size<N>()];
  int bx, j;
  if (!x_tile<T, R>((N / 2 + T) / T, N, bx, j)) return;
  tw = stage_twiddles<N, R>(twl, tw);
  const int p = threadIdx.x % T, t = threadIdx.x / T;
  const int k = bx * T + p;
  const bool ok = k < N / 2 + 1;
  uh += blockIdx.z * g.mS;
  W += blockIdx.z * g.mW;
  cx<R> U[E], V[E], Wv[E];
#pragma unroll
  for (int c = 0; c < E; c++) {
    const long l = srow((long)j * N + t + TPP * c, k, g);
    U[c] = ok ? cvt<R, SI>(ld_state<T>(uh + l)) : zero<R>();
    V[c] = ok ? cvt<R, SI>(ld_state<T>(uh + g.fs + l)) : zero<R>();
    Wv[c] = ok ? cvt<R, SI>(ld_state<T>(uh + 2 * g.fs + l)) : zero<R>();
  }
  curl_inverse_x<N, R, T, WITH_U, E>(U, V, Wv, W, g, t, p, j, k, ok, lds, tw);
}

// ---------------------------------------------------------------------------
// z-pencil helpers.  A block owns PPW contiguous rows (i,j); thread t of a
// row holds positions pos = t + TPP*c.

// Complex pencil Z = X + iY of two half-spectrum rows extended to length N
// by Hermitian symmetry; the imaginary parts of k=0 and k=N/2 are dropped
// (FFTW c2r semantics).  Y == nullptr means Y = 0.  herm_load issues the
// loads (so they can be started early), herm_combine builds the pencil.
template <int N, typename R> struct HermRaw {
  cx<R> x[Sched<N>::E], y[Sched<N>::E];
};
// Value of v in lane `src` of this wave (ds_bpermute; every lane active).
template <typename R> __device__ __forceinline__ cx<R> lane_get(cx<R> v, int src) {
  if constexpr (sizeof(R) == 4) {
    return mk<R>(__int_as_float(__builtin_amdgcn_ds_bpermute(src << 2, __float_as_int(v.x))),
                 __int_as_float(__builtin_amdgcn_ds_bpermute(src << 2, __float_as_int(v.y))));
  } else {
    auto get = [&](double d) {
      const long long b = __double_as_longlong(d);
      const int lo = __builtin_amdgcn_ds_bpermute(src << 2, (int)(b & 0xffffffffLL));
      const int hi = __builtin_amdgcn_ds_bpermute(src << 2, (int)(b >> 32));
      return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
    };
    return mk<R>(get(v.x), get(v.y));
  }
}
template <int N, typename R>
__device__ __forceinline__ void herm_load(HermRaw<N, R> &h, const cx<R> *X, const cx<R> *Y, int t,
                                          int lim = N / 2 + 1) {
  constexpr int E = Sched<N>::E, TPP = N / E, H = N / 2;
  // Coefficients k >= lim are taken as zero and not read.
  if constexpr (TPP <= 64) {
    // Each stored element is loaded once (non-temporal): positions pos <= H
    // directly (pos < H sit at c < E/2: t + TPP c < TPP E/2 = H); pos > H
    // needs k = N - pos = TPP (E - c) - t, which thread TPP - t of the same
    // pencil (same wave) holds at index E - 1 - c (for t = 0: its own index
    // E - c, or the Nyquist element at c = E/2), passed between lanes.
#pragma unroll
    for (int c = 0; c < E / 2; c++) {
      const bool in = t + TPP * c < lim;
      h.x[c] = in ? ldnt(X + t + TPP * c) : zero<R>();
      h.y[c] = Y && in ? ldnt(Y + t + TPP * c) : zero<R>();
    }
    cx<R> nx = zero<R>(), ny = zero<R>();
    if (t == 0 && H < lim) {
      nx = ldnt(X + H);
      if (Y) ny = ldnt(Y + H);
    }
    const int lane = (int)(threadIdx.x & 63);
    const int src = lane - t + (TPP - t) % TPP;
#pragma unroll
    for (int c = E / 2; c < E; c++) {
      const cx<R> mx = lane_get<R>(h.x[E - 1 - c], src);
      const cx<R> my = lane_get<R>(h.y[E - 1 - c], src);
      if (t == 0) {
        h.x[c] = c == E / 2 ? nx : h.x[E - c];
        h.y[c] = c == E / 2 ? ny : h.y[E - c];
      } else {
        h.x[c] = mx;
        h.y[c] = my;
      }
    }
  } else {
    // a pencil over several waves: each thread loads its mirrored positions
    // too, with ordinary loads, so that the second fetch of an element finds
    // it in L2 (the pencil's other half-wave loaded it)
#pragma unroll
    for (int c = 0; c < E; c++) {
      const int pos = t + TPP * c;
      const int k = pos <= H ? pos : N - pos;
      h.x[c] = k < lim ? X[k] : zero<R>();
      h.y[c] = Y && k < lim ? Y[k] : zero<R>();
    }
  }
}
template <int N, typename R>
__device__ __forceinline__ void herm_combine(cx<R> (&z)[Sched<N>::E], const HermRaw<N, R> &h, int t) {
  constexpr int E = Sched<N>::E, TPP = N / E, H = N / 2;
#pragma unroll
  for (int c = 0; c < E; c++) {
    const int pos = t + TPP * c;
    const cx<R> x = h.x[c], y = h.y[c];
    if (pos == 0 || pos == H) z[c] = mk<R>(x.x, y.x);
    else if (pos < H) z[c] = mk<R>(x.x - y.y, x.y + y.x);
    else z[c] = mk<R>(x.x + y.y, y.x - x.y);
  }
}
template <int N, typename R>
__device__ __forceinline__ void herm_pair(cx<R> (&z)[Sched<N>::E], const cx<R> *X, const cx<R> *Y, int t,
                                          int lim = N / 2 + 1) {
  HermRaw<N, R> h;
  herm_load<N, R>(h, X, Y, t, lim);
  herm_combine<N, R>(z, h, t);
}

// Split the forward FFT F of (a + i b), a and b real, into the half spectra
// A_k = (F_k + conj F_{N-k})/2 and B_k = -i (F_k - conj F_{N-k})/2, k <= N/2.
template <int N, typename R, class Slot>
__device__ __forceinline__ void split_store(const cx<R> (&f)[Sched<N>::E], cx<R> *A, cx<R> *B, int t, cx<R> *lds,
                                            Slot slot, int lim = N / 2 + 1) {
  constexpr int E = Sched<N>::E, TPP = N / E, H = N / 2;
#pragma unroll
  for (int c = 0; c < E; c++) lds[slot(t + TPP * c)] = f[c];
  __syncthreads();
#pragma unroll
  for (int c = 0; c <= E / 2; c++) {
    const int pos = t + TPP * c;
    if (pos <= H && pos < lim) {
      const cx<R> a = lds[slot(pos)];
      const cx<R> b = lds[slot(pos == 0 ? 0 : N - pos)];
      const R sx = (R)0.5 * (a.x + b.x), sy = (R)0.5 * (a.y - b.y);
      const R dx = (R)0.5 * (a.x - b.x), dy = (R)0.5 * (a.y + b.y);
      st<R>(A + pos, mk<R>(sx, sy));
      st<R>(B + pos, mk<R>(dy, -dx));
    }
  }
  __syncthreads();
}

// Twiddle pointer for one transform of the fused z pass.  Where a z pencil
// spans several waves in FP64 (n >= 864), the pointer's value is hidden from
// the optimiser before each transform (an empty asm statement, no
// instruction), so that each transform loads its twiddles (from L1/L2)
// instead of the compiler hoisting all of them to the kernel start and
// keeping them in registers across the five transforms: 258 -> 211..250
// registers, two waves per SIMD instead of one (n=1024: z pass 21.7 -> 20.1
// ms).  Elsewhere the hoisted twiddles measured faster (n=512 FP64 z pass
// 1.81 vs 2.04 ms, float 0.97 vs 1.04 ms).
template <int N, typename R> __device__ __forceinline__ const cx<R> *z_twiddles(const cx<R> *tw) {
  if constexpr (sizeof(R) == 8 && N / Sched<N>::E > 64) asm volatile("" : "+s"(tw));
  return tw;
}

// One row of the fused z pass: c2r of u and omega (MODE 0: omega_z given in
// oz0, u from the stale physical velocity), u x omega, forward transform of
// (N_x, N_y) stored; N_z returned in nz.  Loads one Hermitian pair at a
// time (register budget).  MODE 1, 2: h holds the raw loads of the row's
// first pair (W fields 0, 1) on entry.
template <int N, typename R, int MODE, typename UR, bool REC = false>
__device__ __forceinline__ void zrow(cx<R> *__restrict__ w, UR *__restrict__ usr, Geo g, R invn3,
                                     const cx<R> *__restrict__ tw, int t, cx<R> *lds, PencilSlot<N> slot,
                                     const R (&oz0)[Sched<N>::E], R (&nz)[Sched<N>::E], const HermRaw<N, R> &h,
                                     double *__restrict__ rec = nullptr) {
  constexpr int E = Sched<N>::E, TPP = N / E;
  R u[E], v[E], ww[E], ox[E], oy[E], oz[E];
  if constexpr (MODE == 0) {
    cx<R> z1[E];
#pragma unroll
    for (int c = 0; c < E; c++) {
      const long l = t + TPP * c;
      u[c] = (R)usr[l];
      v[c] = (R)usr[g.rs + l];
      ww[c] = (R)usr[2 * g.rs + l];
    }
    herm_pair<N, R>(z1, w + 3 * g.wfs, w + 4 * g.wfs, t);
    fft_pencil<N, R, true>(z1, t, lds, slot, z_twiddles<N, R>(tw));
#pragma unroll
    for (int c = 0; c < E; c++) {
      ox[c] = z1[c].x * invn3;
      oy[c] = z1[c].y * invn3;
      oz[c] = oz0[c];
    }
  } else {
    cx<R> z1[E], z2[E], z3[E];
    herm_combine<N, R>(z1, h, t);
    fft_pencil<N, R, true>(z1, t, lds, slot, z_twiddles<N, R>(tw));
    herm_pair<N, R>(z2, w + 2 * g.wfs, w + 3 * g.wfs, t);
    fft_pencil<N, R, true>(z2, t, lds, slot, z_twiddles<N, R>(tw));
    herm_pair<N, R>(z3, w + 4 * g.wfs, w + 5 * g.wfs, t);
    fft_pencil<N, R, true>(z3, t, lds, slot, z_twiddles<N, R>(tw));
#pragma unroll
    for (int c = 0; c < E; c++) {
      u[c] = z1[c].x * invn3;
      v[c] = z1[c].y * invn3;
      ww[c] = z2[c].x * invn3;
      ox[c] = z2[c].y * invn3;
      oy[c] = z3[c].x * invn3;
      oz[c] = z3[c].y * invn3;
    }
    if constexpr (MODE == 2) {
#pragma unroll
      for (int c = 0; c < E; c++) {
        const long l = t + TPP * c;
        st_real<UR>(usr + l, (UR)u[c]);
        st_real<UR>(usr + g.rs + l, (UR)v[c]);
        st_real<UR>(usr + 2 * g.rs + l, (UR)ww[c]);
      }
    }
  }
  if constexpr (REC) {  // the adjoint's record of u and omega (physical)
#pragma unroll
    for (int c = 0; c < E; c++) {
      const long l = t + TPP * c;
      st_real<double>(rec + l, (double)u[c]);
      st_real<double>(rec + g.rs + l, (double)v[c]);
      st_real<double>(rec + 2 * g.rs + l, (double)ww[c]);
      st_real<double>(rec + 3 * g.rs + l, (double)ox[c]);
      st_real<double>(rec + 4 * g.rs + l, (double)oy[c]);
      st_real<double>(rec + 5 * g.rs + l, (double)oz[c]);
    }
  }
  cx<R> f1[E];
#pragma unroll
  for (int c = 0; c < E; c++) {
    f1[c] = mk<R>(v[c] * oz[c] - ww[c] * oy[c], ww[c] * ox[c] - u[c] * oz[c]);
    nz[c] = u[c] * oy[c] - v[c] * ox[c];
  }
  fft_pencil<N, R, false>(f1, t, lds, slot, z_twiddles<N, R>(tw));
  split_store<N, R>(f1, w, w + g.wfs, t, lds, slot, g.nk);
}

// C: fused z pass of one RK stage, one row per wave.
//  MODE 0 (stage 1): u,v,w from the stale physical velocity `us`,
//                    omega from W fields 3..5;
//  MODE 1 (stages 2,3): u,v,w,omega from W fields 0..5;
//  MODE 2 (stage 4): as MODE 1 and the physical u,v,w are written to `us`.
// Output: N = u x omega (main.c:374-376) transformed along z into W 0..2,
// kz < kmax only (the rest is removed by the dealias mask).
// (N_x, N_y) are transformed and stored before N_z is transformed, so the
// two forward transforms are never live together (9 VGPRs fewer in float
// when measured).  Grid n*n/PPW blocks.
// REC: u and omega of every point are also stored (double) into rec, six
// physical fields [6][n][n][n] (the adjoint's forward recomputation).
template <int N, typename R, int PPW, int MODE, typename UR, int MINW = 1, bool REC = false>
__global__ __launch_bounds__(PPW *(N / Sched<N>::E), MINW) void k_zfused(cx<R> *__restrict__ W, UR *__restrict__ us,
                                                                    Geo g, R invn3,
                                                                    const cx<R> *__restrict__ tw,
                                                                    double *__restrict__ rec = nullptr) {
  constexpr int E = Sched<N>::E, TPP = N / E, H = N / 2;
  __shared__ cx<R> lds[PPW * padlen<N>()];
  const int p = threadIdx.x / TPP, t = threadIdx.x % TPP;
  const long row = (long)blockIdx.x * PPW + p;
  us += blockIdx.z * g.mR;
  cx<R> *w = W + blockIdx.z * g.mW + (row / N) * g.wps + (row % N) * g.nfp;
  const PencilSlot<N> slot{p * padlen<N>()};
  R oz0[E], nz[E];
  HermRaw<N, R> h;
  if constexpr (MODE == 0) {
    cx<R> z[E];
    herm_pair<N, R>(z, w + 5 * g.wfs, nullptr, t);
    fft_pencil<N, R, true>(z, t, lds, slot, z_twiddles<N, R>(tw));
#pragma unroll
    for (int c = 0; c < E; c++) oz0[c] = z[c].x * invn3;
  } else {
    herm_load<N, R>(h, w, w + g.wfs, t);
  }
  zrow<N, R, MODE, UR, REC>(w, us + row * N, g, invn3, tw, t, lds, slot, oz0, nz, h,
                            REC ? rec + row * N : nullptr);
  // N_z is the transform of a real pencil: its first half is the r2c output.
  cx<R> f2[E];
#pragma unroll
  for (int c = 0; c < E; c++) f2[c] = mk<R>(nz[c], (R)0);
  fft_pencil<N, R, false>(f2, t, lds, slot, z_twiddles<N, R>(tw));
#pragma unroll
  for (int c = 0; c <= E / 2; c++) {
    const int pos = t + TPP * c;
    if (pos <= H && pos < g.nk) st<R>(w + 2 * g.wfs + pos, f2[c]);
  }
}

// Plain r2c along z of `nfields` physical double fields into spectral
// fields of precision R (pairs share one complex transform).
template <int N, typename R, int PPW>
__global__ __launch_bounds__(PPW *(N / Sched<N>::E)) void k_zr2c(const double *__restrict__ in, cx<R> *out,
                                                                  int nfields, Geo g,
                                                                  const cx<R> *__restrict__ tw) {
  constexpr int E = Sched<N>::E, TPP = N / E, H = N / 2;
  __shared__ cx<R> lds[PPW * padlen<N>()];
  const int p = threadIdx.x / TPP, t = threadIdx.x % TPP;
  const long row = (long)blockIdx.x * PPW + p;
  const PencilSlot<N> slot{p * padlen<N>()};
  for (int f = 0; f < nfields; f += 2) {
    const bool pair = f + 1 < nfields;
    cx<R> z[E];
#pragma unroll
    for (int c = 0; c < E; c++) {
      const long l = row * N + t + TPP * c;
      z[c] = mk<R>((R)in[f * g.rs + l], pair ? (R)in[(f + 1) * g.rs + l] : (R)0);
    }
    fft_pencil<N, R, false>(z, t, lds, slot, tw);
    cx<R> *o = out + row * g.nfp + f * g.fs;
    if (pair) {
      split_store<N, R>(z, o, o + g.fs, t, lds, slot);
    } else {
#pragma unroll
      for (int c = 0; c <= E / 2; c++) {
        const int pos = t + TPP * c;
        if (pos <= H) st<R>(o + pos, z[c]);
      }
    }
  }
}

// Plain c2r along z of one spectral field into a physical double field,
// multiplied by `scale` (1 for the raw transform, 1/n^3 for snapshots).
template <int N, typename R, int PPW>
__global__ __launch_bounds__(PPW *(N / Sched<N>::E)) void k_zc2r(const cx<R> *__restrict__ in,
                                                                  double *__restrict__ out, Geo g, R scale,
                                                                  const cx<R> *__restrict__ tw) {
  constexpr int E = Sched<N>::E, TPP = N / E;
  __shared__ cx<R> lds[PPW * padlen<N>()];
  const int p = threadIdx.x / TPP, t = threadIdx.x % TPP;
  const long row = (long)blockIdx.x * PPW + p;
  const PencilSlot<N> slot{p * padlen<N>()};
  cx<R> z[E];
  herm_pair<N, R>(z, in + row * g.nfp, nullptr, t);
  fft_pencil<N, R, true>(z, t, lds, slot, tw);
#pragma unroll
  for (int c = 0; c < E; c++) st_real<double>(out + row * N + t + TPP * c, (double)(z[c].x * scale));
}

// ---------------------------------------------------------------------------
// E: forward FFT along x of the three N partials + spectral epilogue.
//  dU = N*(dealias*dt); P = kk>0 ? (dU kx + dV ky + dW kz)/kk : 0;
//  dU -= P kx + (nu dt kk) Ur   (main.c:381-399, same evaluation order)
//  STAGE 0  : Ur = Un;  Ur' = Un + b dU;  acc = Un + a dU
//  STAGE 1,2: Ur' = Un + b dU;  acc += a dU
//  STAGE 3  : Un = acc + a dU                     (main.c:400-420)
// Mixed precision (SR = float) keeps Un and Un+1 in double but accumulates
// the increments in a float running sum: `acc` then holds, per member, a
// float copy of Un and S = a0 dU0 + a1 dU1 + a2 dU2 (each dU is only float-
// accurate anyway, being made from the float nonlinear term); stage 3 forms
// (Un + S) + a3 dU3 in double and also writes its float copy, which stages
// 0..2 of the next step read (whoever sets Un from outside refreshes it).
// S64 (mixed state only) keeps S in double instead: the increments are
// float-accurate, their sum is formed and stored in double.
// The epilogue runs in double for the FP64 state; for the mixed state dU
// and P are evaluated in float (their inputs are float) and only the stage
// 3 sum Un + S + a3 dU is formed in double.  phat (optional, stage 3)
// receives P.
// The new state (Ur' for stages 0..2, Un for stage 3) is returned in
// U,V,Wv as the next stage's pass A would read it back from memory.
// Grid (ceil(nf/T), n): blockIdx.y = j.
// EPI selects the epilogue's batching: 0 = default (below), otherwise
// BC = EPI & 63 positions per batch and prefetch of the next batch if
// EPI & 64.  E = elements per thread (a multiple of every radix of N).
template <int N, typename R, int T, int STAGE, typename SR, int EPI = 0, int E = Sched<N>::E, bool S64 = false>
__device__ __forceinline__ void forward_x_rk(const cx<R> *__restrict__ W, cx<SR> *__restrict__ uhr,
                                             double2 *__restrict__ uhn, double2 *__restrict__ acc,
                                             double2 *__restrict__ phat, Geo g, RKParams q,
                                             const cx<R> *__restrict__ tw, int t, int p, int j, int k, bool ok,
                                             cx<R> *lds, cx<R> (&U)[E], cx<R> (&V)[E], cx<R> (&Wv)[E]) {
#pragma clang fp contract(off)
  constexpr int TPP = N / E;
  const long off = (long)j * g.nfp + k, istr = g.wps;  // step partials
  // state element (i, j, k) at srow(jrow + i, k, g) (state layout)
  const long jrow = (long)j * N;
  const TileSlot<N, T> slot{p};
  constexpr bool MIXED_STATE = sizeof(SR) == 4;
  // two positions per batch (six or nine loads; one position for
  // 16-element threads), next batch prefetched in float tiles of up to 8
  // elements (register budget)
  constexpr int BC = EPI ? (EPI & 63) : E >= 16 ? 1 : 2;
  constexpr bool PF = EPI ? (EPI & 64) != 0 : sizeof(R) == 4 && E < 16;
  static_assert(E % BC == 0, "batch must divide the elements per thread");
  W += blockIdx.z * g.mW;
  uhr += blockIdx.z * g.mS;
  uhn += blockIdx.z * g.mS;
  // mixed: [float Un | S], S float or (S64) double
  float2 *inc = (float2 *)acc + blockIdx.z * (S64 ? 3 : 2) * g.mS;
  double2 *s64 = (double2 *)(inc + 3 * g.fs);
  acc += blockIdx.z * g.mS;
  if (phat) phat += blockIdx.z * g.fs;
  const double2 nd = q.nudt[blockIdx.z];
  const double nu = nd.x, dt = nd.y;
  // The y and z passes produced the nonlinear term only inside the
  // dealiased box (|ky|, kz < kmax); tiles outside it have N = 0 and skip the
  // transform (block-uniform branch).
  const bool nl = in_box(j, N, g.nk) && k - p < g.nk;
  const bool okn = k < g.nk;
  cx<R> F[3][E];
  if (nl) {
#pragma unroll
    for (int f = 0; f < 3; f++) {
#pragma unroll
      for (int c = 0; c < E; c++) F[f][c] = okn ? ld_row<T>(W + f * g.wfs + off + (t + TPP * c) * istr) : zero<R>();
    }
#pragma unroll
    for (int f = 0; f < 3; f++) fft_pencil<N, R, false>(F[f], t, lds, slot, tw);
  } else {
#pragma unroll
    for (int f = 0; f < 3; f++)
#pragma unroll
      for (int c = 0; c < E; c++) F[f][c] = zero<R>();
  }
  const double ky = (double)wave<N>(j), kz = (double)k;
  const double nudt = nu * dt;
  const bool myz = fabs(ky) < q.kmax && fabs(kz) < q.kmax;
  const float nudtf = (float)nudt, af = (float)q.a, bf = (float)q.b;  // mixed stages 0..2
  // The state values of a batch of positions (all three fields) are loaded
  // together and the next batch's loads are issued before this batch's
  // stores: the stores may alias the loads as far as the compiler can tell,
  // so a per-element order would cost one memory round trip per element
  // (and each wait would also drain the earlier stores).
  // Ur as stored: stage 0 reads Un (double; mixed: its float copy)
  using URT = std::conditional_t<STAGE == 0 && !MIXED_STATE, double2, cx<SR>>;
  using URS = std::conditional_t<STAGE == 0 && !MIXED_STATE, double, SR>;
  struct Batch {
    URT ur[3][BC];      // stage state (stage 0: Un)
    double2 un[3][BC];  // FP64 stages 1,2 and mixed stage 3: Un
    float2 unf[3][BC];  // mixed stages 1,2: float copy of Un
    double2 ac[3][BC];  // FP64 stages 1..3: accumulator
    std::conditional_t<S64, double2, float2> sf[3][BC];  // mixed stages 1..3: running sum S
  };
  constexpr int NB = E / BC;
  auto load = [&](Batch &B, auto BB) {
    constexpr int c0 = decltype(BB)::value * BC;
    if (!ok) return;
#pragma unroll
    for (int i = 0; i < BC; i++) {
#pragma unroll
      for (int f = 0; f < 3; f++) {
        const long lf = f * g.fs + srow(jrow + t + TPP * (c0 + i), k, g);
        if constexpr (STAGE == 0 && MIXED_STATE) B.ur[f][i] = ld_state<T>(inc + lf);
        else if constexpr (STAGE == 0) B.ur[f][i] = ld_state<T>(uhn + lf);
        else B.ur[f][i] = ld_state<T>(uhr + lf);
        if constexpr (MIXED_STATE) {
          if constexpr (STAGE == 1 || STAGE == 2) B.unf[f][i] = ld_state<T>(inc + lf);
          if constexpr (STAGE == 3) B.un[f][i] = ld_state<T>(uhn + lf);
          if constexpr (STAGE > 0 && S64) B.sf[f][i] = ld_state<T>(s64 + lf);
          else if constexpr (STAGE > 0) B.sf[f][i] = ld_state<T>(inc + 3 * g.fs + lf);
        } else {
          if constexpr (STAGE == 1 || STAGE == 2) B.un[f][i] = ld_state<T>(uhn + lf);
          if constexpr (STAGE > 0) B.ac[f][i] = ld_state<T>(acc + lf);
        }
      }
    }
  };
  auto apply = [&](const Batch &B, auto BB) {
    constexpr int c0 = decltype(BB)::value * BC;
#pragma unroll
    for (int i = 0; i < BC; i++) {
      const int c = c0 + i;
      U[c] = V[c] = Wv[c] = zero<R>();
      if (!ok) continue;
      const long l = srow(jrow + t + TPP * c, k, g);
      const double kx = (double)wave<N>(t + TPP * c);
      const double kk = kx * kx + ky * ky + kz * kz;
      const double fac = (myz && fabs(kx) < q.kmax) ? dt : 0.0;
      if constexpr (MIXED_STATE) {
        // Mixed state: dU is made from the float nonlinear term and the float
        // stage state, so the expressions of main.c:381-399 are evaluated in
        // float (same order; wavenumbers and kk are exact integers in float).
        // Stages 0..2 store float results (Ur', S); stage 3 forms
        // Un+1 = (Un + S) + a3 dU in double (main.c:415-420).
        const float kxf = (float)kx, kkf = (float)kk, facf = (float)fac;
        const float kvf[3] = {kxf, (float)ky, (float)kz};
        float2 e[3];
#pragma unroll
        for (int f = 0; f < 3; f++) e[f] = make_float2((float)F[f][c].x * facf, (float)F[f][c].y * facf);
        float2 Pf = make_float2(0.f, 0.f);
        if (kkf > 0) {
          Pf.x = (e[0].x * kvf[0] + e[1].x * kvf[1] + e[2].x * kvf[2]) / kkf;
          Pf.y = (e[0].y * kvf[0] + e[1].y * kvf[1] + e[2].y * kvf[2]) / kkf;
        }
        if constexpr (STAGE == 3) {
          if (phat) st<double>(phat + l, make_double2((double)Pf.x, (double)Pf.y));
        }
        const float vkf = nudtf * kkf;
        cx<R> nxt[3];
#pragma unroll
        for (int f = 0; f < 3; f++) {
          const long lf = f * g.fs + l;
          const float2 ur = cvt<float, URS>(B.ur[f][i]);
          e[f].x -= Pf.x * kvf[f] + vkf * ur.x;
          e[f].y -= Pf.y * kvf[f] + vkf * ur.y;
          const float2 du = e[f];
          if constexpr (STAGE == 3) {
            const double2 un = B.un[f][i];
            const auto sa = B.sf[f][i];
            double2 nn = make_double2(un.x + (double)sa.x, un.y + (double)sa.y);
            nn = make_double2(nn.x + q.a * (double)du.x, nn.y + q.a * (double)du.y);
            st<double>(uhn + lf, nn);
            st<float>(inc + lf, cvt<float, double>(nn));
            nxt[f] = cvt<R, double>(nn);
            continue;
          }
          const float2 un = STAGE == 0 ? ur : B.unf[f][i];
          const float2 nr = make_float2(un.x + bf * du.x, un.y + bf * du.y);
          st<SR>(uhr + lf, cvt<SR, float>(nr));
          if constexpr (S64) {
            const double2 ad = make_double2(q.a * (double)du.x, q.a * (double)du.y);
            if constexpr (STAGE == 0) {
              st<double>(s64 + lf, ad);
            } else {
              const double2 sa = B.sf[f][i];
              st<double>(s64 + lf, make_double2(sa.x + ad.x, sa.y + ad.y));
            }
          } else {
            const float2 sa = STAGE == 0 ? make_float2(0.f, 0.f) : B.sf[f][i];
            st<float>(inc + 3 * g.fs + lf,
                      STAGE == 0 ? make_float2(af * du.x, af * du.y) : make_float2(sa.x + af * du.x, sa.y + af * du.y));
          }
          nxt[f] = cvt<R, float>(nr);
        }
        U[c] = nxt[0];
        V[c] = nxt[1];
        Wv[c] = nxt[2];
      } else {
        double2 d[3];
#pragma unroll
        for (int f = 0; f < 3; f++) d[f] = make_double2((double)F[f][c].x * fac, (double)F[f][c].y * fac);
        double2 P = make_double2(0.0, 0.0);
        if (kk > 0) {
          P.x = (d[0].x * kx + d[1].x * ky + d[2].x * kz) / kk;
          P.y = (d[0].y * kx + d[1].y * ky + d[2].y * kz) / kk;
        }
        if (phat) st<double>(phat + l, P);
        const double vk = nudt * kk;
        const double kv[3] = {kx, ky, kz};
        cx<R> nxt[3];
#pragma unroll
        for (int f = 0; f < 3; f++) {
          const long lf = f * g.fs + l;
          const double2 ur = cvt<double, URS>(B.ur[f][i]);
          d[f].x -= P.x * kv[f] + vk * ur.x;
          d[f].y -= P.y * kv[f] + vk * ur.y;
          const double2 du = d[f];
          if constexpr (STAGE == 0) {
            const cx<SR> nr = cvt<SR, double>(make_double2(ur.x + q.b * du.x, ur.y + q.b * du.y));
            st<SR>(uhr + lf, nr);
            st<double>(acc + lf, make_double2(ur.x + q.a * du.x, ur.y + q.a * du.y));
            nxt[f] = cvt<R, SR>(nr);
          } else if constexpr (STAGE < 3) {
            const double2 un = B.un[f][i];
            const cx<SR> nr = cvt<SR, double>(make_double2(un.x + q.b * du.x, un.y + q.b * du.y));
            st<SR>(uhr + lf, nr);
            const double2 ac = B.ac[f][i];
            st<double>(acc + lf, make_double2(ac.x + q.a * du.x, ac.y + q.a * du.y));
            nxt[f] = cvt<R, SR>(nr);
          } else {
            const double2 ac = B.ac[f][i];
            const double2 nn = make_double2(ac.x + q.a * du.x, ac.y + q.a * du.y);
            st<double>(uhn + lf, nn);
            nxt[f] = cvt<R, double>(nn);
          }
        }
        U[c] = nxt[0];
        V[c] = nxt[1];
        Wv[c] = nxt[2];
      }
    }
  };
  if constexpr (PF) {
    Batch cur;
    load(cur, std::integral_constant<int, 0>{});
    static_for<0, NB>([&](auto BB) {
      constexpr int b = decltype(BB)::value;
      Batch nxb;
      if constexpr (b + 1 < NB) load(nxb, std::integral_constant<int, b + 1>{});
      apply(cur, BB);
      if constexpr (b + 1 < NB) cur = nxb;
    });
  } else {
    static_for<0, NB>([&](auto BB) {
      Batch cur;
      load(cur, BB);
      apply(cur, BB);
    });
  }
}

// E of stage STAGE fused with A of the following stage (stage STAGE+1, or
// stage 1 of the next step after STAGE 3): the new state goes from the RK
// update straight into the curl prologue and inverse x transform, saving
// the re-read of the state.  W fields 0..2 (this stage's N) are read before
// the same pencils' fields 0..5 are overwritten with the next partials.
template <int N, typename R, int T, int STAGE, typename SR, int MINW = 1, int EPI = 0, int E = Sched<N>::E,
          bool CX = false, bool S64 = false>
__global__ __launch_bounds__(T *(N / E), MINW) void k_xfwd_rk_xinv(cx<R> *__restrict__ W,
                                                                       cx<SR> *__restrict__ uhr,
                                                                       double2 *__restrict__ uhn,
                                                                       double2 *__restrict__ acc,
                                                                       double2 *__restrict__ phat, Geo g,
                                                                       RKParams q, const cx<R> *__restrict__ tw) {
  __shared__ cx<R> lds[T * padlen<N>()];
  __shared__ cx<R> twl[tw_size<N>()];
  int bx, j;
  if (!x_tile<T, R>((N / 2 + T) / T, N, bx, j)) return;
  tw = stage_twiddles<N, R>(twl, tw);
  const int p = threadIdx.x % T, t = threadIdx.x / T;
  const int k = bx * T + p;
  const bool ok = k < N / 2 + 1;
  cx<R> U[E], V[E], Wv[E];
  forward_x_rk<N, R, T, STAGE, SR, EPI, E, S64>(W, uhr, uhn, acc, phat, g, q, tw, t, p, j, k, ok, lds, U, V, Wv);
  curl_inverse_x<N, R, T, (STAGE < 3), E, CX>(U, V, Wv, W + blockIdx.z * g.mW, g, t, p, j, k, ok, lds, tw);
}

// ---------------------------------------------------------------------------
// Energy / enstrophy (main.c:257-266): per element
//   e = |U|^2 + |V|^2 + |W|^2 (double, left to right), z = kk * e,
// summed with compensated (two-sum) accumulation per thread, then a
// compensated tree per block; partial sums go to part[4*block].
// XMAJOR: the three fields are in the x-major layout of the host API
// ((i*n + j)*nfp + k) instead of the state layout; the elements are visited
// in the same order either way, so both give bitwise the same sums.
__device__ __forceinline__ void two_sum_acc(double &s, double &c, double x) {
  const double t = s + x;
  const double bp = t - s;
  c += (s - (t - bp)) + (x - bp);
  s = t;
}

template <int BS, bool XMAJOR = false>
__global__ __launch_bounds__(BS) void k_diag(const double2 *__restrict__ uh, int n, Geo g,
                                             double *__restrict__ part) {
#pragma clang fp contract(off)
  const int nf = n / 2 + 1;
  uh += blockIdx.z * g.mS;
  part += 4L * blockIdx.z * gridDim.x;
  double es = 0, ec = 0, zs = 0, zc = 0;
  for (long row = blockIdx.x; row < (long)n * n; row += gridDim.x) {
    int i, j;
    state_row(row, n, i, j);
    const double kx = i < n / 2 ? i : i - n, ky = j < n / 2 ? j : j - n;
    for (int k = threadIdx.x; k < nf; k += BS) {
      const double2 *r = uh + (XMAJOR ? ((long)i * n + j) * g.nfp + k : srow(row, k, g));
      const double2 a = r[0], b = r[g.fs], c = r[2 * g.fs];
      const double e = (a.x * a.x + a.y * a.y) + (b.x * b.x + b.y * b.y) + (c.x * c.x + c.y * c.y);
      const double kk = kx * kx + ky * ky + (double)k * (double)k;
      two_sum_acc(es, ec, e);
      two_sum_acc(zs, zc, kk * e);
    }
  }
  __shared__ double sh[4][BS];
  sh[0][threadIdx.x] = es;
  sh[1][threadIdx.x] = ec;
  sh[2][threadIdx.x] = zs;
  sh[3][threadIdx.x] = zc;
  __syncthreads();
  for (int s = BS / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      double a = sh[0][threadIdx.x], ac = sh[1][threadIdx.x];
      two_sum_acc(a, ac, sh[0][threadIdx.x + s]);
      sh[0][threadIdx.x] = a;
      sh[1][threadIdx.x] = ac + sh[1][threadIdx.x + s];
      double z = sh[2][threadIdx.x], zcc = sh[3][threadIdx.x];
      two_sum_acc(z, zcc, sh[2][threadIdx.x + s]);
      sh[2][threadIdx.x] = z;
      sh[3][threadIdx.x] = zcc + sh[3][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < 4; q++) part[4 * blockIdx.x + q] = sh[q][0];
  }
}

// Final compensated reduction of `nb` block partials by one block: out[0..3]
// = (energy sum, its correction, enstrophy sum, its correction), or with
// SCALED out[0..1] = ((sum + correction) * s2) of each, s2 = (1/n^3)^2, the
// same operations the host applies to the four values (main.c:265-266).
template <int BS, bool SCALED = false>
__global__ __launch_bounds__(BS) void k_diag_final(const double *__restrict__ part, int nb,
                                                   double *__restrict__ out, double s2 = 0.0) {
#pragma clang fp contract(off)
  part += 4L * blockIdx.z * nb;
  out += 4 * blockIdx.z;
  double es = 0, ec = 0, zs = 0, zc = 0;
  for (int b = threadIdx.x; b < nb; b += BS) {
    two_sum_acc(es, ec, part[4 * b]);
    ec += part[4 * b + 1];
    two_sum_acc(zs, zc, part[4 * b + 2]);
    zc += part[4 * b + 3];
  }
  __shared__ double sh[4][BS];
  sh[0][threadIdx.x] = es;
  sh[1][threadIdx.x] = ec;
  sh[2][threadIdx.x] = zs;
  sh[3][threadIdx.x] = zc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0, ac = 0, z = 0, zcc = 0;
    for (int q = 0; q < BS; q++) {
      two_sum_acc(a, ac, sh[0][q]);
      ac += sh[1][q];
      two_sum_acc(z, zcc, sh[2][q]);
      zcc += sh[3][q];
    }
    if (SCALED) {
      out[0] = (a + ac) * s2;
      out[1] = (z + zcc) * s2;
    } else {
      out[0] = a;
      out[1] = ac;
      out[2] = z;
      out[3] = zcc;
    }
  }
}

}  // namespace dnsk
