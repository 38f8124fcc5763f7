// Device-side Stockham FFT engine for gfx950 (CDNA4).
//
// A length-N pencil is owned by TPP = N/E threads; thread t holds the E
// elements at positions t + TPP*c (c = 0..E-1) in registers.  Each stage of
// radix r performs E/r butterflies per thread:
//   butterfly j = t + q*TPP reads x[j + m*N/r] (= register q + m*E/r),
//   multiplies by w^(m*(j mod Ns)) with w = exp(-+2 pi i/(Ns r)), applies a
//   radix-r DFT and writes y[(j/Ns)*Ns*r + j%Ns + m*Ns].
// Between stages the values go through LDS; after the last stage
// (Ns = N/r) the outputs land back on the load positions t + TPP*c, so a
// pencil is read from and written to HBM with the same, coalesced pattern.
//
// Twiddles come from a per-N table (generated on the host in long double,
// rounded once to the working precision); stage 0 needs none.
#pragma once
#include <hip/hip_runtime.h>
#include <utility>

namespace dnsk {

template <typename R> struct cx_of;
template <> struct cx_of<double> { using type = double2; };
template <> struct cx_of<float> { using type = float2; };
template <typename R> using cx = typename cx_of<R>::type;

template <typename R> __device__ __forceinline__ cx<R> mk(R a, R b) {
  cx<R> z;
  z.x = a;
  z.y = b;
  return z;
}
template <typename R> __device__ __forceinline__ cx<R> cadd(cx<R> a, cx<R> b) { return mk<R>(a.x + b.x, a.y + b.y); }
template <typename R> __device__ __forceinline__ cx<R> csub(cx<R> a, cx<R> b) { return mk<R>(a.x - b.x, a.y - b.y); }
template <typename R> __device__ __forceinline__ cx<R> cmul(cx<R> a, cx<R> b) {
  return mk<R>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
template <typename R> __device__ __forceinline__ cx<R> cmulc(cx<R> a, cx<R> b) {  // a * conj(b)
  return mk<R>(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
// multiply by -i (forward) or +i (inverse)
template <typename R, bool INV> __device__ __forceinline__ cx<R> rot(cx<R> a) {
  return INV ? mk<R>(-a.y, a.x) : mk<R>(a.y, -a.x);
}

// Streaming stores: every pass writes each element exactly once and the next
// pass reads it from HBM, so stores use the non-temporal path.
#ifndef DNS_NT_STORES
#define DNS_NT_STORES 1
#endif
template <typename R> __device__ __forceinline__ void st(cx<R> *p, cx<R> v) {
#if DNS_NT_STORES
  typedef R v2 __attribute__((ext_vector_type(2)));
  v2 t = {v.x, v.y};
  __builtin_nontemporal_store(t, (v2 *)p);
#else
  *p = v;
#endif
}
// Non-temporal load of a complex element that its pass reads once: the x
// pass with these loads runs 3.7% (FP64) and 5% (float) faster at n=512,
// the inverse y pass 4%, the z pass 3-5% (its Hermitian loads fetch each
// element once and pass the mirrored half between lanes; loading both
// halves non-temporally instead costs 7%).
template <typename T> __device__ __forceinline__ T ldnt(const T *p) {
  typedef decltype(p->x) E;
  typedef E v2 __attribute__((ext_vector_type(2)));
  v2 t = __builtin_nontemporal_load((const v2 *)p);
  T r;
  r.x = t.x;
  r.y = t.y;
  return r;
}
// Load of one element of a tile row of COLS consecutive elements: non-
// temporal when the row covers whole 128-byte lines.  Narrower rows share
// their lines with the neighbouring tile's block, which finds the other
// half in L2 only with ordinary loads (n=1024, 64-byte rows: non-temporal
// loads cost 13%).
template <int COLS, typename T> __device__ __forceinline__ T ld_row(const T *p) {
  if constexpr (COLS * sizeof(T) >= 128) return ldnt(p);
  else return *p;
}
// Load of one element of a state-layout tile row (two x rows share each
// line; see srow): non-temporal when the tile's two rows cover whole lines.
template <int COLS, typename T> __device__ __forceinline__ T ld_state(const T *p) {
  if constexpr (2 * COLS * sizeof(T) >= 128) return ldnt(p);
  else return *p;
}
template <typename R> __device__ __forceinline__ void st_real(R *p, R v) {
#if DNS_NT_STORES
  __builtin_nontemporal_store(v, p);
#else
  *p = v;
#endif
}

// ---------------------------------------------------------------------------
// Radix schedules.  E = elements per thread (divisible by every radix), so
// a pencil takes N/E threads; radix-3 grids keep E at 6 or 12 (registers).
// The radix order of each size is the fastest of the fewest-stage orderings
// timed by dns_amd/tune_radix.py (one build per candidate); a build may
// replace one size's entry with -DDNS_SCHED_N=n -DDNS_SCHED_E=e
// -DDNS_SCHED_R=r0,r1,... (the tuner's candidate builds).
template <int N> struct SchedTable;
template <int N> struct Sched : SchedTable<N> {};
#define DNS_SCHED(N_, E_, ...)                                \
  template <> struct SchedTable<N_> {                         \
    static constexpr int E = E_;                              \
    static constexpr int r[] = {__VA_ARGS__};                 \
    static constexpr int NS = sizeof(r) / sizeof(r[0]);       \
  };
DNS_SCHED(8, 4, 4, 2)
DNS_SCHED(12, 12, 4, 3)
DNS_SCHED(16, 4, 4, 4)
DNS_SCHED(18, 6, 2, 3, 3)
DNS_SCHED(24, 12, 4, 2, 3)
DNS_SCHED(32, 8, 8, 4)
DNS_SCHED(36, 12, 4, 3, 3)
DNS_SCHED(48, 12, 4, 4, 3)
DNS_SCHED(54, 6, 2, 3, 3, 3)
DNS_SCHED(64, 8, 8, 8)
DNS_SCHED(72, 12, 4, 2, 3, 3)
DNS_SCHED(96, 12, 4, 2, 4, 3)
DNS_SCHED(108, 12, 4, 3, 3, 3)
DNS_SCHED(128, 8, 8, 8, 2)
DNS_SCHED(144, 12, 4, 4, 3, 3)
DNS_SCHED(162, 6, 2, 3, 3, 3, 3)
DNS_SCHED(192, 12, 4, 4, 3, 4)
DNS_SCHED(216, 12, 4, 2, 3, 3, 3)
DNS_SCHED(256, 8, 8, 4, 8)
DNS_SCHED(288, 12, 4, 4, 2, 3, 3)
DNS_SCHED(324, 12, 4, 3, 3, 3, 3)
DNS_SCHED(384, 12, 4, 4, 4, 2, 3)
DNS_SCHED(432, 12, 4, 4, 3, 3, 3)
DNS_SCHED(486, 6, 2, 3, 3, 3, 3, 3)
DNS_SCHED(512, 8, 8, 8, 8)
DNS_SCHED(576, 12, 4, 4, 4, 3, 3)
DNS_SCHED(648, 12, 4, 2, 3, 3, 3, 3)
DNS_SCHED(768, 12, 4, 4, 4, 4, 3)
DNS_SCHED(864, 12, 4, 4, 2, 3, 3, 3)
DNS_SCHED(972, 12, 4, 3, 3, 3, 3, 3)
DNS_SCHED(1024, 8, 8, 8, 4, 4)
#undef DNS_SCHED
#ifdef DNS_SCHED_N
template <> struct Sched<DNS_SCHED_N> {
  static constexpr int E = DNS_SCHED_E;
  static constexpr int r[] = {DNS_SCHED_R};
  static constexpr int NS = sizeof(r) / sizeof(r[0]);
};
#endif

template <int N> constexpr int stage_Ns(int s) {
  int ns = 1;
  for (int q = 0; q < s; q++) ns *= Sched<N>::r[q];
  return ns;
}
template <int N> constexpr int stage_twoff(int s) {
  int off = 0, ns = 1;
  for (int q = 0; q < s; q++) {
    off += ns * (Sched<N>::r[q] - 1);
    ns *= Sched<N>::r[q];
  }
  return off;
}
template <int N> constexpr int tw_size() { return stage_twoff<N>(Sched<N>::NS); }

template <int I, int END, class F> __device__ __forceinline__ void static_for(F &&f) {
  if constexpr (I < END) {
    f(std::integral_constant<int, I>{});
    static_for<I + 1, END>(f);
  }
}

// ---------------------------------------------------------------------------
// Radix-r DFTs in registers (unnormalised; INV selects exp(+i)).
template <typename R, bool INV> __device__ __forceinline__ void dft2(cx<R> *a) {
  cx<R> t = a[0];
  a[0] = cadd<R>(t, a[1]);
  a[1] = csub<R>(t, a[1]);
}
template <typename R, bool INV> __device__ __forceinline__ void dft4(cx<R> *a) {
  cx<R> s0 = cadd<R>(a[0], a[2]), d0 = csub<R>(a[0], a[2]);
  cx<R> s1 = cadd<R>(a[1], a[3]), d1 = rot<R, INV>(csub<R>(a[1], a[3]));
  a[0] = cadd<R>(s0, s1);
  a[2] = csub<R>(s0, s1);
  a[1] = cadd<R>(d0, d1);
  a[3] = csub<R>(d0, d1);
}
template <typename R, bool INV> __device__ __forceinline__ void dft3(cx<R> *a) {
  const R s3 = (R)0.86602540378443864676372317075294;  // sin(2 pi / 3)
  cx<R> t1 = cadd<R>(a[1], a[2]);
  cx<R> t2 = mk<R>(a[0].x - (R)0.5 * t1.x, a[0].y - (R)0.5 * t1.y);
  cx<R> t3 = rot<R, INV>(mk<R>(s3 * (a[1].x - a[2].x), s3 * (a[1].y - a[2].y)));
  a[0] = cadd<R>(a[0], t1);
  a[1] = cadd<R>(t2, t3);
  a[2] = csub<R>(t2, t3);
}
template <typename R, bool INV> __device__ __forceinline__ void dft8(cx<R> *a) {
  const R h = (R)0.70710678118654752440084436210485;  // 1/sqrt(2)
  cx<R> s[4], d[4];
#pragma unroll
  for (int m = 0; m < 4; m++) {
    s[m] = cadd<R>(a[m], a[m + 4]);
    d[m] = csub<R>(a[m], a[m + 4]);
  }
  // d[m] *= w8^m, w8 = exp(-+ i pi/4)
  if (INV) {
    d[1] = mk<R>(h * (d[1].x - d[1].y), h * (d[1].x + d[1].y));
    d[2] = mk<R>(-d[2].y, d[2].x);
    d[3] = mk<R>(-h * (d[3].x + d[3].y), h * (d[3].x - d[3].y));
  } else {
    d[1] = mk<R>(h * (d[1].x + d[1].y), h * (d[1].y - d[1].x));
    d[2] = mk<R>(d[2].y, -d[2].x);
    d[3] = mk<R>(h * (d[3].y - d[3].x), -h * (d[3].x + d[3].y));
  }
  dft4<R, INV>(s);
  dft4<R, INV>(d);
#pragma unroll
  for (int p = 0; p < 4; p++) {
    a[2 * p] = s[p];
    a[2 * p + 1] = d[p];
  }
}
template <typename R, int RAD, bool INV> __device__ __forceinline__ void dft(cx<R> *a) {
  if constexpr (RAD == 2) dft2<R, INV>(a);
  else if constexpr (RAD == 3) dft3<R, INV>(a);
  else if constexpr (RAD == 4) dft4<R, INV>(a);
  else dft8<R, INV>(a);
}

// Twiddle table staged in LDS by the whole block.  LDS reads are waited for
// separately from global memory operations, so a transform that follows a
// batch of global stores does not stall until those stores complete (a
// global twiddle load would: loads and stores share one counter).
template <int N, typename R>
__device__ __forceinline__ const cx<R> *stage_twiddles(cx<R> *twl, const cx<R> *__restrict__ tw) {
  for (int i = threadIdx.x; i < tw_size<N>(); i += blockDim.x) twl[i] = tw[i];
  __syncthreads();
  return twl;
}

// ---------------------------------------------------------------------------
// LDS slot maps.  One dummy row every 8 positions breaks the power-of-two
// strides of the first stages' scattered writes.
template <int N> __device__ __forceinline__ int padpos(int pos) { return pos + (pos >> 3); }
template <int N> constexpr int padlen() { return N + N / 8; }

// Tile of T pencils, pencil index fastest (strided-axis passes).
template <int N, int T> struct TileSlot {
  int p;
  __device__ __forceinline__ int operator()(int pos) const { return padpos<N>(pos) * T + p; }
};
// Pencil-major (contiguous-axis passes).
template <int N> struct PencilSlot {
  int base;
  __device__ __forceinline__ int operator()(int pos) const { return base + padpos<N>(pos); }
};

// In-register FFT of one pencil.  v[c] holds position t + TPP*c on entry
// and the transform's coefficient at the same position on exit.
// All threads of the block must call it together (it uses __syncthreads).
// E (elements per thread) defaults to the schedule's; any multiple of every
// radix of the schedule works (fewer threads per pencil, more registers).
template <int N, typename R, bool INV, class Slot, int E>
__device__ __forceinline__ void fft_pencil(cx<R> (&v)[E], int t, cx<R> *lds, Slot slot,
                                           const cx<R> *__restrict__ tw) {
  constexpr int NS = Sched<N>::NS, TPP = N / E;
  static_for<0, NS>([&](auto S) { static_assert(E % Sched<N>::r[decltype(S)::value] == 0, "radix must divide E"); });
  static_for<0, NS>([&](auto S) {
    constexpr int s = decltype(S)::value;
    constexpr int r = Sched<N>::r[s];
    constexpr int Ns = stage_Ns<N>(s);
    constexpr int toff = stage_twoff<N>(s);
    constexpr int Q = E / r;
#pragma unroll
    for (int q = 0; q < Q; q++) {
      const int j = t + q * TPP;
      cx<R> a[r];
#pragma unroll
      for (int m = 0; m < r; m++) a[m] = v[q + m * Q];
      if constexpr (s > 0) {
        const int k = j % Ns;
        const cx<R> *w = tw + toff + k * (r - 1);
#pragma unroll
        for (int m = 1; m < r; m++) a[m] = INV ? cmulc<R>(a[m], w[m - 1]) : cmul<R>(a[m], w[m - 1]);
      }
      dft<R, r, INV>(a);
      if constexpr (s == NS - 1) {
#pragma unroll
        for (int m = 0; m < r; m++) v[q + m * Q] = a[m];
      } else {
        const int base = (j / Ns) * Ns * r + (j % Ns);
#pragma unroll
        for (int m = 0; m < r; m++) lds[slot(base + m * Ns)] = a[m];
      }
    }
    if constexpr (s < NS - 1) {
      __syncthreads();
#pragma unroll
      for (int c = 0; c < E; c++) v[c] = lds[slot(t + TPP * c)];
      __syncthreads();
    }
  });
}

}  // namespace dnsk
