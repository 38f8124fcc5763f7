// Fused passes of the backward RK4 stage (FP64): the transposes of the
// forward passes A..E, sharing their tiles, layouts and transforms.  With
// the notation of adjoint.hpp, backward stage r (r = 3..0) is
//   XA  lam_d = a_r lamU + b_r lamNext;  lamV = -(nu dt kk) lam_d;
//       lamNh = dt M P lam_d / W, inverse x transform (tiles of the dealiased
//       box only; the rest of lamNh is 0)                  [k_adj_xa, k_adj_xe]
//   YB  inverse y transform of the three lamN partials (rows of the box read,
//       all rows written)                                  [k_ypass INV TRUNC]
//   ZC  c2r along z of lamN (kz < kmax), with the recorded u_r, omega_r:
//       lam_u = sc_u (omega x lamN + lamS), lam_w = sc (lamN x u);
//       r > 0: r2c along z of lam_w (fields 0..2) and lam_u (3..5);
//       r = 0: lam_u is the adjoint of the stale velocity (physical),
//       lam_w alone is transformed                          [k_adj_z]
//   YD  forward y transform of the six (three) partials    [k_ypass]
//   XE  forward x transform of F_w, F_u; lamUr = lamV + W (Ck F_w + F_u);
//       lamUn = (r == 3 ? lamU : lamUn) + lamUr; then XA of stage r - 1 with
//       lamNext = lamUr from registers                      [k_adj_xe]
// u_r and omega_r are the physical fields the forward z pass computed at
// stage r (u_0 is the stale velocity), recorded by the forward recomputation
// (k_zfused with a record pointer), so no transform of the forward state is
// repeated here.
#pragma once
#include "adjoint.hpp"

namespace dnsk {

struct AdjStage {
  double a, b;     // a_r, b_r of the stage whose lamNh is formed (b = 0: no lamNext)
  double dt, nu, kmax;
};

// XA on registers: d[f][c] = lam_d of field f at x position t + TPP c of
// pencil (j, k).  Stores lamV (state layout) and the inverse x transform of
// lamNh into W fields 0..2 (partials layout).  d is overwritten.
template <int N, int T, int E>
__device__ __forceinline__ void adj_pre_x(double2 (&d)[3][E], double2 *__restrict__ lamV, double2 *__restrict__ W,
                                          Geo g, AdjStage q, int t, int p, int j, int k, bool ok, double2 *lds,
                                          const double2 *__restrict__ tw) {
#pragma clang fp contract(off)
  constexpr int TPP = N / E;
  const long jrow = (long)j * N;  // state rows (srow)
  const double ky = wave<N>(j), kz = k;
  const bool zy = fabs(ky) < q.kmax && fabs(kz) < q.kmax;
  const double sN0 = q.dt / hweight(k, N);
#pragma unroll
  for (int c = 0; c < E; c++) {
    const int i = t + TPP * c;
    const double kx = wave<N>(i);
    const double kk = kx * kx + ky * ky + kz * kz;
    const double kv[3] = {kx, ky, kz};
    const double sU = -(q.nu * q.dt * kk);
    if (ok) {
#pragma unroll
      for (int f = 0; f < 3; f++)
        st<double>(lamV + f * g.fs + srow(jrow + i, k, g), make_double2(sU * d[f][c].x, sU * d[f][c].y));
    }
    double2 pr[3];
    if (kk > 0) {
      const double px = (d[0][c].x * kx + d[1][c].x * ky + d[2][c].x * kz) / kk;
      const double py = (d[0][c].y * kx + d[1][c].y * ky + d[2][c].y * kz) / kk;
#pragma unroll
      for (int f = 0; f < 3; f++) pr[f] = make_double2(d[f][c].x - px * kv[f], d[f][c].y - py * kv[f]);
    } else {
#pragma unroll
      for (int f = 0; f < 3; f++) pr[f] = d[f][c];
    }
    const double sN = (zy && fabs(kx) < q.kmax) ? sN0 : 0.0;
#pragma unroll
    for (int f = 0; f < 3; f++) d[f][c] = make_double2(sN * pr[f].x, sN * pr[f].y);
  }
  // tiles outside the box in (ky, kz) carry lamNh = 0: nothing to transform
  // or store (the y pass reads the box only); block-uniform branch
  if (in_box(j, N, g.nk) && k - p < g.nk) {
    const long off = (long)j * g.nfp + k, istr = g.wps;
#pragma unroll
    for (int f = 0; f < 3; f++) {
      fft_pencil<N, double, true>(d[f], t, lds, TileSlot<N, T>{p}, tw);
      if (ok && k < g.nk) {
#pragma unroll
        for (int c = 0; c < E; c++) st<double>(W + f * g.wfs + off + (t + TPP * c) * istr, d[f][c]);
      }
    }
  }
}

// XA of backward stage 3 (no lamNext): grid (ceil(nf/T), n), blockIdx.y = j.
template <int N, int T, int E = Sched<N>::E>
__global__ __launch_bounds__(T *(N / E)) void k_adj_xa(const double2 *__restrict__ lamU, double2 *__restrict__ lamV,
                                                       double2 *__restrict__ W, Geo g, AdjStage q,
                                                       const double2 *__restrict__ tw) {
#pragma clang fp contract(off)
  constexpr int TPP = N / E;
  __shared__ double2 lds[T * padlen<N>()];
  __shared__ double2 twl[tw_size<N>()];
  tw = stage_twiddles<N, double>(twl, tw);
  const int p = threadIdx.x % T, t = threadIdx.x / T;
  const int j = blockIdx.y, k = blockIdx.x * T + p;
  const bool ok = k < N / 2 + 1;
  const long jrow = (long)j * N;  // state rows (srow)
  double2 d[3][E];
#pragma unroll
  for (int f = 0; f < 3; f++)
#pragma unroll
    for (int c = 0; c < E; c++) {
      const double2 x = ok ? lamU[f * g.fs + srow(jrow + t + TPP * c, k, g)] : make_double2(0, 0);
      d[f][c] = make_double2(q.a * x.x, q.a * x.y);
    }
  adj_pre_x<N, T, E>(d, lamV, W, g, q, t, p, j, k, ok, lds, tw);
}

// XE of backward stage r, fused with XA of stage r - 1 when NEXT.
//  FU   : F_u present (r > 0); F_w in W fields 0..2, F_u in 3..5
//  FIRST: r == 3, lamUn = lamU + lamUr (else lamUn += lamUr)
// qn: the parameters of stage r - 1 (a, b, dt, nu, kmax).
template <int N, int T, bool FU, bool FIRST, bool NEXT, int E = Sched<N>::E>
__global__ __launch_bounds__(T *(N / E)) void k_adj_xe(double2 *__restrict__ W, const double2 *__restrict__ lamU,
                                                       double2 *__restrict__ lamUn, double2 *__restrict__ lamV, Geo g,
                                                       AdjStage qn, const double2 *__restrict__ tw) {
#pragma clang fp contract(off)
  constexpr int TPP = N / E;
  __shared__ double2 lds[T * padlen<N>()];
  __shared__ double2 twl[tw_size<N>()];
  tw = stage_twiddles<N, double>(twl, tw);
  const int p = threadIdx.x % T, t = threadIdx.x / T;
  const int j = blockIdx.y, k = blockIdx.x * T + p;
  const bool ok = k < N / 2 + 1;
  const long off = (long)j * g.nfp + k, istr = g.wps;          // partials
  const long jrow = (long)j * N;                               // state rows (srow)
  const TileSlot<N, T> slot{p};
  double2 C[3][E];
#pragma unroll
  for (int f = 0; f < 3; f++) {
#pragma unroll
    for (int c = 0; c < E; c++)
      C[f][c] = ok ? ld_row<T>(W + f * g.wfs + off + (t + TPP * c) * istr) : make_double2(0, 0);
    fft_pencil<N, double, false>(C[f], t, lds, slot, tw);
  }
  const double ky = wave<N>(j), kz = k;
  // C = Ck F_w (i k x, same expressions as the forward curl, main.c:359-361)
#pragma unroll
  for (int c = 0; c < E; c++) {
    const double kx = wave<N>(t + TPP * c);
    const double2 X = C[0][c], Y = C[1][c], Z = C[2][c];
    C[0][c] = make_double2(-(ky * Z.y - kz * Y.y), ky * Z.x - kz * Y.x);
    C[1][c] = make_double2(-(kz * X.y - kx * Z.y), kz * X.x - kx * Z.x);
    C[2][c] = make_double2(-(kx * Y.y - ky * X.y), kx * Y.x - ky * X.x);
  }
  if constexpr (FU) {
#pragma unroll
    for (int f = 0; f < 3; f++) {
      double2 v[E];
#pragma unroll
      for (int c = 0; c < E; c++)
        v[c] = ok ? ld_row<T>(W + (3 + f) * g.wfs + off + (t + TPP * c) * istr) : make_double2(0, 0);
      fft_pencil<N, double, false>(v, t, lds, slot, tw);
#pragma unroll
      for (int c = 0; c < E; c++) C[f][c] = make_double2(C[f][c].x + v[c].x, C[f][c].y + v[c].y);
    }
  }
  const double wk = hweight(k, N);
#pragma unroll
  for (int c = 0; c < E; c++) {
    const long l = srow(jrow + t + TPP * c, k, g);
    double2 r[3], base[3], lu[3];
    if (ok) {
#pragma unroll
      for (int f = 0; f < 3; f++) {
        r[f] = lamV[f * g.fs + l];
        base[f] = FIRST ? lamU[f * g.fs + l] : lamUn[f * g.fs + l];
        if (NEXT) lu[f] = FIRST ? base[f] : lamU[f * g.fs + l];
      }
#pragma unroll
      for (int f = 0; f < 3; f++) {
        r[f].x += wk * C[f][c].x;
        r[f].y += wk * C[f][c].y;
        st<double>(lamUn + f * g.fs + l, make_double2(base[f].x + r[f].x, base[f].y + r[f].y));
        if (NEXT) {
          double2 dd = make_double2(qn.a * lu[f].x, qn.a * lu[f].y);
          dd.x += qn.b * r[f].x;
          dd.y += qn.b * r[f].y;
          C[f][c] = dd;
        }
      }
    } else {
#pragma unroll
      for (int f = 0; f < 3; f++) C[f][c] = make_double2(0, 0);
    }
  }
  if constexpr (NEXT) adj_pre_x<N, T, E>(C, lamV, W, g, qn, t, p, j, k, ok, lds, tw);
}

// ZC of backward stage r: one z row per pencil slot, PPW rows per block.
//  R0: r == 0 (lam_u to lamSout, only lam_w transformed); LS: r == 3 (lamS in).
// rec: u_r (fields 0..2) and omega_r (3..5), physical, [6][n][n][n].
template <int N, int PPW, bool R0, bool LS>
__global__ __launch_bounds__(PPW *(N / Sched<N>::E)) void k_adj_z(double2 *__restrict__ W,
                                                                   const double *__restrict__ rec,
                                                                   const double *__restrict__ lamS,
                                                                   double *__restrict__ lamSout, Geo g, double sc,
                                                                   const double2 *__restrict__ tw) {
#pragma clang fp contract(off)
  constexpr int E = Sched<N>::E, TPP = N / E, H = N / 2;
  __shared__ double2 lds[PPW * padlen<N>()];
  const int p = threadIdx.x / TPP, t = threadIdx.x % TPP;
  const long row = (long)blockIdx.x * PPW + p;
  double2 *w = W + (row / N) * g.wps + (row % N) * g.nfp;
  const PencilSlot<N> slot{p * padlen<N>()};
  double L0[E], L1[E], L2[E];
  {
    double2 z[E];
    herm_pair<N, double>(z, w, w + g.wfs, t, g.nk);
    fft_pencil<N, double, true>(z, t, lds, slot, tw);
#pragma unroll
    for (int c = 0; c < E; c++) {
      L0[c] = z[c].x;
      L1[c] = z[c].y;
    }
    herm_pair<N, double>(z, w + 2 * g.wfs, nullptr, t, g.nk);
    fft_pencil<N, double, true>(z, t, lds, slot, tw);
#pragma unroll
    for (int c = 0; c < E; c++) L2[c] = z[c].x;
  }
  const double sc_u = R0 ? 1.0 : sc;
  double2 f1[E], f2[E], f3[E];
#pragma unroll
  for (int c = 0; c < E; c++) {
    const long l = row * N + t + TPP * c;
    const double u0 = rec[l], u1 = rec[g.rs + l], u2 = rec[2 * g.rs + l];
    const double w0 = rec[3 * g.rs + l], w1 = rec[4 * g.rs + l], w2 = rec[5 * g.rs + l];
    double a0 = w1 * L2[c] - w2 * L1[c], a1 = w2 * L0[c] - w0 * L2[c], a2 = w0 * L1[c] - w1 * L0[c];
    if (LS) {
      a0 += lamS[l];
      a1 += lamS[g.rs + l];
      a2 += lamS[2 * g.rs + l];
    }
    const double lu0 = sc_u * a0, lu1 = sc_u * a1, lu2 = sc_u * a2;
    const double lw0 = sc * (L1[c] * u2 - L2[c] * u1);
    const double lw1 = sc * (L2[c] * u0 - L0[c] * u2);
    const double lw2 = sc * (L0[c] * u1 - L1[c] * u0);
    if (R0) {
      st_real<double>(lamSout + l, lu0);
      st_real<double>(lamSout + g.rs + l, lu1);
      st_real<double>(lamSout + 2 * g.rs + l, lu2);
      f1[c] = make_double2(lw0, lw1);
      f2[c] = make_double2(lw2, 0.0);
    } else {
      f1[c] = make_double2(lw0, lw1);
      f2[c] = make_double2(lw2, lu0);
      f3[c] = make_double2(lu1, lu2);
    }
  }
  fft_pencil<N, double, false>(f1, t, lds, slot, tw);
  split_store<N, double>(f1, w, w + g.wfs, t, lds, slot);
  fft_pencil<N, double, false>(f2, t, lds, slot, tw);
  if (R0) {
#pragma unroll
    for (int c = 0; c <= E / 2; c++) {
      const int pos = t + TPP * c;
      if (pos <= H) st<double>(w + 2 * g.wfs + pos, f2[c]);
    }
  } else {
    split_store<N, double>(f2, w + 2 * g.wfs, w + 3 * g.wfs, t, lds, slot);
    fft_pencil<N, double, false>(f3, t, lds, slot, tw);
    split_store<N, double>(f3, w + 4 * g.wfs, w + 5 * g.wfs, t, lds, slot);
  }
}

}  // namespace dnsk
