// Discrete adjoint of the RK4 step (FP64) -- pointwise kernels.
//
// Forward step (main.c:334-420), with s = 1/n^3, C2R/R2C the unnormalised
// 3-D transforms and Ck = i k x (curl):
//   u_r  = r == 0 ? u_stale : s C2R(U_r)        w_r = s C2R(Ck U_r)
//   N_r  = u_r x w_r                            Nh_r = R2C(N_r)
//   dU_r = dt P M Nh_r - nu dt K U_r            (M mask, P Leray, K = |k|^2)
//   U_{r+1} = U_n + b_r dU_r (r < 3);  U_{n+1} = U_n + sum a_r dU_r;
//   u_stale' = u_3.
// With <a,b> = Re sum conj(a) b over the stored half spectrum and the plain
// sum over physical points, the adjoints of the building blocks are
//   R2C* = C2R o W^-1,  C2R* = W o R2C    (W = 2 on 0 < kz < n/2, else 1),
//   Ck* = Ck (i [k]x is Hermitian), P, M, K symmetric,
//   (u x w)*: lam_u = w x lam_N,  lam_w = lam_N x u.
// Backward through a step, stages r = 3..0:
//   lam_dU = a_r lam_U' + b_r lam_U_{r+1}            (b term for r < 3)
//   lam_Nh = dt M P lam_dU;  lam_U_r = -nu dt K lam_dU
//   lam_N  = C2R(W^-1 lam_Nh)
//   lam_u  = w_r x lam_N (+ lam_stale' at r = 3);  lam_w = lam_N x u_r
//   lam_U_r += W (Ck R2C(s lam_w) + R2C(s lam_u))      (r > 0)
//   lam_stale = lam_u                                  (r = 0)
//   lam_U_n  = lam_U' + sum_r lam_U_r.
#pragma once
#include "kernels.hpp"

namespace dnsk {

__device__ __forceinline__ double hweight(int k, int n) { return (k == 0 || 2 * k == n) ? 1.0 : 2.0; }

// Spectral pre-step of one backward stage.  Elements (row, k) of the padded
// layout; grid covers n*n rows x nf columns.
//   ld = a * lamU + b * lamNext (lamNext may be null)
//   outN = dt M P ld / W      outU = -(nu dt kk) ld
static __global__ void k_adj_spec_pre(const double2 *__restrict__ lamU, const double2 *__restrict__ lamNext,
                               double2 *__restrict__ outN, double2 *__restrict__ outU, int n, Geo g, double a,
                               double b, double dt, double nu, double kmax) {
  const int nf = n / 2 + 1;
  const long row = blockIdx.x;
  int i, j;
  state_row(row, n, i, j);
  const double kx = i < n / 2 ? i : i - n, ky = j < n / 2 ? j : j - n;
  for (int k = threadIdx.x; k < nf; k += blockDim.x) {
    const long l = row * g.nfp + k;
    const double kz = k;
    const double kk = kx * kx + ky * ky + kz * kz;
    const double kv[3] = {kx, ky, kz};
    double2 d[3];
#pragma unroll
    for (int f = 0; f < 3; f++) {
      const double2 x = lamU[f * g.fs + l];
      d[f] = make_double2(a * x.x, a * x.y);
      if (lamNext) {
        const double2 y = lamNext[f * g.fs + l];
        d[f].x += b * y.x;
        d[f].y += b * y.y;
      }
    }
    const bool keep = fabs(kx) < kmax && fabs(ky) < kmax && fabs(kz) < kmax;
    double2 pr[3];
    if (kk > 0) {
      const double px = (d[0].x * kx + d[1].x * ky + d[2].x * kz) / kk;
      const double py = (d[0].y * kx + d[1].y * ky + d[2].y * kz) / kk;
#pragma unroll
      for (int f = 0; f < 3; f++) pr[f] = make_double2(d[f].x - px * kv[f], d[f].y - py * kv[f]);
    } else {
#pragma unroll
      for (int f = 0; f < 3; f++) pr[f] = d[f];
    }
    const double sN = keep ? dt / hweight(k, n) : 0.0;
    const double sU = -(nu * dt * kk);
#pragma unroll
    for (int f = 0; f < 3; f++) {
      outN[f * g.fs + l] = make_double2(sN * pr[f].x, sN * pr[f].y);
      outU[f * g.fs + l] = make_double2(sU * d[f].x, sU * d[f].y);
    }
  }
}

// Physical cross products of one backward stage (n^3 points):
//   lu = sc * (w x lamN + lamS),  lw = sc * (lamN x u)   (lamS may be null)
static __global__ void k_adj_cross(const double *__restrict__ lamN, const double *__restrict__ u,
                            const double *__restrict__ w, const double *__restrict__ lamS, double *__restrict__ lu,
                            double *__restrict__ lw, long rs, double sc_u, double sc_w) {
  for (long p = blockIdx.x * (long)blockDim.x + threadIdx.x; p < rs; p += (long)gridDim.x * blockDim.x) {
    const double L0 = lamN[p], L1 = lamN[rs + p], L2 = lamN[2 * rs + p];
    const double u0 = u[p], u1 = u[rs + p], u2 = u[2 * rs + p];
    const double w0 = w[p], w1 = w[rs + p], w2 = w[2 * rs + p];
    double a0 = w1 * L2 - w2 * L1, a1 = w2 * L0 - w0 * L2, a2 = w0 * L1 - w1 * L0;
    if (lamS) {
      a0 += lamS[p];
      a1 += lamS[rs + p];
      a2 += lamS[2 * rs + p];
    }
    lu[p] = sc_u * a0;
    lu[rs + p] = sc_u * a1;
    lu[2 * rs + p] = sc_u * a2;
    lw[p] = sc_w * (L1 * u2 - L2 * u1);
    lw[rs + p] = sc_w * (L2 * u0 - L0 * u2);
    lw[2 * rs + p] = sc_w * (L0 * u1 - L1 * u0);
  }
}

// Spectral post-step: lamUr += W (Ck Fw + Fu) (Fu may be null), then
// lamUn += lamUr.
static __global__ void k_adj_spec_post(const double2 *__restrict__ Fw, const double2 *__restrict__ Fu,
                                double2 *__restrict__ lamUr, double2 *__restrict__ lamUn, int n, Geo g) {
  const int nf = n / 2 + 1;
  const long row = blockIdx.x;
  int i, j;
  state_row(row, n, i, j);
  const double kx = i < n / 2 ? i : i - n, ky = j < n / 2 ? j : j - n;
  for (int k = threadIdx.x; k < nf; k += blockDim.x) {
    const long l = row * g.nfp + k;
    const double kz = k, wk = hweight(k, n);
    const double2 X = Fw[l], Y = Fw[g.fs + l], Z = Fw[2 * g.fs + l];
    // Ck = i k x: (i(ky Z - kz Y), i(kz X - kx Z), i(kx Y - ky X)); i z = (-Im, Re)
    double2 c[3];
    c[0] = make_double2(-(ky * Z.y - kz * Y.y), ky * Z.x - kz * Y.x);
    c[1] = make_double2(-(kz * X.y - kx * Z.y), kz * X.x - kx * Z.x);
    c[2] = make_double2(-(kx * Y.y - ky * X.y), kx * Y.x - ky * X.x);
#pragma unroll
    for (int f = 0; f < 3; f++) {
      double2 v = c[f];
      if (Fu) {
        const double2 q = Fu[f * g.fs + l];
        v.x += q.x;
        v.y += q.y;
      }
      double2 r = lamUr[f * g.fs + l];
      r.x += wk * v.x;
      r.y += wk * v.y;
      lamUr[f * g.fs + l] = r;
      double2 t = lamUn[f * g.fs + l];
      t.x += r.x;
      t.y += r.y;
      lamUn[f * g.fs + l] = t;
    }
  }
}

// Spectral curl into a separate array (forward recomputation of omega's
// spectrum): out = Ck in.
static __global__ void k_curl_spec(const double2 *__restrict__ in, double2 *__restrict__ out, int n, Geo g) {
  const int nf = n / 2 + 1;
  const long row = blockIdx.x;
  int i, j;
  state_row(row, n, i, j);
  const double kx = i < n / 2 ? i : i - n, ky = j < n / 2 ? j : j - n;
  for (int k = threadIdx.x; k < nf; k += blockDim.x) {
    const long l = row * g.nfp + k;
    const double kz = k;
    const double2 U = in[l], V = in[g.fs + l], Wz = in[2 * g.fs + l];
    // same expressions as main.c:359-361
    double2 z = make_double2(ky * Wz.x - kz * V.x, ky * Wz.y - kz * V.y);
    out[l] = make_double2(-z.y, z.x);
    z = make_double2(kz * U.x - kx * Wz.x, kz * U.y - kx * Wz.y);
    out[g.fs + l] = make_double2(-z.y, z.x);
    z = make_double2(kx * V.x - ky * U.x, kx * V.y - ky * U.y);
    out[2 * g.fs + l] = make_double2(-z.y, z.x);
  }
}

// Gradient of the enstrophy sum  Z = sum kk |U|^2 / n^6  (main.c:263-266):
// lam = 2 kk U / n^6 on the valid columns, 0 on the padding.
static __global__ void k_enstrophy_grad(const double2 *__restrict__ U, double2 *__restrict__ lam, int n, Geo g,
                                 double inv6) {
  const int nf = n / 2 + 1;
  const long row = blockIdx.x;
  int i, j;
  state_row(row, n, i, j);
  const double kx = i < n / 2 ? i : i - n, ky = j < n / 2 ? j : j - n;
  for (int k = threadIdx.x; k < g.nfp; k += blockDim.x) {
    const long l = row * g.nfp + k;
    const double kk = kx * kx + ky * ky + (double)k * k;
#pragma unroll
    for (int f = 0; f < 3; f++) {
      const double2 x = U[f * g.fs + l];
      lam[f * g.fs + l] = k < nf ? make_double2(2 * kk * x.x * inv6, 2 * kk * x.y * inv6) : make_double2(0, 0);
    }
  }
}

// out = W^-1 in (spectral, 3 fields) -- used for the final R2C adjoint.
static __global__ void k_hweight_inv(const double2 *__restrict__ in, double2 *__restrict__ out, int n, Geo g) {
  const int nf = n / 2 + 1;
  const long row = blockIdx.x;
  for (int k = threadIdx.x; k < nf; k += blockDim.x) {
    const long l = row * g.nfp + k;
    const double s = 1.0 / hweight(k, n);
#pragma unroll
    for (int f = 0; f < 3; f++) {
      const double2 x = in[f * g.fs + l];
      out[f * g.fs + l] = make_double2(s * x.x, s * x.y);
    }
  }
}

// y += a x (real arrays of `count` elements)
static __global__ void k_axpy_scaled(const double *__restrict__ x, double *__restrict__ y, double a, long count) {
  for (long p = blockIdx.x * (long)blockDim.x + threadIdx.x; p < count; p += (long)gridDim.x * blockDim.x)
    y[p] += a * x[p];
}

}  // namespace dnsk
