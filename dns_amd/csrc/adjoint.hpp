// Discrete adjoint of the RK4 step (FP64): its algebra, and the pointwise
// kernels of the sweep (the backward stage itself: adjoint_fused.hpp).
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
    const long l = srow(row, k, g);
    const double kk = kx * kx + ky * ky + (double)k * k;
#pragma unroll
    for (int f = 0; f < 3; f++) {
      const double2 x = U[f * g.fs + l];
      lam[f * g.fs + l] = k < nf ? make_double2(2 * kk * x.x * inv6, 2 * kk * x.y * inv6) : make_double2(0, 0);
    }
  }
}

// out = W^-1 in (spectral state layout, 3 fields) -- used for the final
// R2C adjoint.
static __global__ void k_hweight_inv(const double2 *__restrict__ in, double2 *__restrict__ out, int n, Geo g) {
  const int nf = n / 2 + 1;
  const long row = blockIdx.x;
  for (int k = threadIdx.x; k < nf; k += blockDim.x) {
    const long l = srow(row, k, g);
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
