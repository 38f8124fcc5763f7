/*
 * CPU oracle: restatement of the reference RK4 loop -- TEST INFRASTRUCTURE
 * ONLY (see oracle.h).  Every block cites the main.c lines it follows; the
 * floating-point expression order of main.c is kept, and the file is
 * compiled with -ffp-contract=off like the reference's `c99` build line.
 *
 * Reference facts kept on purpose:
 *  - stage 1 of every step forms u x omega from the physical velocity left
 *    by stage 4 of the previous step (U,V,W are refreshed only for rk > 0,
 *    main.c:344-353); at step 0 that is the raw input (main.c:180-185);
 *  - energy/enstrophy are plain sums over the half spectrum divided by n^6
 *    with long double accumulation (main.c:257-266);
 *  - the dealias mask is the box |k_axis| < 2/3 (n/2+1) (main.c:231-239);
 *  - x/y Nyquist wavenumber is -n/2, z Nyquist is +n/2 (main.c:224-230).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"
#ifdef _OPENMP
#include <omp.h>
#endif

int orc__c2r_work(int n, const double *in, double *out, double *work);

typedef struct {
  double re, im;
} cpx;

struct orc_dns {
  long n, nf, n3, n3f;
  double nu, dt, invn3, kmax;
  double *kx, *kz, *kk;
  int *dealias;
  double *U, *V, *W, *CU, *CV, *CW, *Ut, *Vt, *Wt;
  cpx *Uh, *Vh, *Wh, *Uh0, *Vh0, *Wh0, *Uh1, *Vh1, *Wh1, *dU, *dV, *dW;
  cpx *cX, *cY, *cZ, *Ph;
};

static const double a_rk[] = {1 / 6.0, 1 / 3.0, 1 / 3.0, 1 / 6.0}; /* main.c:15 */
static const double b_rk[] = {0.5, 0.5, 1.0};                      /* main.c:16 */

static void r2c(const orc_dns *s, const double *in, cpx *out) {
  orc_rfft3((int)s->n, in, (double *)out);
}
/* main.c:23-30: copy to work, then c2r (work is destroyed). */
static void c2r(const orc_dns *s, const cpx *hat, double *real, cpx *work) {
  orc__c2r_work((int)s->n, (const double *)hat, real, (double *)work);
}

orc_dns *orc_create(int n_, double nu, double dt) {
  orc_dns *s = calloc(1, sizeof *s);
  long n = n_, nf = n / 2 + 1;
  s->n = n;
  s->nf = nf;
  s->n3 = n * n * n;
  s->n3f = n * n * nf;
  s->nu = nu;
  s->dt = dt;
  s->invn3 = 1.0 / s->n3; /* main.c:188 */
  size_t R = sizeof(double) * (size_t)s->n3, C = sizeof(cpx) * (size_t)s->n3f;
  double **reals[] = {&s->U, &s->V, &s->W, &s->CU, &s->CV, &s->CW, &s->Ut, &s->Vt, &s->Wt};
  for (unsigned q = 0; q < sizeof reals / sizeof *reals; q++)
    *reals[q] = calloc(1, R);
  cpx **cpxs[] = {&s->Uh, &s->Vh,  &s->Wh, &s->Uh0, &s->Vh0, &s->Wh0, &s->Uh1, &s->Vh1,
                  &s->Wh1, &s->dU, &s->dV, &s->dW,  &s->cX,  &s->cY,  &s->cZ,  &s->Ph};
  for (unsigned q = 0; q < sizeof cpxs / sizeof *cpxs; q++)
    *cpxs[q] = calloc(1, C);
  s->kx = malloc(sizeof(double) * n);
  s->kz = malloc(sizeof(double) * nf);
  s->kk = malloc(sizeof(double) * s->n3f);
  s->dealias = malloc(sizeof(int) * s->n3f);
  /* main.c:224-231 */
  for (long i = 0; i < n / 2; i++) {
    s->kx[i] = i;
    s->kz[i] = i;
  }
  s->kz[n / 2] = n / 2;
  for (long i = -n / 2; i < 0; i++)
    s->kx[i + n] = i;
  s->kmax = 2. / 3. * (n / 2 + 1);
  /* main.c:232-247 */
#pragma omp parallel for collapse(3)
  for (long i = 0; i < n; i++)
    for (long j = 0; j < n; j++)
      for (long k = 0; k < nf; k++) {
        long l = (i * n + j) * nf + k;
        s->dealias[l] = (fabs(s->kx[i]) < s->kmax) && (fabs(s->kx[j]) < s->kmax) &&
                        (fabs(s->kz[k]) < s->kmax);
        s->kk[l] = s->kx[i] * s->kx[i] + s->kx[j] * s->kx[j] + s->kz[k] * s->kz[k];
      }
  return s;
}

void orc_destroy(orc_dns *s) {
  if (!s)
    return;
  double *reals[] = {s->U, s->V, s->W, s->CU, s->CV, s->CW, s->Ut, s->Vt, s->Wt};
  for (unsigned q = 0; q < sizeof reals / sizeof *reals; q++)
    free(reals[q]);
  cpx *cpxs[] = {s->Uh, s->Vh, s->Wh, s->Uh0, s->Vh0, s->Wh0, s->Uh1, s->Vh1,
                 s->Wh1, s->dU, s->dV, s->dW, s->cX, s->cY, s->cZ, s->Ph};
  for (unsigned q = 0; q < sizeof cpxs / sizeof *cpxs; q++)
    free(cpxs[q]);
  free(s->kx);
  free(s->kz);
  free(s->kk);
  free(s->dealias);
  free(s);
}

int orc_n(const orc_dns *s) { return (int)s->n; }

int orc_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

int orc_load(orc_dns *s, const double *uvw) {
  size_t R = sizeof(double) * (size_t)s->n3;
  memcpy(s->U, uvw, R);
  memcpy(s->V, uvw + s->n3, R);
  memcpy(s->W, uvw + 2 * s->n3, R);
  /* main.c:249-251 */
  r2c(s, s->U, s->Uh);
  r2c(s, s->V, s->Vh);
  r2c(s, s->W, s->Wh);
  return 0;
}

int orc_step(orc_dns *s, long nsteps) {
  const long n = s->n, nf = s->nf, n3 = s->n3, n3f = s->n3f;
  const double invn3 = s->invn3, dt = s->dt, nu = s->nu;
  const double *kx = s->kx, *kz = s->kz, *kk = s->kk;
  const int *dealias = s->dealias;
  double *U = s->U, *V = s->V, *W = s->W, *CU = s->CU, *CV = s->CV, *CW = s->CW;
  cpx *Uh = s->Uh, *Vh = s->Vh, *Wh = s->Wh, *dU = s->dU, *dV = s->dV, *dW = s->dW;
  for (long st = 0; st < nsteps; st++) {
    /* main.c:334-342 */
#pragma omp parallel for
    for (long k = 0; k < n3f; k++) {
      s->Uh0[k] = Uh[k];
      s->Vh0[k] = Vh[k];
      s->Wh0[k] = Wh[k];
      s->Uh1[k] = Uh[k];
      s->Vh1[k] = Vh[k];
      s->Wh1[k] = Wh[k];
    }
    for (int rk = 0; rk < 4; rk++) {
      if (rk > 0) { /* main.c:344-353 */
        c2r(s, Uh, U, s->cX);
        c2r(s, Vh, V, s->cX);
        c2r(s, Wh, W, s->cX);
#pragma omp parallel for
        for (long k = 0; k < n3; k++) {
          U[k] *= invn3;
          V[k] *= invn3;
          W[k] *= invn3;
        }
      }
      /* main.c:354-362: curl = I * (...) = (-Im, Re) of the bracket */
#pragma omp parallel for collapse(3)
      for (long i = 0; i < n; i++)
        for (long j = 0; j < n; j++)
          for (long k = 0; k < nf; k++) {
            long l = (i * n + j) * nf + k;
            cpx z;
            z.re = kx[i] * Vh[l].re - kx[j] * Uh[l].re;
            z.im = kx[i] * Vh[l].im - kx[j] * Uh[l].im;
            s->cZ[l].re = -z.im;
            s->cZ[l].im = z.re;
            z.re = kz[k] * Uh[l].re - kx[i] * Wh[l].re;
            z.im = kz[k] * Uh[l].im - kx[i] * Wh[l].im;
            s->cY[l].re = -z.im;
            s->cY[l].im = z.re;
            z.re = kx[j] * Wh[l].re - kz[k] * Vh[l].re;
            z.im = kx[j] * Wh[l].im - kz[k] * Vh[l].im;
            s->cX[l].re = -z.im;
            s->cX[l].im = z.re;
          }
      /* main.c:363-371 (the c2r destroys curl*, which is dead after) */
      c2r(s, s->cX, CU, s->cX);
      c2r(s, s->cY, CV, s->cY);
      c2r(s, s->cZ, CW, s->cZ);
#pragma omp parallel for
      for (long k = 0; k < n3; k++) {
        CU[k] *= invn3;
        CV[k] *= invn3;
        CW[k] *= invn3;
      }
      /* main.c:372-377 */
#pragma omp parallel for
      for (long k = 0; k < n3; k++) {
        s->Ut[k] = V[k] * CW[k] - W[k] * CV[k];
        s->Vt[k] = W[k] * CU[k] - U[k] * CW[k];
        s->Wt[k] = U[k] * CV[k] - V[k] * CU[k];
      }
      /* main.c:378-380 */
      r2c(s, s->Ut, dU);
      r2c(s, s->Vt, dV);
      r2c(s, s->Wt, dW);
      /* main.c:381-386 */
#pragma omp parallel for
      for (long k = 0; k < n3f; k++) {
        double f = dealias[k] * dt;
        dU[k].re *= f;
        dU[k].im *= f;
        dV[k].re *= f;
        dV[k].im *= f;
        dW[k].re *= f;
        dW[k].im *= f;
      }
      /* main.c:387-399 */
#pragma omp parallel for collapse(3)
      for (long i = 0; i < n; i++)
        for (long j = 0; j < n; j++)
          for (long k = 0; k < nf; k++) {
            long l = (i * n + j) * nf + k;
            cpx P;
            if (kk[l] > 0) {
              P.re = (dU[l].re * kx[i] + dV[l].re * kx[j] + dW[l].re * kz[k]) / kk[l];
              P.im = (dU[l].im * kx[i] + dV[l].im * kx[j] + dW[l].im * kz[k]) / kk[l];
            } else {
              P.re = 0.0;
              P.im = 0.0;
            }
            s->Ph[l] = P;
            double vk = nu * dt * kk[l];
            dU[l].re -= P.re * kx[i] + vk * Uh[l].re;
            dU[l].im -= P.im * kx[i] + vk * Uh[l].im;
            dV[l].re -= P.re * kx[j] + vk * Vh[l].re;
            dV[l].im -= P.im * kx[j] + vk * Vh[l].im;
            dW[l].re -= P.re * kz[k] + vk * Wh[l].re;
            dW[l].im -= P.im * kz[k] + vk * Wh[l].im;
          }
      if (rk < 3) { /* main.c:400-407 */
#pragma omp parallel for
        for (long k = 0; k < n3f; k++) {
          Uh[k].re = s->Uh0[k].re + b_rk[rk] * dU[k].re;
          Uh[k].im = s->Uh0[k].im + b_rk[rk] * dU[k].im;
          Vh[k].re = s->Vh0[k].re + b_rk[rk] * dV[k].re;
          Vh[k].im = s->Vh0[k].im + b_rk[rk] * dV[k].im;
          Wh[k].re = s->Wh0[k].re + b_rk[rk] * dW[k].re;
          Wh[k].im = s->Wh0[k].im + b_rk[rk] * dW[k].im;
        }
      }
      /* main.c:408-413 */
#pragma omp parallel for
      for (long k = 0; k < n3f; ++k) {
        s->Uh1[k].re += a_rk[rk] * dU[k].re;
        s->Uh1[k].im += a_rk[rk] * dU[k].im;
        s->Vh1[k].re += a_rk[rk] * dV[k].re;
        s->Vh1[k].im += a_rk[rk] * dV[k].im;
        s->Wh1[k].re += a_rk[rk] * dW[k].re;
        s->Wh1[k].im += a_rk[rk] * dW[k].im;
      }
    }
    /* main.c:415-420 */
#pragma omp parallel for
    for (long k = 0; k < n3f; k++) {
      Uh[k] = s->Uh1[k];
      Vh[k] = s->Vh1[k];
      Wh[k] = s->Wh1[k];
    }
  }
  return 0;
}

static double cabs2(cpx z) { return z.re * z.re + z.im * z.im; } /* main.c:31-33 */

void orc_diag(const orc_dns *s, long double *E, long double *Z) {
  long double energy = 0.0, Omega = 0.0;
  const long n3f = s->n3f;
  /* main.c:258-266 */
#pragma omp parallel for reduction(+ : energy, Omega)
  for (long k = 0; k < n3f; k++) {
    energy += cabs2(s->Uh[k]) + cabs2(s->Vh[k]) + cabs2(s->Wh[k]);
    Omega += s->kk[k] * (cabs2(s->Uh[k]) + cabs2(s->Vh[k]) + cabs2(s->Wh[k]));
  }
  energy *= s->invn3 * s->invn3;
  Omega *= s->invn3 * s->invn3;
  *E = energy;
  *Z = Omega;
}

void orc_diag_d(const orc_dns *s, double *E, double *Z) {
  long double e, z;
  orc_diag(s, &e, &z);
  *E = (double)e;
  *Z = (double)z;
}

void orc_get_spectral(const orc_dns *s, double *out) {
  size_t C = sizeof(cpx) * (size_t)s->n3f;
  memcpy(out, s->Uh, C);
  memcpy((char *)out + C, s->Vh, C);
  memcpy((char *)out + 2 * C, s->Wh, C);
}

void orc_set_spectral(orc_dns *s, const double *in) {
  size_t C = sizeof(cpx) * (size_t)s->n3f;
  memcpy(s->Uh, in, C);
  memcpy(s->Vh, (const char *)in + C, C);
  memcpy(s->Wh, (const char *)in + 2 * C, C);
}

void orc_get_ustale(const orc_dns *s, double *out) {
  size_t R = sizeof(double) * (size_t)s->n3;
  memcpy(out, s->U, R);
  memcpy((char *)out + R, s->V, R);
  memcpy((char *)out + 2 * R, s->W, R);
}

void orc_set_ustale(orc_dns *s, const double *in) {
  size_t R = sizeof(double) * (size_t)s->n3;
  memcpy(s->U, in, R);
  memcpy(s->V, (const char *)in + R, R);
  memcpy(s->W, (const char *)in + 2 * R, R);
}

void orc_get_phat(const orc_dns *s, double *out) {
  memcpy(out, s->Ph, sizeof(cpx) * (size_t)s->n3f);
}

/* main.c:272-279: per variable copy, c2r, scale by 1/n^3. */
void orc_snapshot(const orc_dns *s, double *uvwp) {
  const cpx *vars[4] = {s->Uh, s->Vh, s->Wh, s->Ph};
  cpx *work = malloc(sizeof(cpx) * (size_t)s->n3f);
  for (int v = 0; v < 4; v++) {
    double *d = uvwp + (size_t)v * s->n3;
    c2r(s, vars[v], d, work);
#pragma omp parallel for
    for (long i = 0; i < s->n3; i++)
      d[i] *= s->invn3;
  }
  free(work);
}
