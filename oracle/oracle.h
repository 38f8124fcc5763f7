/*
 * CPU oracle for the dns RK4 time step -- TEST INFRASTRUCTURE ONLY.
 *
 * This library is a plain-C (OpenMP) restatement of the arithmetic of the
 * reference solver (main.c of slitvinov/dns).  It exists so that tests,
 * the smoke check and the CPU-baseline leg of bench.py can check the
 * product (the HIP library under dns_amd/) against an independent
 * implementation.  Nothing in the product links, loads or calls it.
 *
 * The FFTs are this project's own (mixed radix 2/3/4 Stockham), because
 * FFTW is not available; their conventions follow FFTW's r2c/c2r
 * (main.c:221-223): forward = exp(-i...), unnormalised, half spectrum on
 * the last (fastest, z) axis, layout [x][y][z/2+1] interleaved re/im.
 *
 * Parity pins (tests/): brute-force DFT, numpy rfftn/irfftn, closed-form
 * single modes, and the reference's own golden trajectories (directories
 * 0128 and 0256 of the reference).
 */
#ifndef DNS_ORACLE_H
#define DNS_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

/* 1-D complex FFT of `howmany` contiguous pencils of length n (interleaved
 * re/im doubles), in place; sign = -1 forward, +1 backward, unnormalised.
 * n must be 2^a 3^b. Returns 0 or -1 on bad n. */
int orc_fft1(int n, double *data, long howmany, int sign);

/* 3-D real-to-complex (forward) transform of a real [n][n][n] array into
 * the half spectrum [n][n][n/2+1] complex (FFTW r2c, main.c:221,249-251). */
int orc_rfft3(int n, const double *in, double *out);

/* 3-D complex-to-real (backward) transform, unnormalised, FFTW c2r
 * semantics (main.c:223): c2c on x and y, then c2r on z that ignores the
 * imaginary parts of the kz=0 and kz=n/2 coefficients. Input preserved. */
int orc_irfft3(int n, const double *in, double *out);

/* Whole-solver restatement of main.c:174-423 (the RK4 loop). */
typedef struct orc_dns orc_dns;
orc_dns *orc_create(int n, double nu, double dt);
void orc_destroy(orc_dns *s);
/* uvw: host [3][n][n][n]; Û = r2c(u), stale physical u = input (main.c:180-185,249-251). */
int orc_load(orc_dns *s, const double *uvw);
/* nsteps RK4 steps (main.c:334-422). */
int orc_step(orc_dns *s, long nsteps);
/* energy / enstrophy as main.c:257-266 (long double accumulation). */
void orc_diag(const orc_dns *s, long double *E, long double *Z);
void orc_diag_d(const orc_dns *s, double *E, double *Z);
/* Copies of internal state for tests: spectral Û,V̂,Ŵ [3][n][n][n/2+1]
 * complex; stale physical velocity [3][n][n][n]; P̂ [n][n][n/2+1]. */
void orc_get_spectral(const orc_dns *s, double *uvw_hat);
void orc_set_spectral(orc_dns *s, const double *uvw_hat);
void orc_get_ustale(const orc_dns *s, double *uvw);
void orc_set_ustale(orc_dns *s, const double *uvw);
void orc_get_phat(const orc_dns *s, double *phat);
/* Snapshot [4][n][n][n] = c2r(U,V,W,P)/n^3 (main.c:269-280). */
void orc_snapshot(const orc_dns *s, double *uvwp);
int orc_n(const orc_dns *s);
/* OpenMP threads the oracle uses (1 without OpenMP). */
int orc_threads(void);

#ifdef __cplusplus
}
#endif
#endif
