/*
 * dns-ic: initial conditions in the raw input layout of `dns`,
 * [4][n][n][n] float64 (U, V, W, P), i = x slowest, k = z fastest.
 *
 *   dns-ic [-t tgv|iso] (-l <level> | -N <n>) -o <out.raw>
 *          [-S <seed>] [-k <k0>] [-e <energy>]
 *
 * tgv (default): Taylor-Green vortex as tgv.py:14-27 of the reference; with
 *   dx = 2 pi/n, s_i = sin(dx i), c_i = cos(dx i):
 *     U = s_i c_j c_k,  V = (-c_i) s_j c_k,  W = P = 0.
 * iso: seeded random-phase, divergence-free, isotropic velocity with the
 *   spectrum E(k) ~ k^4 exp(-2 (k/k0)^2) (default k0 = 4) and kinetic energy
 *   <|u|^2>/2 = energy (default 0.5).  Each Fourier mode k of the half
 *   spectrum gets a complex Gaussian vector from a counter-based SplitMix64
 *   stream keyed by (seed, mode, component), projected onto the plane normal
 *   to k and scaled to sqrt(E(|k|)/(4 pi |k|^2)); the mean mode and the
 *   Nyquist planes (i = n/2, j = n/2, k = n/2) are zero.  The real field is
 *   the complex-to-real transform of that half spectrum (on the GPU, through
 *   libdns_hip), rescaled to the requested energy.  The same seed gives the
 *   same file on any machine.  The "energy" line printed by dns is the
 *   half-spectrum sum of main.c:257-266, which counts the kz=0 plane once
 *   and so reads slightly above <|u|^2>/2 for such fields.
 *
 * DNS_DEVICE selects the GPU for iso (default 0).
 */
#include <errno.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "dns.h"

static void usage(void) {
  fprintf(stderr, "Usage: dns-ic [-t tgv|iso] (-l <level> | -N <n>) -o <output.raw>\n"
                  "              [-S <seed>] [-k <k0>] [-e <energy>]\n"
                  "  -t tgv|iso   Taylor-Green vortex (default) or isotropic random field\n"
                  "  -l <level>   grid n = 2^level\n"
                  "  -N <n>       grid n\n"
                  "  -o <path>    output raw file [4][n][n][n] float64\n"
                  "  -S <seed>    iso: random seed (default 20261202)\n"
                  "  -k <k0>      iso: spectrum peak parameter (default 4)\n"
                  "  -e <energy>  iso: kinetic energy <|u|^2>/2 (default 0.5)\n");
  exit(1);
}

/* SplitMix64 finaliser (Steele, Lea, Flood 2014) of a counter. */
static uint64_t mix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ULL;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}
static double uniform01(uint64_t seed, uint64_t ctr) { /* (0, 1] */
  uint64_t r = mix64(mix64(seed) ^ ctr);
  return ((double)(r >> 11) + 1.0) * (1.0 / 9007199254740992.0);
}
/* Standard normal by Box-Muller from two uniforms of counter 2c, 2c+1. */
static double gauss(uint64_t seed, uint64_t c) {
  const double two_pi = 6.283185307179586476925286766559;
  double a = uniform01(seed, 2 * c), b = uniform01(seed, 2 * c + 1);
  return sqrt(-2.0 * log(a)) * cos(two_pi * b);
}

static int write_raw(const char *out, const double *uvw, long n) {
  FILE *f = fopen(out, "wb");
  if (f == NULL) {
    fprintf(stderr, "dns-ic: error: cannot open '%s': %s\n", out, strerror(errno));
    return 1;
  }
  long n3 = n * n * n;
  double *zero = calloc(n * n, sizeof(double));
  int bad = fwrite(uvw, sizeof(double), 3 * n3, f) != (size_t)(3 * n3);
  for (long r = 0; r < n && !bad; r++)
    bad = fwrite(zero, sizeof(double), n * n, f) != (size_t)(n * n);
  free(zero);
  if (fclose(f) != 0 || bad) {
    fprintf(stderr, "dns-ic: error: write failed\n");
    return 1;
  }
  return 0;
}

static int tgv(long n, const char *out) {
  const double L = 2 * 3.141592653589793, dx = L / n;
  long n3 = n * n * n;
  double *s = malloc(sizeof(double) * n), *c = malloc(sizeof(double) * n);
  double *uvw = malloc(sizeof(double) * 3 * n3);
  if (!s || !c || !uvw) {
    fprintf(stderr, "dns-ic: error: out of memory\n");
    return 1;
  }
  for (long i = 0; i < n; i++) {
    s[i] = sin(dx * i);
    c[i] = cos(dx * i);
  }
  for (long i = 0; i < n; i++)
    for (long j = 0; j < n; j++)
      for (long k = 0; k < n; k++) {
        long l = (i * n + j) * n + k;
        uvw[l] = s[i] * c[j] * c[k];
        uvw[n3 + l] = -c[i] * s[j] * c[k];
        uvw[2 * n3 + l] = 0.0;
      }
  int rc = write_raw(out, uvw, n);
  free(s);
  free(c);
  free(uvw);
  return rc;
}

static int iso(long n, const char *out, uint64_t seed, double k0, double energy) {
  const double pi = 3.141592653589793238;
  long nf = n / 2 + 1, n3 = n * n * n, n3f = n * n * nf;
  double *spec = malloc(sizeof(double) * 2 * n3f);
  double *uvw = malloc(sizeof(double) * 3 * n3);
  if (!spec || !uvw) {
    fprintf(stderr, "dns-ic: error: out of memory\n");
    return 1;
  }
  int device = getenv("DNS_DEVICE") ? atoi(getenv("DNS_DEVICE")) : 0;
  dns_ctx *ctx = NULL;
  int rc = dns_create(&ctx, (int)n, 0.0, 0.0, DNS_FP64, device);
  if (rc != DNS_OK) {
    fprintf(stderr, "dns-ic: error: %s (%s)\n", dns_strerror(rc), dns_last_error());
    return 1;
  }
  /* one complex Gaussian 3-vector per mode, projected normal to k */
  double *hat = malloc(sizeof(double) * 6 * n3f);
  for (long i = 0; i < n; i++)
    for (long j = 0; j < n; j++)
      for (long k = 0; k < nf; k++) {
        long l = (i * n + j) * nf + k;
        double *h = hat + 6 * l;
        double kx = i < n / 2 ? i : i - n, ky = j < n / 2 ? j : j - n, kz = k;
        double kk = kx * kx + ky * ky + kz * kz;
        if (kk == 0 || i == n / 2 || j == n / 2 || k == n / 2) {
          memset(h, 0, 6 * sizeof(double));
          continue;
        }
        double g[6], kv[3] = {kx, ky, kz};
        for (int q = 0; q < 6; q++)
          g[q] = gauss(seed, (uint64_t)l * 8 + q);
        for (int part = 0; part < 2; part++) { /* real and imaginary parts */
          double dot = (g[part] * kx + g[2 + part] * ky + g[4 + part] * kz) / kk;
          for (int a = 0; a < 3; a++)
            g[2 * a + part] -= dot * kv[a];
        }
        double norm = 0;
        for (int q = 0; q < 6; q++)
          norm += g[q] * g[q];
        double kmag = sqrt(kk);
        double Ek = kk * kk * exp(-2.0 * (kmag / k0) * (kmag / k0));
        double amp = norm > 0 ? sqrt(Ek / (4 * pi * kk) / norm) : 0.0;
        for (int q = 0; q < 6; q++)
          h[q] = amp * g[q];
      }
  for (int a = 0; a < 3 && rc == DNS_OK; a++) {
    for (long l = 0; l < n3f; l++) {
      spec[2 * l] = hat[6 * l + 2 * a];
      spec[2 * l + 1] = hat[6 * l + 2 * a + 1];
    }
    rc = dns_irfft3(ctx, spec, uvw + a * n3);
  }
  dns_destroy(ctx);
  free(hat);
  if (rc != DNS_OK) {
    fprintf(stderr, "dns-ic: error: %s (%s)\n", dns_strerror(rc), dns_last_error());
    return 1;
  }
  long double e = 0;
  for (long l = 0; l < 3 * n3; l++)
    e += (long double)uvw[l] * uvw[l];
  double ke = (double)(e / (2.0L * n3));
  double scale = ke > 0 ? sqrt(energy / ke) : 0.0;
  for (long l = 0; l < 3 * n3; l++)
    uvw[l] *= scale;
  rc = write_raw(out, uvw, n);
  free(spec);
  free(uvw);
  return rc;
}

int main(int argc, char **argv) {
  long n = 0;
  const char *out = NULL, *type = "tgv";
  uint64_t seed = 20261202;
  double k0 = 4.0, energy = 0.5;
  for (int a = 1; a < argc; a++) {
    if (strcmp(argv[a], "-l") == 0 && a + 1 < argc)
      n = 1L << atoi(argv[++a]);
    else if (strcmp(argv[a], "-N") == 0 && a + 1 < argc)
      n = atol(argv[++a]);
    else if (strcmp(argv[a], "-o") == 0 && a + 1 < argc)
      out = argv[++a];
    else if (strcmp(argv[a], "-t") == 0 && a + 1 < argc)
      type = argv[++a];
    else if (strcmp(argv[a], "-S") == 0 && a + 1 < argc)
      seed = strtoull(argv[++a], NULL, 10);
    else if (strcmp(argv[a], "-k") == 0 && a + 1 < argc)
      k0 = atof(argv[++a]);
    else if (strcmp(argv[a], "-e") == 0 && a + 1 < argc)
      energy = atof(argv[++a]);
    else
      usage();
  }
  if (n < 2 || out == NULL)
    usage();
  fprintf(stderr, "dns-ic: n=%ld\n", n);
  if (strcmp(type, "tgv") == 0)
    return tgv(n, out);
  if (strcmp(type, "iso") == 0)
    return iso(n, out, seed, k0, energy);
  usage();
  return 1;
}
