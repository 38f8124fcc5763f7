/*
 * dns-ic: initial conditions in the raw input layout of `dns`,
 * [4][n][n][n] float64 (U, V, W, P), i = x slowest, k = z fastest.
 *
 *   dns-ic -l <level> -o <out.raw>          Taylor-Green vortex, n = 2^level
 *   dns-ic -N <n> -o <out.raw>              Taylor-Green vortex, any n
 *
 * Taylor-Green follows tgv.py:14-27 of the reference: with dx = 2 pi/n,
 * s_i = sin(dx i), c_i = cos(dx i):
 *   U = s_i c_j c_k,  V = (-c_i) s_j c_k,  W = P = 0.
 */
#include <errno.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static void usage(void) {
  fprintf(stderr, "Usage: dns-ic (-l <level> | -N <n>) -o <output.raw>\n"
                  "  -l <level>   grid n = 2^level\n"
                  "  -N <n>       grid n\n"
                  "  -o <path>    output raw file [4][n][n][n] float64\n");
  exit(1);
}

int main(int argc, char **argv) {
  long n = 0;
  const char *out = NULL;
  for (int a = 1; a < argc; a++) {
    if (strcmp(argv[a], "-l") == 0 && a + 1 < argc)
      n = 1L << atoi(argv[++a]);
    else if (strcmp(argv[a], "-N") == 0 && a + 1 < argc)
      n = atol(argv[++a]);
    else if (strcmp(argv[a], "-o") == 0 && a + 1 < argc)
      out = argv[++a];
    else
      usage();
  }
  if (n < 2 || out == NULL)
    usage();
  fprintf(stderr, "dns-ic: n=%ld\n", n);
  const double L = 2 * 3.141592653589793, dx = L / n;
  double *s = malloc(sizeof(double) * n), *c = malloc(sizeof(double) * n);
  double *row = malloc(sizeof(double) * n);
  for (long i = 0; i < n; i++) {
    s[i] = sin(dx * i);
    c[i] = cos(dx * i);
  }
  FILE *f = fopen(out, "wb");
  if (f == NULL) {
    fprintf(stderr, "dns-ic: error: cannot open '%s': %s\n", out, strerror(errno));
    return 1;
  }
  for (int var = 0; var < 4; var++)
    for (long i = 0; i < n; i++)
      for (long j = 0; j < n; j++) {
        for (long k = 0; k < n; k++) {
          if (var == 0)
            row[k] = s[i] * c[j] * c[k];
          else if (var == 1)
            row[k] = -c[i] * s[j] * c[k];
          else
            row[k] = 0.0;
        }
        if (fwrite(row, sizeof(double), n, f) != (size_t)n) {
          fprintf(stderr, "dns-ic: error: write failed\n");
          return 1;
        }
      }
  fclose(f);
  free(s);
  free(c);
  free(row);
  return 0;
}
