/*
 * dns_cpu: the CPU oracle behind the reference's command line -- TEST
 * INFRASTRUCTURE ONLY.  Restates main.c:34-423 (argument parsing,
 * raw-file sizing and reading, the diagnostics/dump/step loop and its
 * stdout format) on top of the oracle library (orc_solver.c).
 */
#define _GNU_SOURCE
#include <fenv.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "oracle.h"

enum { nvars = 4 };
static const double pi = 3.141592653589793238;

static void usage(void) { /* main.c:61-79 */
  fprintf(stderr, "Usage: dns [-v] [-d] -i <input.raw> -n <viscosity> -t "
                  "<end time> -s <time step>\n"
                  "\n"
                  "Options:\n"
                  "  -i <input.raw>    Input file\n"
                  "  -n <viscosity>    Viscosity\n"
                  "  -t <end time>     End time\n"
                  "  -s <time step>    Time step\n"
                  "  -v                Verbose output\n"
                  "  -d                Dump snapshots\n"
                  "  -h                Show this help message\n"
                  "\n"
                  "Example:\n"
                  "  dns -i tgv.raw -n 0.01 -t 1.0 -s 0.001 -v\n");
#ifdef _OPENMP
  fprintf(stderr, "\nBuild Info:\n"
                  "  OpenMP is enabled.\n");
#endif
  exit(1);
}

static double number(char **argv, const char *flag) {
  char *end;
  if (*argv == NULL) {
    fprintf(stderr, "dns: error: %s needs an argument\n", flag);
    exit(1);
  }
  double v = strtod(*argv, &end);
  if (*end != '\0') {
    fprintf(stderr, "dns: error: '%s' is not a number\n", *argv);
    exit(1);
  }
  return v;
}

static void write_xdmf(long idump, double t, long n, double dx) { /* main.c:281-328 */
  char path[FILENAME_MAX];
  static const char *names[nvars] = {"U", "V", "W", "P"};
  sprintf(path, "a.%08ld.xdmf2", idump);
  FILE *file = fopen(path, "w");
  fprintf(file,
          "<Xdmf\n"
          "    Version=\"2\">\n"
          "  <Domain>\n"
          "    <Grid>\n"
          "      <Time\n"
          "          Value=\"%+.16e\"/>\n"
          "      <Topology\n"
          "          TopologyType=\"3DCoRectMesh\"\n"
          "          Dimensions=\"%ld %ld %ld\"/>\n"
          "      <Geometry\n"
          "          GeometryType=\"ORIGIn_DXDYDZ\">\n"
          "        <DataItem\n"
          "            Dimensions=\"3\">\n"
          "          0\n"
          "          0\n"
          "          0\n"
          "        </DataItem>\n"
          "        <DataItem\n"
          "            Dimensions=\"3\">\n"
          "          %.16e\n"
          "          %.16e\n"
          "          %.16e\n"
          "        </DataItem>\n"
          "      </Geometry>\n",
          t, n, n, n, dx, dx, dx);
  long offset = 0;
  for (int v = 0; v < nvars; v++) {
    fprintf(file,
            "      <Attribute\n"
            "          name=\"%s\">\n"
            "        <DataItem\n"
            "            Format=\"Binary\"\n"
            "            Seek=\"%ld\"\n"
            "            Precision=\"8\"\n"
            "            Dimensions=\"%ld %ld %ld\">\n"
            "          %08ld.raw\n"
            "        </DataItem>\n"
            "      </Attribute>\n",
            names[v], offset, n, n, n, idump);
    offset += n * n * n * (long)sizeof(double);
  }
  fprintf(file, "    </Grid>\n"
                "  </Domain>\n"
                "</Xdmf>\n");
  fclose(file);
}

int main(int argc, char **argv) {
  (void)argc;
  char *input_path = NULL;
  double dt = -1, T = 0, nu = -1;
  int Verbose = 0, Dump = 0;
  FILE *file;
  feclearexcept(FE_ALL_EXCEPT);
  feenableexcept(FE_DIVBYZERO | FE_INVALID | FE_OVERFLOW); /* main.c:49-50 */
  while (*++argv != NULL && argv[0][0] == '-') {            /* main.c:58-134 */
    switch (argv[0][1]) {
    case 'h':
      usage();
      break;
    case 'v':
      Verbose = 1;
      break;
    case 'd':
      Dump = 1;
      break;
    case 'i':
      argv++;
      if (*argv == NULL) {
        fprintf(stderr, "dns: error: -i needs an argument\n");
        exit(1);
      }
      input_path = *argv;
      break;
    case 'n':
      argv++;
      nu = number(argv, "-n");
      break;
    case 's':
      argv++;
      dt = number(argv, "-s");
      break;
    case 't':
      argv++;
      T = number(argv, "-t");
      break;
    default:
      fprintf(stderr, "dns: error: unknown option '%s'\n", *argv);
      exit(1);
    }
  }
  if (T == 0) { /* main.c:135-150 */
    fprintf(stderr, "dns: error: -t is not set or invalid\n");
    exit(1);
  }
  if (nu == -1) {
    fprintf(stderr, "dns: error: -n is not set or invalid\n");
    exit(1);
  }
  if (dt == -1) {
    fprintf(stderr, "dns: error: -s is not set or invalid\n");
    exit(1);
  }
  if (input_path == NULL) {
    fprintf(stderr, "dns: error: -i is not set\n");
    exit(1);
  }
  if ((file = fopen(input_path, "r")) == NULL) {
    fprintf(stderr, "dns: error: fail to open '%s'\n", input_path);
    exit(1);
  }
#ifdef _OPENMP
  if (Verbose)
    fprintf(stderr, "dns: omp_get_max_threads: %d\n", omp_get_max_threads());
#endif
  fseek(file, 0, SEEK_END); /* main.c:163-173 */
  size_t offset = ftell(file);
  rewind(file);
  long n = offset / sizeof(double) / nvars;
  n = round(powf(n, 1.0 / 3));
  if (n * n * n * nvars * sizeof(double) != offset) {
    fprintf(stderr, "dns: error: wrong file '%s'\n", input_path);
    exit(1);
  }
  if (Verbose)
    fprintf(stderr, "dns: n = %ld\n", n);
  long n3 = n * n * n;
  double *uvw = malloc(sizeof(double) * 3 * n3);
  if (fread(uvw, sizeof(double), n3, file) != (size_t)n3 ||
      fread(uvw + n3, sizeof(double), n3, file) != (size_t)n3 ||
      fread(uvw + 2 * n3, sizeof(double), n3, file) != (size_t)n3 || fclose(file) != 0) {
    fprintf(stderr, "dns: error: fail to read '%s'\n", input_path);
    exit(1);
  }
  orc_dns *s = orc_create((int)n, nu, dt);
  orc_load(s, uvw);
  free(uvw);
  double *dump = Dump ? malloc(sizeof(double) * nvars * n3) : NULL;
  double dx = 2 * pi / n, t = 0.0;
  long tstep = 0, idump = 0;
  for (;;) { /* main.c:256-423 */
    if (tstep % 10 == 0) {
      long double energy, Omega;
      orc_diag(s, &energy, &Omega);
      printf("% 10ld % .16e % .16Le % .16Le\n", tstep, t, energy, Omega);
      fflush(stdout);
      if (Dump) {
        char path[FILENAME_MAX];
        sprintf(path, "%08ld.raw", tstep);
        file = fopen(path, "w");
        orc_snapshot(s, dump);
        fwrite(dump, n3 * nvars, sizeof(double), file);
        fclose(file);
        write_xdmf(idump, t, n, dx);
        idump++;
      }
    }
    if (t > T)
      break;
    orc_step(s, 1);
    t += dt;
    tstep++;
  }
  free(dump);
  orc_destroy(s);
  return 0;
}
