/*
 * dns: drop-in command line of the reference solver (main.c:34-423) over
 * libdns_hip.  Same flags, raw input layout, stdout lines, snapshot files
 * and exit codes; the arithmetic runs on the GPU through include/dns.h.
 *
 *   dns [-v] [-d] -i <input.raw> -n <viscosity> -t <end time> -s <time step>
 *
 * Environment (opt-in, defaults keep the reference behaviour):
 *   DNS_PRECISION=fp64|mixed|mixed_acc64   arithmetic of the step (default fp64)
 *   DNS_DEVICE=<ordinal>       HIP device (default 0)
 *
 * Differences that cannot be avoided: under -v the reference prints
 * "dns: omp_get_max_threads: N"; this prints "dns: device: <ordinal>",
 * and at the end "dns: <steps> steps in <s> s (<steps/s> steps/s)" (wall
 * time of the step loop, diagnostics and dumps included).
 * The reference dies with SIGFPE at the first overflow/invalid operation
 * (main.c:49-50); here the diagnostics are checked every 10 steps and a
 * non-finite energy or enstrophy raises SIGFPE before its line is printed.
 *
 * Snapshots (-d) are written while the next steps run: the transforms and
 * the device-to-host copy are queued (dns_snapshot_begin), a writer thread
 * waits for the copy and writes the raw file (in parallel chunks) and the
 * XDMF file, and the main thread joins it before the next diagnostics line
 * is printed.  Files, stdout, error messages and exit codes are those of
 * the synchronous dump of main.c:269-330.
 */
#define _GNU_SOURCE
#include <fcntl.h>
#include <math.h>
#include <pthread.h>
#include <signal.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>
#include "dns.h"

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
  exit(1);
}

static double number(char **argv, const char *flag) { /* main.c:95-129 */
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

static void fail(int rc) {
  const char *detail = dns_last_error();
  if (detail && *detail)
    fprintf(stderr, "dns: error: %s (%s)\n", dns_strerror(rc), detail);
  else
    fprintf(stderr, "dns: error: %s\n", dns_strerror(rc));
  exit(1);
}

static double wall_seconds(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static void nonfinite(void) {
  fflush(stdout);
  signal(SIGFPE, SIG_DFL);
  raise(SIGFPE);
  exit(128 + SIGFPE);
}

/* main.c:281-328, including its quirks: the file is numbered by idump,
 * it references "<idump>.raw", and the geometry type is spelt as there. */
static int write_xdmf(long idump, double t, long n, double dx, char *path) {
  static const char *names[nvars] = {"U", "V", "W", "P"};
  sprintf(path, "a.%08ld.xdmf2", idump);
  FILE *file = fopen(path, "w");
  if (file == NULL)
    return -1;
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
  return 0;
}

/* One chunk of a raw file written with pwrite (like fwrite in main.c:275,
 * short writes are not reported). */
struct chunk {
  int fd;
  const char *p;
  size_t len;
  off_t off;
};

static void *write_chunk(void *arg) {
  struct chunk *c = arg;
  while (c->len > 0) {
    ssize_t w = pwrite(c->fd, c->p, c->len, c->off);
    if (w <= 0)
      break;
    c->p += w;
    c->len -= (size_t)w;
    c->off += w;
  }
  return NULL;
}

/* A pending snapshot: written by dump_writer, checked by dump_finish. */
struct dump_job {
  dns_ctx *ctx;
  long tstep, idump, n;
  double t, dx;
  int rc;                    /* library error of dns_snapshot_wait */
  int open_failed;           /* path could not be opened */
  char path[FILENAME_MAX];
};

enum { max_writers = 8 };

static void *dump_writer(void *arg) {
  struct dump_job *j = arg;
  const double *buf;
  if ((j->rc = dns_snapshot_wait(j->ctx, &buf)) != DNS_OK)
    return NULL;
  sprintf(j->path, "%08ld.raw", j->tstep);
  int fd = open(j->path, O_WRONLY | O_CREAT | O_TRUNC, 0666);
  if (fd < 0) {
    j->open_failed = 1;
    return NULL;
  }
  const size_t bytes = (size_t)nvars * j->n * j->n * j->n * sizeof(double);
  int nw = (int)(bytes >> 26) + 1; /* one writer per 64 MiB, at most max_writers */
  if (nw > max_writers)
    nw = max_writers;
  struct chunk c[max_writers];
  pthread_t th[max_writers];
  const size_t per = (bytes / nw + 4095) & ~(size_t)4095;
  int started = 0;
  for (int i = 0; i < nw; i++) {
    size_t off = per * i;
    c[i].fd = fd;
    c[i].p = (const char *)buf + (off < bytes ? off : bytes);
    c[i].len = off < bytes ? (bytes - off < per ? bytes - off : per) : 0;
    c[i].off = (off_t)(off < bytes ? off : bytes);
    if (i > 0 && pthread_create(&th[i], NULL, write_chunk, &c[i]) == 0)
      started |= 1 << i;
    else if (i > 0)
      write_chunk(&c[i]);
  }
  write_chunk(&c[0]);
  for (int i = 1; i < nw; i++)
    if (started & (1 << i))
      pthread_join(th[i], NULL);
  close(fd);
  if (write_xdmf(j->idump, j->t, j->n, j->dx, j->path) != 0)
    j->open_failed = 1;
  return NULL;
}

/* Wait for the pending snapshot, if any; report its failure as main.c does. */
static void dump_finish(pthread_t *th, int *live, struct dump_job *j) {
  if (!*live)
    return;
  pthread_join(*th, NULL);
  *live = 0;
  if (j->rc != DNS_OK)
    fail(j->rc);
  if (j->open_failed) {
    fprintf(stderr, "dns: error: fail to open '%s'\n", j->path);
    exit(1);
  }
}

int main(int argc, char **argv) {
  (void)argc;
  char *input_path = NULL;
  double dt = -1, T = 0, nu = -1;
  int Verbose = 0, Dump = 0, rc;
  FILE *file;
  while (*++argv != NULL && argv[0][0] == '-') { /* main.c:58-134 */
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
  int precision = DNS_FP64, device = 0;
  const char *env = getenv("DNS_PRECISION");
  if (env && strcmp(env, "mixed") == 0)
    precision = DNS_MIXED;
  else if (env && strcmp(env, "mixed_acc64") == 0)
    precision = DNS_MIXED_ACC64;
  else if (env && *env && strcmp(env, "fp64") != 0) {
    fprintf(stderr, "dns: error: DNS_PRECISION must be fp64, mixed or mixed_acc64\n");
    exit(1);
  }
  if ((env = getenv("DNS_DEVICE")) != NULL && *env)
    device = atoi(env);
  fseek(file, 0, SEEK_END); /* main.c:163-173 */
  size_t offset = ftell(file);
  rewind(file);
  long n = offset / sizeof(double) / nvars;
  n = round(powf(n, 1.0 / 3));
  if (n * n * n * nvars * sizeof(double) != offset) {
    fprintf(stderr, "dns: error: wrong file '%s'\n", input_path);
    exit(1);
  }
  dns_ctx *ctx = NULL;
  if ((rc = dns_create(&ctx, (int)n, nu, dt, precision, device)) != DNS_OK)
    fail(rc);
  if (Verbose) {
    fprintf(stderr, "dns: device: %d\n", device);
    fprintf(stderr, "dns: n = %ld\n", n);
  }
  long n3 = n * n * n;
  double *buf = malloc(sizeof(double) * 3 * n3);
  if (buf == NULL) {
    fprintf(stderr, "dns: error: out of host memory\n");
    exit(1);
  }
  if (fread(buf, sizeof(double), n3, file) != (size_t)n3 || /* main.c:180-185 */
      fread(buf + n3, sizeof(double), n3, file) != (size_t)n3 ||
      fread(buf + 2 * n3, sizeof(double), n3, file) != (size_t)n3 || fclose(file) != 0) {
    fprintf(stderr, "dns: error: fail to read '%s'\n", input_path);
    exit(1);
  }
  if ((rc = dns_load_velocity(ctx, buf)) != DNS_OK)
    fail(rc);
  free(buf);
  if (Dump && (rc = dns_set_pressure_capture(ctx, 1)) != DNS_OK)
    fail(rc);
  double dx = 2 * pi / n, t = 0.0;
  long tstep = 0, idump = 0, pending = 0;
  pthread_t writer;
  int writer_live = 0;
  struct dump_job job;
  const double t_start = wall_seconds();
  for (;;) { /* main.c:256-423; steps are queued and run before each line */
    if (tstep % 10 == 0) {
      double energy, Omega;
      if (pending && (rc = dns_step(ctx, pending)) != DNS_OK)
        fail(rc);
      pending = 0;
      rc = dns_diagnostics(ctx, &energy, &Omega);
      /* the previous snapshot is complete before this line, as in main.c */
      dump_finish(&writer, &writer_live, &job);
      if (rc == DNS_ENONFINITE)
        nonfinite();
      if (rc != DNS_OK)
        fail(rc);
      printf("% 10ld % .16e % .16Le % .16Le\n", tstep, t, (long double)energy, (long double)Omega);
      fflush(stdout);
      if (Dump) { /* main.c:269-330, written while the next steps run */
        if ((rc = dns_snapshot_begin(ctx)) != DNS_OK)
          fail(rc);
        memset(&job, 0, sizeof job);
        job.ctx = ctx;
        job.tstep = tstep;
        job.idump = idump;
        job.n = n;
        job.t = t;
        job.dx = dx;
        if (pthread_create(&writer, NULL, dump_writer, &job) != 0) {
          fprintf(stderr, "dns: error: cannot start the snapshot writer\n");
          exit(1);
        }
        writer_live = 1;
        idump++;
      }
    }
    if (t > T)
      break;
    pending++;
    t += dt;
    tstep++;
  }
  dump_finish(&writer, &writer_live, &job);
  if (pending) { /* the reference runs these steps too before exiting */
    double e, z;
    if ((rc = dns_step(ctx, pending)) != DNS_OK)
      fail(rc);
    rc = dns_diagnostics(ctx, &e, &z);
    if (rc == DNS_ENONFINITE)
      nonfinite();
  }
  if (Verbose) {
    const double el = wall_seconds() - t_start;
    fprintf(stderr, "dns: %ld steps in %.3f s (%.2f steps/s)\n", tstep, el, el > 0 ? tstep / el : 0.0);
  }
  dns_destroy(ctx);
  return 0;
}
