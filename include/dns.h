/*
 * libdns_hip: MI355X (gfx950) implementation of the RK4 time step of the
 * pseudo-spectral Navier-Stokes solver `dns` (slitvinov/dns, main.c) for
 * the triply periodic box [0, 2 pi)^3.
 *
 * C ABI: plain pointers and sizes.  Every call returns DNS_OK (0) or a
 * negative DNS_E* code; the library never exits the process.  One context
 * owns all device memory and one HIP stream; calls on a context must be
 * serialised, separate contexts may be used from separate host threads.
 * Host buffers are owned by the caller and only read or written during
 * the call.
 *
 * Which reference interface each entry point replaces (main.c of the
 * reference, which has no library boundary; the boundary it crosses for
 * this path is FFTW's plan/execute API):
 *   dns_create          main.c:174-251  sizing, allocation, fftw_plan_dft_*_3d
 *                                       (221-223), kx/kz/kmax/dealias/kk
 *                                       tables (224-247)
 *   dns_load_velocity   main.c:180-185,249-251  read U,V,W and the initial
 *                                       fftw_execute_dft_r2c of each
 *   dns_step            main.c:334-422  one RK4 step per call unit (33
 *                                       fftw_execute_dft_{r2c,c2r} + the
 *                                       OpenMP loops of the stage)
 *   dns_diagnostics     main.c:257-266  energy / enstrophy reduction
 *   dns_snapshot        main.c:269-280  c2r of U,V,W,P scaled by 1/n^3
 *   dns_rfft3/irfft3    main.c:221-223  fftw_plan_dft_r2c_3d / c2r_3d
 *                                       executed on caller data
 *   dns_destroy         main.c:424-460  fftw_destroy_plan / fftw_free
 */
#ifndef DNS_H
#define DNS_H
#ifdef __cplusplus
extern "C" {
#endif

enum {
  DNS_OK = 0,
  DNS_EINVAL = -1,      /* bad argument (grid size, null pointer, ...) */
  DNS_ENODEV = -2,      /* no usable GPU */
  DNS_ENOMEM = -3,      /* device allocation failed */
  DNS_EHIP = -4,        /* other HIP runtime error */
  DNS_ENONFINITE = -5   /* energy or enstrophy is NaN or Inf */
};

/* Arithmetic of the step.
 * DNS_FP64 : everything in double (the reference's arithmetic).
 * DNS_MIXED: FFTs, curl, u x omega, projection and viscous term in float
 *            with twiddles generated in double; the stage increments are
 *            summed in float (S), and the step-start spectral state U_n,
 *            the new state U_n+1 = (U_n + S) + a_4 dU_4, and the
 *            energy/enstrophy reductions are double.
 * DNS_MIXED_ACC64: as DNS_MIXED, but the running sum S of the stage
 *            increments is formed and stored in double (FP64 RK4
 *            accumulation; 3 more float fields of device memory and 10.5%
 *            more bytes per step at n=512). */
enum dns_precision { DNS_FP64 = 0, DNS_MIXED = 1, DNS_MIXED_ACC64 = 2 };

typedef struct dns_ctx dns_ctx;

/* Grid n x n x n with n even, n = 2^a 3^b and 8 <= n <= 1024 (31 sizes:
 * 8, 12, 16, 18, 24, 32, 36, 48, 54, 64, 72, 96, 108, 128, 144, 162, 192,
 * 216, 256, 288, 324, 384, 432, 486, 512, 576, 648, 768, 864, 972, 1024);
 * other n give DNS_EINVAL.  nu = viscosity (-n), dt = time step (-s).
 * device = HIP device ordinal.  On success *out holds a new context. */
int dns_create(dns_ctx **out, int n, double nu, double dt, int precision, int device);
void dns_destroy(dns_ctx *ctx);

/* Grid size of the context. */
int dns_get_n(const dns_ctx *ctx);

/* uvw: host [3][n][n][n] doubles (U, V, W of the raw input file).  Sets the
 * spectral state to the r2c transform of u (no projection, no dealiasing,
 * as main.c:249-251) and the stale physical velocity used by the first
 * stage of the next step to u itself (main.c:180-185). */
int dns_load_velocity(dns_ctx *ctx, const double *uvw);

/* Advance nsteps RK4 steps (asynchronous on the context's stream). */
int dns_step(dns_ctx *ctx, long nsteps);

/* Ask the last step of every following dns_step call to keep the stage-4
 * pressure P = (k . dU)/|k|^2 for dns_snapshot (main.c:392-395).  Off by
 * default: it costs one extra field write per call. */
int dns_set_pressure_capture(dns_ctx *ctx, int on);

/* Energy sum |U|^2+|V|^2+|W|^2 / n^6 and enstrophy sum |k|^2(...) / n^6
 * over the half spectrum (main.c:257-266), accumulated in double with
 * compensated summation.  Synchronises the stream.  Returns
 * DNS_ENONFINITE (after storing the values) if either is NaN or Inf. */
int dns_diagnostics(dns_ctx *ctx, double *energy, double *enstrophy);

/* Host [4][n][n][n] doubles: c2r(U,V,W,P)/n^3 (main.c:269-280).  P is the
 * captured stage-4 pressure, or zeros if none was captured. */
int dns_snapshot(dns_ctx *ctx, double *uvwp);

/* Asynchronous snapshot, the same values as dns_snapshot (main.c:269-280).
 * dns_snapshot_begin queues the four c2r transforms into a device staging
 * buffer on the context's stream and their copy into a pinned host buffer
 * owned by the context on a second stream, and returns at once: dns_step
 * calls made afterwards run while the copy proceeds.  dns_snapshot_wait
 * blocks until the copy is complete and sets *uvwp to the pinned buffer
 * ([4][n][n][n] doubles), which stays valid until the next
 * dns_snapshot_begin or dns_destroy; it only waits on an event, so another
 * host thread may call it while the owning thread issues dns_step.  The
 * first call allocates the staging and pinned buffers (4 n^3 doubles each).
 * main.c dumps synchronously; these let a caller write the files while the
 * next steps run. */
int dns_snapshot_begin(dns_ctx *ctx);
int dns_snapshot_wait(dns_ctx *ctx, const double **uvwp);

/* State access in the reference layouts, for checkpointing and tests.
 * Spectral: host [3][n][n][n/2+1] complex (interleaved re,im doubles).
 * Stale physical velocity: host [3][n][n][n] doubles. */
int dns_get_spectral(dns_ctx *ctx, double *uvw_hat);
int dns_set_spectral(dns_ctx *ctx, const double *uvw_hat);
int dns_get_stale_velocity(dns_ctx *ctx, double *uvw);
int dns_set_stale_velocity(dns_ctx *ctx, const double *uvw);

/* Stand-alone 3-D transforms on host data with the context's kernels and
 * precision, FFTW conventions (unnormalised; r2c forward exp(-i), c2r
 * backward exp(+i) ignoring Im of the kz=0 and kz=n/2 coefficients).
 * real: [n][n][n] doubles; spec: [n][n][n/2+1] complex doubles. */
int dns_rfft3(dns_ctx *ctx, const double *real, double *spec);
int dns_irfft3(dns_ctx *ctx, const double *spec, double *real);

/* Launch shapes.  A context's kernels can run with several block shapes
 * (strided-pass column tile width, z-pass pencils per block) that all give
 * bitwise identical results; grids up to 512 carry 4 shapes, larger grids 1.
 * dns_autotune times each shape for `steps` steps from the current state,
 * keeps the fastest and restores the state. */
int dns_launch_shapes(dns_ctx *ctx);
int dns_set_launch_shape(dns_ctx *ctx, int shape);
int dns_get_launch_shape(dns_ctx *ctx);
int dns_autotune(dns_ctx *ctx, int steps, int *best_shape);

/* Kernel timing.  When on, every kernel launched by dns_step is bracketed
 * by HIP events on the context's stream; times accumulate per slot
 * 6*stage + pass (stage 0..3; pass 0..5 = x-inverse with curl, y-inverse,
 * fused z pass, y-forward, x-forward with the RK update, and x-forward with
 * the RK update fused with the next stage's x-inverse).  Switching on or off
 * resets the sums.  dns_kernel_timing synchronises and fills
 * ms[DNS_KERNEL_SLOTS] and launches[DNS_KERNEL_SLOTS]; dns_kernel_slot_name
 * gives "<pass>.s<stage 1..4>". */
#define DNS_KERNEL_SLOTS 24
int dns_set_kernel_timing(dns_ctx *ctx, int on);
int dns_kernel_timing(dns_ctx *ctx, double *ms, long *launches);
const char *dns_kernel_slot_name(int slot);

/* Discrete adjoint (FP64 contexts).  J = final enstrophy sum as printed
 * by dns (main.c:263-266) after nsteps RK4 steps from the initial velocity
 * u0 ([3][n][n][n], loaded as dns_load_velocity does).
 * dns_adjoint_forward runs the steps, keeping (spectral state, stale
 * velocity) checkpoints every `every` steps on the device, and returns J.
 * dns_adjoint_backward then returns grad = seed * dJ/du0 by the exact
 * transpose of the discrete step (adjoint FFTs with the half-spectrum
 * weights, adjoint curl, projection, viscous term and u x omega, through the
 * stale-velocity coupling of stage 1), recomputing each segment from its
 * checkpoint.  After the sweep the context holds an unspecified state.
 * dns_adjoint_step applies the adjoint of one step at the state
 * (uvw_hat [3][n][n][n/2+1] complex, uvw_stale [3][n][n][n]) to
 * (lam_hat, lam_stale) in the same layouts.  Pointers may be host or
 * device memory.  No reference counterpart (main.c has no derivative). */
int dns_adjoint_forward(dns_ctx *ctx, const double *u0, long nsteps, long every, double *enstrophy);
int dns_adjoint_backward(dns_ctx *ctx, double seed, double *grad);
int dns_adjoint_step(dns_ctx *ctx, const double *uvw_hat, const double *uvw_stale, const double *lam_hat,
                     const double *lam_stale, double *lam_hat_out, double *lam_stale_out);

/* Batched ensembles: `members` independent simulations, member m on grid
 * n[m] (any mix of supported sizes) with viscosity nu[m] and time step
 * dt[m], all in one precision on one device.  Members of equal n share
 * every kernel launch (one more grid dimension); the sizes run as
 * concurrent branches of one HIP graph captured at the first
 * dns_batch_step and replayed once per step.  Each member follows exactly
 * the arithmetic of a single context (dns_create) with the same n, nu, dt
 * and precision.  No reference counterpart (main.c runs one simulation per
 * process; run.sh:1-10 and hal.sh:1-16 run the Re sweep as separate
 * processes); these replace such sweeps. */
typedef struct dns_batch dns_batch;
int dns_batch_create(dns_batch **out, int members, const int *n, const double *nu, const double *dt,
                     int precision, int device);
void dns_batch_destroy(dns_batch *b);
int dns_batch_load_velocity(dns_batch *b, int member, const double *uvw);
int dns_batch_step(dns_batch *b, long nsteps);
int dns_batch_diagnostics(dns_batch *b, int member, double *energy, double *enstrophy);
int dns_batch_get_spectral(dns_batch *b, int member, double *uvw_hat);
int dns_batch_get_stale_velocity(dns_batch *b, int member, double *uvw);
int dns_batch_member_n(dns_batch *b, int member);
/* Autotune every grid size of the batch (as dns_autotune); the step graph
 * is recaptured with the chosen shapes.  dns_batch_launch_shape reports the
 * shape used for a member's grid size. */
int dns_batch_autotune(dns_batch *b, int steps);
int dns_batch_launch_shape(dns_batch *b, int member);
long long dns_batch_device_bytes(const dns_batch *b);

/* Energy and enstrophy sums of dns_diagnostics (main.c:257-266) of a
 * spectral field in caller-owned device memory, allocating nothing.  No
 * counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops).
 * uvw_hat: device [3][n][n][n/2+1] complex doubles in the layout of
 * dns_get_spectral; out: device double[2] = (energy, enstrophy), bitwise
 * equal to what dns_diagnostics gives for the same state.  Queued on
 * `stream` (a hipStream_t, NULL for the null stream) of the current device;
 * the caller synchronises and checks finiteness.  n even, 2 <= n <= 1024,
 * else DNS_EINVAL before any device work.  The block partial sums live in
 * one fixed 32 KiB device array per GPU, shared by every call: calls that
 * overlap on one GPU race and must be serialised by the caller. */
int dns_diagnostics_noalloc(const double *uvw_hat, int n, void *stream, double *out);

/* Bytes of device memory held by the context. */
long long dns_device_bytes(const dns_ctx *ctx);

/* Static text for an error code. */
const char *dns_strerror(int code);

/* Detail of the last error on this thread (HIP message), or "". */
const char *dns_last_error(void);

#ifdef __cplusplus
}
#endif
#endif
