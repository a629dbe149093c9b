/*
 * npbx.h -- C ABI of libnpbx.so, the MI355X (gfx950) structured-grid stencil
 * library behind the npbench "mi355x" framework entry.
 *
 * Conventions
 *  - Every entry point returns int: NPBX_OK (0) or a negative npbx_status.
 *    npbx_last_error() copies the calling thread's last message.
 *  - Array arguments are device pointers (HBM) to C-contiguous row-major
 *    arrays, last axis fastest, exactly the NumPy layouts of the reference
 *    kernels.  Sizes are element counts.  `stream` is a hipStream_t (NULL =
 *    the default stream); kernels are enqueued and the call returns without
 *    waiting; use npbx_sync() to wait with a deadline.
 *  - Library scratch: iterated hdiff (nsteps > 1), vadv outside its on-chip
 *    shapes (K > 160 or columns not whole 16-byte vectors), ADI, cavity and
 *    channel flow, and single tiled FP32 heat grids use one grow-only
 *    device workspace per GPU.  Calls that use it are not ordered against
 *    each other by the library: enqueue them on one stream (or wait between
 *    them); two such calls in flight on different streams corrupt each
 *    other's results.  Growing the workspace synchronises the device, so a
 *    call that grows it must not run under stream capture (run it once
 *    outside the capture first).  npbx_adi_noalloc never touches it.
 *  - dtype selects the storage type: NPBX_F64 (the drop-in, FP64 arithmetic,
 *    bitwise equal to the reference NumPy code), NPBX_F32 and NPBX_BF16
 *    (reduced-precision storage, FP32 arithmetic).  vadv always eliminates
 *    in FP64, whatever the field storage type.
 *
 * Reference interfaces replaced (paths relative to the npbench source tree):
 *  - npbx_jacobi2d  <- npbench/benchmarks/polybench/jacobi_2d/jacobi_2d_numpy.py:4
 *                      kernel(TSTEPS, A, B)
 *  - npbx_heat3d    <- npbench/benchmarks/polybench/heat_3d/heat_3d_numpy.py:4
 *                      kernel(TSTEPS, A, B)
 *  - npbx_hdiff     <- npbench/benchmarks/weather_stencils/hdiff/hdiff_numpy.py:5
 *                      hdiff(in_field, out_field, coeff)
 *  - npbx_vadv      <- npbench/benchmarks/weather_stencils/vadv/vadv_numpy.py:9
 *                      vadv(utens_stage, u_stage, wcon, u_pos, utens, dtr_stage)
 *  - copy / sync / timing entry points back the framework adapter's
 *    copy_func / copy_back_func / exec_str synchronisation
 *    (npbench/infrastructure/framework.py:42-50,140-162,
 *     npbench/infrastructure/cupy_framework.py:26-58).
 *  - The batch entry points have no reference counterpart: they run many
 *    independent grids of uneven size in one launch for the batched and
 *    serving forms of the same kernels.
 */
#ifndef NPBX_H
#define NPBX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum npbx_status {
  NPBX_OK = 0,
  NPBX_EINVAL = -1,       /* bad argument (shape, pointer, dtype, aliasing) */
  NPBX_EHIP = -2,         /* a HIP runtime call failed */
  NPBX_ETIMEDOUT = -3,    /* npbx_sync deadline passed; work still queued */
  NPBX_ENOMEM = -4,       /* device allocation failed */
  NPBX_EUNSUPPORTED = -5  /* valid request the library does not implement */
} npbx_status;

typedef enum npbx_dtype {
  NPBX_F64 = 0,
  NPBX_F32 = 1,
  NPBX_BF16 = 2
} npbx_dtype;

/* ---- library / device ---------------------------------------------------- */
int npbx_version(char *buf, size_t len);
int npbx_last_error(char *buf, size_t len);
int npbx_device_count(int *count);
/* Selects the device for this thread and checks it is a gfx950 part. */
int npbx_init(int device);
/* Releases the library-owned scratch of the current device (waits for the
 * device first); later calls allocate it again.  Replaces nothing in the
 * reference (host-side teardown). */
int npbx_finalize(void);

/* ---- memory, streams, events (framework.py:42-50 copy_func/copy_back_func) */
int npbx_malloc(void **ptr, size_t bytes);
int npbx_free(void *ptr);
int npbx_memcpy_h2d(void *dst, const void *src, size_t bytes, void *stream);
int npbx_memcpy_d2h(void *dst, const void *src, size_t bytes, void *stream);
int npbx_stream_create(void **stream);
int npbx_stream_destroy(void *stream);
/* Waits for `stream` until it drains or the monotonic clock passes
 * deadline_ns (CLOCK_MONOTONIC nanoseconds; 0 = no deadline).  Polls, so a
 * host-side timeout can interrupt between calls (timeout_decorator.py:46-72). */
int npbx_sync(void *stream, int64_t deadline_ns);
int npbx_event_create(void **event);
int npbx_event_destroy(void *event);
int npbx_event_record(void *event, void *stream);
int npbx_event_elapsed_ms(void *start, void *stop, float *ms);

/* ---- storage conversion ---------------------------------------------------
 * Element-wise dst = convert(src), round-to-nearest-even (FP64->BF16 is
 * rounded once, not twice).  No reference counterpart: used by the FP32 and
 * BF16 storage forms of the drop-in kernels. */
int npbx_cast(int src_dtype, const void *src, int dst_dtype, void *dst,
              int64_t count, void *stream);

/* Streaming device copy of `bytes` (a multiple of 16, 16-byte aligned
 * pointers).  No reference counterpart: calibrates the achievable HBM rate
 * and the HBM traffic counters that the stencil roofline figures use. */
int npbx_stream_copy(void *dst, const void *src, int64_t bytes, void *stream);

/* Bandwidth probes over `bytes` (a multiple of 16, 16-byte aligned
 * pointers).  No reference counterpart: they measure the device's practical
 * HBM ceiling.  mode 0 copies src to dst, 1 reads src (dst: at least 16
 * bytes, never written in practice), 2 writes dst (src unused).  variant
 * picks the launch shape: one pass, a workgroup of 256 threads per 256 x U
 * vectors, U = 4 (0), 8 (1), and with non-temporal loads and stores U = 4
 * (2), 8 (3), 16 (4). */
int npbx_bandwidth(int mode, int variant, void *dst, const void *src,
                   int64_t bytes, void *stream);

/* ---- 2-D 5-point Jacobi (jacobi_2d_numpy.py:4-10) -------------------------
 * A, B: N x N.  Runs TSTEPS-1 ping-pong iterations (2*(TSTEPS-1) sweeps,
 * A->B then B->A), interior only, boundaries never written.
 * tblock: sweeps fused per HBM round trip (0 or 1 = one launch per sweep;
 * 2..4 FP32, 2 FP64/BF16 = temporal-blocked tiles, bitwise the same result;
 * needs N a multiple of the elements per 16 bytes).
 */
int npbx_jacobi2d(int dtype, void *A, void *B, int64_t N, int64_t tsteps,
                  int tblock, void *stream);
/* Same sweep, counted in sweeps: sweep s (s = first..first+nsweeps-1) reads
 * A and writes B when s is even, and reads B and writes A when s is odd. */
int npbx_jacobi2d_sweeps(int dtype, void *A, void *B, int64_t N,
                         int64_t first, int64_t nsweeps, int tblock,
                         void *stream);

/* ---- 3-D 7-point heat (heat_3d_numpy.py:4-20) ----------------------------- */
int npbx_heat3d(int dtype, void *A, void *B, int64_t N, int64_t tsteps,
                int tblock, void *stream);
int npbx_heat3d_sweeps(int dtype, void *A, void *B, int64_t N, int64_t first,
                       int64_t nsweeps, int tblock, void *stream);

/* ---- batched grids: many independent grids in persistent launches --------
 * No reference counterpart: the batched and serving forms of the two
 * explicit kernels.  A plan fixes a list of grids (each with its own side,
 * arrays and sweep range, all of one dtype) and a temporal-blocking depth
 * `tblock` (2..4 FP32, 2 FP64/BF16); every launch of the plan runs one round
 * of up to `tblock` sweeps for every grid that still has one, as one
 * persistent kernel over a device work-tile queue.  Results are bitwise
 * those of the per-grid entry points.  `seg` = heat planes per tile (0 =
 * default).  Each grid's side must be a multiple of 16 bytes of elements.
 * npbx_plan_capture records the plan's launches into a HIP graph on a
 * non-default stream; later npbx_plan_launch calls replay the graph. */
typedef enum npbx_kind {
  NPBX_KIND_JACOBI2D = 0,
  NPBX_KIND_HEAT3D = 1
} npbx_kind;

typedef struct npbx_grid_spec {
  int32_t kind;    /* npbx_kind */
  int32_t dtype;   /* npbx_dtype, the same for every grid of a plan */
  int64_t n;       /* side length */
  void *A;         /* the reference's A and B arrays (device) */
  void *B;
  int64_t first;   /* first sweep index (its parity picks the source array) */
  int64_t nsweeps; /* sweeps to run */
} npbx_grid_spec;

int npbx_plan_create(const npbx_grid_spec *grids, int32_t count,
                     int32_t tblock, int32_t seg, void **plan);
int npbx_plan_launch(void *plan, void *stream);
int npbx_plan_capture(void *plan, void *stream);
int npbx_plan_info(void *plan, int64_t *launches, int64_t *tiles,
                   int64_t *lups);
int npbx_plan_destroy(void *plan);

/* Serving form: one launch advances each listed grid by one round of its own
 * schedule (continuous batching: the caller changes the list between
 * launches).  Round i reads level `level0` of grid i from X (A or B) and
 * writes level level0+depth to the other array; depth <= tblock.  Rounds
 * that follow npbx_round_depths() for a grid's sweep range leave A and B
 * exactly as npbx_jacobi2d_sweeps / npbx_heat3d_sweeps would.  `queue` is a
 * work-queue state from npbx_queue_create, used by one stream at a time. */
typedef struct npbx_round_spec {
  int32_t kind;
  int32_t dtype;
  int64_t n;
  void *A;
  void *B;
  const void *X;
  void *Y;
  int64_t level0;
  int32_t depth;
  int32_t reserved;
  /* FP32 heat only, optional: the grid's boundary columns packed by
   * npbx_pack_cols (rounds then read them from here instead of from A and
   * B); NULL otherwise.  Results do not depend on it. */
  const void *cols;
} npbx_round_spec;

int npbx_round_depths(int64_t nsweeps, int32_t tblock, int32_t *depths,
                      int32_t capacity, int32_t *count);
int npbx_queue_create(void **queue);
int npbx_queue_destroy(void *queue);
int npbx_rounds_launch(const npbx_round_spec *rounds, int32_t count,
                       int32_t tblock, int32_t seg, void *queue, void *stream);

/* Boundary columns of an FP32 heat grid (columns 0 and n-1 of every row of A
 * and of B), packed into `cols` of npbx_cols_elems(n) = 4*n*n + 1 floats
 * (the last word records whether A's and B's boundaries are equal), for
 * npbx_round_spec.cols.  Pack again after A's or B's boundary changes.  No
 * counterpart in the reference (a layout for the batched/serving forms). */
int npbx_cols_elems(int64_t n, int64_t *elems);
int npbx_pack_cols(int dtype, const void *A, const void *B, int64_t n,
                   void *cols, void *stream);

/* ---- tuning ------------------------------------------------------------------
 * Process-wide knobs read when tile launches are planned (the autotuner in
 * npbench_amd/tune.py sets them): "heat.rows" = rows per wave of a heat tile
 * (2, 4, 6, 8; built for T=2: 2..8 (FP32), 2..6 (FP64), 2 and 4 (BF16); T=3:
 * 4 and 6; T=4: 4; and for FP32 10 and 12 at T=2, 12 and 14 at T=3, which
 * select the wave-private heat tile: rows per wave including its own 2T halo
 * rows),
 * "heat.seg" = default planes per heat tile, "sweep.shape" = the plain
 * sweep's planes per workgroup / planes loaded ahead (0: 16/4, 1: 16/8,
 * 2: 32/8, 3: 8/4), "hdiff.tblock" = iterated hdiff steps fused per HBM round
 * trip (1, the default, 2: the two-step register march with strips of
 * "hdiff.fuse_rows" output rows (0, the default: 16 for BF16, 32
 * otherwise), or 4: LDS blocks; BF16 at most 2), "adi.path" = ADI launches
 * (0, the default: by size; 1: all time steps in one single-workgroup
 * launch, up to 512 lines; 2: one launch per half step), "vadv.path" = vadv
 * kernel (0,
 * the default: by shape; 1: c' and d' through a device workspace; 2: c' and
 * d' in LDS where K <= 160 and columns hold whole 16-byte vectors; 3: c' in
 * registers, d' in LDS, 80 < K <= 160; 4: loads coalesced and staged in
 * LDS, c' and d' in registers, 80 < K <= 160 with whole 64-byte chunks of
 * levels -- the default where it applies), "hdiff.shape" = hdiff marching
 * shape (0, the default: by storage type; 1 .. 13: columns per lane, rows
 * per thread, bytes of k per lane),
 * "seidel.blocks" = wavefront
 * slot-blocks (slots x 256-column blocks) per seidel_2d launch (default 2048,
 * about 40 ms), "jacobi.deep" = 2-D Jacobi sweeps per launch at tblock <= 1
 * (default 8: LDS-blocked, 64 x 48 points per workgroup; 0 or 1: one launch
 * per sweep), "jacobi.deep_waves" = its waves per workgroup (0, the default:
 * by shape; 4, 8 or 16).  Results do not depend on them. */
int npbx_set_tuning(const char *key, int64_t value);
int npbx_get_tuning(const char *key, int64_t *value);

/* ---- horizontal diffusion (hdiff_numpy.py:5-29) ----------------------------
 * in_field: (I+4, J+4, K); out_field, coeff: (I, J, K).  nsteps == 1 is the
 * reference; nsteps > 1 iterates, feeding out back into the interior of the
 * input with its 2-cell halo held fixed (in_field itself is not modified). */
int npbx_hdiff(int dtype, const void *in_field, void *out_field,
               const void *coeff, int64_t I, int64_t J, int64_t K,
               int64_t nsteps, void *stream);

/* ---- implicit vertical advection (vadv_numpy.py:9-78) ----------------------
 * utens_stage (in/out), u_stage, u_pos, utens: (I, J, K); wcon: (I+1, J, K).
 * field_dtype NPBX_F64 or NPBX_F32; the elimination is FP64 in both.  K >= 2.
 */
int npbx_vadv(int field_dtype, void *utens_stage, const void *u_stage,
              const void *wcon, const void *u_pos, const void *utens,
              double dtr_stage, int64_t I, int64_t J, int64_t K, void *stream);

/* ---- 1-D three-point Jacobi (jacobi_1d_numpy.py:4-8) -----------------------
 * A, B: (N,) FP64 (dtype NPBX_F64), updated in place by TSTEPS - 1 ping-pong
 * iterations; bitwise the NumPy result.  Boundaries are never written. */
int npbx_jacobi1d(int dtype, void *A, void *B, int64_t N, int64_t tsteps,
                  void *stream);

/* ---- 2-D finite-difference time domain (fdtd_2d_numpy.py:4-11) ------------
 * ex, ey, hz: (NX, NY) FP64 (dtype NPBX_F64), updated in place by TMAX time
 * steps; fict: (>= TMAX,) FP64, read.  Bitwise the NumPy result. */
int npbx_fdtd2d(int dtype, void *ex, void *ey, void *hz, const void *fict,
                int64_t NX, int64_t NY, int64_t tmax, void *stream);

/* ---- lid-driven cavity flow (cavity_flow_numpy.py:13-83) -------------------
 * u, v, p: (ny, nx) FP64 (dtype NPBX_F64), updated in place by nt time steps
 * of nit pressure iterations each, with the reference's argument order and
 * scalars (cavity_flow_numpy.py:41); bitwise the NumPy result.  3 <= nx, ny
 * and nx * ny <= 2^26; nt == 0 is a no-op.  One workgroup runs the steps,
 * in launches of about 50 ms of work each; grids of up to 10240 points keep
 * the pressure in LDS.  A device
 * workspace of 4 nx ny doubles is reused between calls. */
int npbx_cavity_flow(int dtype, int64_t nx, int64_t ny, int64_t nt, int64_t nit,
                     void *u, void *v, double dt, double dx, double dy, void *p,
                     double rho, double nu, void *stream);

/* ---- channel flow (channel_flow_numpy.py:13-152) --------------------------
 * u, v, p: (ny, nx) FP64 (dtype NPBX_F64), updated in place by time steps of
 * nit pressure iterations each, periodic in x, until the relative change of
 * sum(u) is at most 0.001 (the reference's while loop); the number of steps
 * is returned in *stepcount.  The reference's arguments, plus nx, ny and
 * max_steps, the most time steps one kernel launch may run (>= 1): launches
 * repeat until the criterion stops one, and the call returns after a stream
 * synchronisation.  Bitwise the NumPy result when every step's stopping
 * decision matches NumPy's (the sums' order differs from NumPy's pairwise
 * blocks; only a udiff within rounding of 0.001 could tell them apart).
 * 3 <= nx, ny; nx * ny <= 2^26. */
int npbx_channel_flow(int dtype, int64_t nit, void *u, void *v, double dt,
                      double dx, double dy, void *p, double rho, double nu,
                      double F, int64_t nx, int64_t ny, int64_t max_steps,
                      int64_t *stepcount, void *stream);

/* One bounded step of npbx_channel_flow: a single launch of at most
 * max_steps time steps continuing from u, v, p, then a stream
 * synchronisation; *steps = the steps it ran, *converged = 1 when the
 * reference's criterion stopped it.  Repeating it until *converged is
 * npbx_channel_flow, with control back in the caller between launches (so
 * a caller's timeout can stop an input that never converges, where the
 * reference's loop runs until interrupted). */
int npbx_channel_flow_launch(int dtype, int64_t nit, void *u, void *v,
                             double dt, double dx, double dy, void *p,
                             double rho, double nu, double F, int64_t nx,
                             int64_t ny, int64_t max_steps, int64_t *steps,
                             int32_t *converged, void *stream);

/* ---- 2-D Gauss-Seidel (seidel_2d_numpy.py:4-13) ----------------------------
 * A: (N, N) FP64 (dtype NPBX_F64), updated in place by TSTEPS - 1 sweeps;
 * bitwise the NumPy result.  N <= 16384; N < 3 or TSTEPS <= 1 is a no-op.
 * Rows run in wavefront slots (row i of sweep t with row i+2 of sweep t-1),
 * one workgroup, at most 64 sweeps per slot schedule; the rows of a slot run
 * their recurrences side by side, one lane each.  A launch runs a range of
 * slots of at most "seidel.blocks" slot-blocks, so its duration is bounded
 * whatever N is. */
int npbx_seidel2d(int dtype, void *A, int64_t N, int64_t tsteps, void *stream);

/* ---- alternating-direction implicit solver (adi_numpy.py:6-52) -------------
 * u: (N, N) FP64 (dtype NPBX_F64), updated in place by TSTEPS time steps of
 * the reference, each a column sweep into library scratch and a row sweep
 * back into u; bitwise the NumPy result.  N < 3 or TSTEPS < 0 leave u
 * unchanged; TSTEPS == 0 is an error (the reference divides by it). */
int npbx_adi(int dtype, void *u, int64_t N, int64_t tsteps, void *stream);

/* npbx_adi without allocation.  No counterpart in the reference: for callers
 * that cannot allocate device memory (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc).  Same arguments, kernels and
 * results as npbx_adi (bitwise), but its scratch is a fixed 16 MiB array
 * inside the library, one per GPU, held until the process ends, instead of
 * the grow-only library workspace: N up to 1000 fits, and a larger N is
 * rejected with NPBX_EINVAL before anything is enqueued.  Calls no hipMalloc
 * or hipFree.  Every call on one GPU shares that GPU's array, so calls that
 * overlap on one GPU (different streams, or threads) race: serialise them.
 * The array's device address is resolved once per device, by npbx_init or
 * else by the first call on that device. */
int npbx_adi_noalloc(int dtype, void *u, int64_t N, int64_t tsteps, void *stream);

/* Self-check of npbx_adi's short division (no counterpart in the reference;
 * for tests): `pairs` seeded random FP64 quotients, half of them next to a
 * rounding midpoint, through the shortened IEEE division sequence the ADI
 * forward passes use; *checked counts the pairs inside its no-rescale range,
 * *wrong those whose bits differ from the division operator's.  Synchronous,
 * on the null stream. */
int npbx_adi_division_check(int64_t pairs, uint64_t seed, int64_t *checked,
                            int64_t *wrong);

#ifdef __cplusplus
}
#endif

#endif /* NPBX_H */
