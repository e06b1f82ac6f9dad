/*
 * tasep_hip.h -- C ABI of the MI355X (gfx950) master-equation engine.
 *
 * Drop-in for the compute section of KilianPiochUBT/TASEP's masterEquation:
 * the generator build (gencrs), the stage kernels (rk4_step,
 * gen_meq_row + rk4_step_reduced), the update/checksum/early-stop loop and
 * the moment reductions (condense, exp_val_eval, the covariance loops).
 * The reference calls those as plain C functions on host buffers
 * (src/lib/matgen.h:12-19, src/lib/solv.h:34-62, src/lib/matvec.h:9-34) from
 * src/masterEquation.c:586-992 and :1009-1139,1291-1303.  Here the state
 * vector lives in HBM for the whole run; the host hands over rates and x0
 * once and reads back n + 2n^2 + 1 doubles (plus x only when asked).
 *
 * Conventions (as the reference): state s in [0, 2^n), bit k of s is site
 * n-k counted from the entry (bit n-1 entry site, bit 0 exit site);
 * r[0] = alpha, r[n] = beta, r[j] = hop rate from site j to j+1.
 *
 * Every call returns 0 on success or a negative TASEP_ERR_* code
 * (tasep_strerror).  A context is single-threaded: calls on one context must
 * not overlap.  Each context owns its device memory and one HIP stream of
 * default priority.
 */
#ifndef TASEP_HIP_H
#define TASEP_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TASEP_OK               0
#define TASEP_ERR_ARG         -1  /* invalid argument / combination          */
#define TASEP_ERR_NOMEM       -2  /* device or host allocation failed        */
#define TASEP_ERR_HIP         -3  /* HIP runtime error                       */
#define TASEP_ERR_STATE       -4  /* call out of order (e.g. no rates set)   */
#define TASEP_ERR_NODEV       -5  /* no usable gfx950 device                 */

/* precision of the state vector */
#define TASEP_F64   0  /* FP64 vector, FP64 arithmetic                        */
#define TASEP_F32   1  /* FP32 vector and FP32 generator rows, stage
                          combination and update in every kernel; checksum,
                          mass, moments and time-variant rates in FP64        */
/* integrator */
#define TASEP_REF   0  /* the reference's loop: all four stages read the
                          unmodified x (src/lib/solv.c:12-13), update
                          x += (((k + 2k) + 2k) + k)/6, bit-faithful in FP64  */
#define TASEP_RK4   1  /* classical fourth-order Runge-Kutta                  */
/* generator */
#define TASEP_MF    0  /* matrix-free: rows derived from the state bitmask    */
#define TASEP_CSR   1  /* stored CSR built on the device (the -f path)        */

typedef struct tasep_ctx tasep_ctx;

typedef struct {
    uint64_t steps_done;     /* steps executed by this call                   */
    uint64_t steps_at_stop;  /* index of the last sub-threshold step          */
    int32_t  early_stop;     /* consecutive sub-threshold steps at the end    */
    int32_t  stopped;        /* 1: left the loop on early_stop > patience     */
    double   t;              /* time after the last step (t0 + h + h + ...)   */
    double   last_checksum;  /* sum_s |dx_s| of the last step                 */
    double   mass;           /* sum_s x_s after the last step (FP64 sum)      */
} tasep_run_info;

/* Create a context for a lattice of n sites (2 <= n <= 34).
 * Valid combinations: REF requires F64; CSR requires F64 and n <= 30.
 * Replaces the allocations of src/masterEquation.c:349,472-505,593-614. */
int tasep_ctx_create(tasep_ctx **ctx, int n, int precision, int integrator,
                     int generator);
int tasep_ctx_destroy(tasep_ctx *ctx);

/* n+1 constant jump rates (src/masterEquation.c:432-461). */
int tasep_set_rates(tasep_ctx *ctx, const double *r);

/* Time-variant rates r_i(t) = max(0, A_i*f(w_i*t + phi_i) + c_i), f = sin
 * (type 0), cos (type 1) or A_i*t + c_i (type 2); each array has n+1
 * entries.  Replaces timeVariantProbs (src/lib/matvec.c:44-68), evaluated on
 * the device in FP64 once per step at the step's start time for REF
 * (src/masterEquation.c:886) and at t, t+h/2, t+h for RK4.  MF only. */
int tasep_set_time_variant(tasep_ctx *ctx, const int32_t *type,
                           const double *amp, const double *period,
                           const double *offset, const double *c);

/* Initial states: uniform 2^-n (src/masterEquation.c:402-409), a single
 * occupied basis state e_s, or an arbitrary host FP64 vector of 2^n. */
int tasep_set_x0_uniform(tasep_ctx *ctx);
int tasep_set_x0_state(tasep_ctx *ctx, uint64_t s);
int tasep_set_x(tasep_ctx *ctx, const double *x);
/* Product measure x_s = prod_k (b_k(s) ? p[k] : 1 - p[k]) over the n bits
 * (p[k] = occupation probability of site n-k, in [0, 1]), built on the
 * device in FP64 with k ascending.  p[k] = 1/2 for every k is the
 * reference's uniform start (src/masterEquation.c:402-409); the reference
 * has no other (its x0 is edited in the source, :411). */
int tasep_set_x0_product(tasep_ctx *ctx, const double *p);
/* Copy the current state to a host FP64 vector of 2^n entries. */
int tasep_get_x(tasep_ctx *ctx, double *x);
/* Copy entries [offset, offset + count) of the current state to host FP64
 * (reads a large state in pieces, as the probabilities file of
 * src/masterEquation.c:1021-1040 walks it in order). */
int tasep_get_x_range(tasep_ctx *ctx, uint64_t offset, uint64_t count,
                      double *x);

/* Build the CSR generator on the device (replaces gencrs,
 * src/lib/matgen.c:7-74; arrays bit-identical to it).  Requires rates. */
int tasep_build_csr(tasep_ctx *ctx);
/* Copy the device CSR to host arrays sized 2^n+1 / tasep_csr_nnz(n). */
int tasep_get_csr(tasep_ctx *ctx, uint64_t *rowPtr, uint64_t *colInd,
                  double *val);
/* nnz = 2*2^n + (n-1)*2^(n-2)  (src/masterEquation.c:593). */
uint64_t tasep_csr_nnz(int n);

/* Same signature meaning as the reference's gencrs (src/lib/matgen.h:12-19):
 * fills host rowPtr[dim+1], valArr[nnz], colInd[nnz] from jump_rate_Array,
 * building on the device.  No scratch `bit` array is needed. */
int tasep_gencrs(uint64_t *rowPtr, double *valArr, uint64_t *colInd,
                 const double *jump_rate_Array, uint64_t dim, int n);

/* Run `steps` steps of size h from time t0 (replaces the loops at
 * src/masterEquation.c:628-682, :883-921, :929-982).  With constant rates the
 * loop leaves after more than `patience` consecutive steps whose checksum
 * sum_s|dx_s| is below thr*h (reference: thr = 1e-7, patience = 100,
 * :186,670-679); patience < 0 disables the exit.  Time-variant runs never
 * exit early (as the reference).  The state stays on the device. */
int tasep_run(tasep_ctx *ctx, double t0, double h, uint64_t steps, double thr,
              int patience, tasep_run_info *info);

/* tasep_run in pieces: with resume != 0 the loop continues where the last
 * call stopped (t, the step index and the early-exit counters are kept;
 * t0 is ignored), so a caller can take snapshots every few steps as the
 * reference's -a option does inside its loop (src/masterEquation.c:646-666).
 * info->steps_done counts all steps since the last call with resume == 0. */
int tasep_run_resume(tasep_ctx *ctx, double t0, double h, uint64_t steps,
                     double thr, int patience, int resume,
                     tasep_run_info *info);

/* Moments of the current state in one device pass with FP64 sums
 * (replaces condense/exp_val_eval, src/lib/matvec.c:10-91, and the two
 * O(n^2 2^n) loops at src/masterEquation.c:1127-1139,1291-1303):
 *   y[i]       = sum_s b_i(s) x_s                  (i = bit, 0 = exit site)
 *   cov[i*n+k] = sum_s (b_k - y_k)(b_i - y_i) x_s
 *   mul[i*n+k] = y_i y_k + cov[i*n+k]
 *   *mass      = sum_s x_s
 * For TASEP_F32 contexts x is first divided by its FP64 mass.  Any output
 * pointer may be NULL.  Index by bit; the reference writes the matrices
 * with both axes reversed (entry site first). */
int tasep_moments(tasep_ctx *ctx, double *y, double *cov, double *mul,
                  double *mass);

/* tasep_moments without any allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  x is a
 * caller-owned device vector of 2^n entries (double for TASEP_F64, float for
 * TASEP_F32) on the current device, e.g. from tasep_get_x_device; results as
 * tasep_moments (any
 * output pointer may be NULL), 2 <= n <= 34.  The block partials live in a
 * static device array of 1.17 MiB per GPU (256 blocks x 596 doubles plus
 * 596 sums) instead of allocated memory; every call on a GPU shares that
 * GPU's copy, so calls that overlap on one GPU race.  The work is enqueued
 * on `stream` (a hipStream_t; NULL = the default stream), which is
 * synchronized before the call returns. */
int tasep_moments_noalloc(const void *x, int n, int precision, double *y,
                          double *cov, double *mul, double *mass,
                          void *stream);

/* Window marginals of the current state in one device pass with FP64 sums:
 *   W[o*16 + p] = sum_s x_s [((s >> o) & 15) == p],  o = 0..n-4, p = 0..15
 * ((n-3)*16 doubles; n >= 4; FP32 contexts divided by the FP64 mass).
 * Every high-order output the reference writes for the final state --
 * key_fit/high_order_corr, error guess, closures, point functions
 * (src/masterEquation.c:32-143, :1422-1514), each a loop over all 2^n
 * states per printed value -- is a sum of these pattern probabilities over
 * windows of <= 4 sites. */
int tasep_window_marginals(tasep_ctx *ctx, double *W);
/* The same for windows of `width` = 4 or 6 sites: W[o*2^width + p],
 * o = 0..n-width ((n-width+1)*2^width doubles; n >= width).  Width 6 covers
 * the timed correlation of `-a` runs (high_order_corr with a 4-site key,
 * src/masterEquation.c:1516-1537), whose sums span 6 sites. */
int tasep_window_marginals_w(tasep_ctx *ctx, int width, double *W);

/* Per-launch timing of one step, for measurement: runs `reps` steps from t0
 * (advancing the state like tasep_run with patience < 0), timing every
 * kernel launch of a step with HIP events on the context's stream, and
 * returns for each launch position its name, mean duration and algorithmic
 * HBM bytes per state.  No reference counterpart (the reference times whole
 * phases with gettimeofday, src/masterEquation.c:589,618-626,993-1001). */
typedef struct {
    char   name[32];
    double ms;
    double bytes_per_state;
} tasep_kernel_time;
int tasep_step_profile(tasep_ctx *ctx, double t0, double h, int reps,
                       tasep_kernel_time *out, int max, int *count);

/* Engine options (no reference counterpart; defaults in brackets):
 *   TASEP_OPT_TILED        [1] tiled matrix-free passes: 1 for n >= 18
 *                              (RK4) / n >= 19 (reference loop), 2 wherever
 *                              a lattice has two tiles (n >= 14 FP32,
 *                              n >= 13 FP64), 0 = one lane per state
 *   TASEP_OPT_XCD_ORDER    [4] order of the tiles in the tiled passes: 0 grid
 *                              order, 1 contiguous eighths per XCD, 2 eighths
 *                              per XCD each swept in order by the workgroups
 *                              resident on it, 3 one sweep over all tiles,
 *                              4 as 2 over a level walk: tiles grouped by the
 *                              particles on their upper tile-index bits (by
 *                              count, then position sum), so that a tile and
 *                              the tiles it reads across tile bonds are
 *                              walked close together (a table of 4 bytes per
 *                              tile, built on first use)
 *   TASEP_OPT_RK4_STAGED   [0] RK4 with four stored stages even for constant
 *                              rates (default: the Horner form of the same
 *                              step, 3 work vectors instead of 5)
 *   TASEP_OPT_TILE_BLOCKS  [0] workgroups of the tiled passes (multiple of
 *                              8, at most 16384); 0 = one per workgroup slot
 *                              resident at once (256 on 256 CUs) up to 2^15
 *                              tiles (2^16 for the reference loop), 4 per
 *                              slot above that, 4096 / 16384 for RK4 passes
 *                              over 2^17 / at least 2^18 tiles, and at most
 *                              one per 8 tiles but at least one per CU
 *   TASEP_OPT_LDS          [1] matrix-free runs that keep two copies of the
 *                              state vector in one workgroup's LDS and run
 *                              all their steps there, in one launch per
 *                              ~2^27 state-updates (<= 65536 steps):
 *                              0 never, 1 for n <= 12, 2 wherever it fits
 *                              (n <= 13)
 *   TASEP_OPT_LDS_THREADS  [0] workgroup size of those runs (power of two in
 *                              [64, 1024]); 0 = 2^n clamped to that range
 *   TASEP_OPT_GRAPH        [0] runs of at least two chunks (a chunk is
 *                              ~2^26 state-updates, 4..4096 steps) replay a
 *                              HIP graph of one chunk's launches (measured
 *                              no faster than direct launches on MI355X)
 *   TASEP_OPT_STORE        [0] output stores of the tiled passes: 0 by
 *                              lattice size, 1 non-temporal, 2 ordinary
 *                              (the results are the same bits) */
#define TASEP_OPT_TILED        1
#define TASEP_OPT_XCD_ORDER    2
#define TASEP_OPT_RK4_STAGED   3
#define TASEP_OPT_TILE_BLOCKS  4
#define TASEP_OPT_LDS          5
#define TASEP_OPT_LDS_THREADS  6
#define TASEP_OPT_GRAPH        7
#define TASEP_OPT_STORE        8
int tasep_set_option(tasep_ctx *ctx, int key, int64_t value);

/* The context's HIP stream (hipStream_t) and device pointer of the current
 * state vector (double* or float* by precision), for callers that time or
 * consume the state on the device. */
int tasep_get_stream(tasep_ctx *ctx, void **stream);
int tasep_get_x_device(tasep_ctx *ctx, void **ptr);

/* ------------------------------------------------------------------------
 * Batches: many independent instances (own n, rates or rate functions,
 * state and run state) driven together.  No reference counterpart: the
 * reference runs one masterEquation process per instance
 * (src/masterEquation.c:main); a batch replaces a loop over such runs with
 * classical RK4 (time-variant instances with stored stages).  Instances
 * with n <= 13 run all their steps inside one workgroup each (one launch
 * for all of them); larger ones share one launch per RK4 pass, replayed
 * from a captured HIP graph.  Run info and early exit are per instance.
 * ---------------------------------------------------------------------- */
typedef struct tasep_batch tasep_batch;

/* count instances of n[i] sites (2 <= n[i] <= 30), TASEP_F32 or TASEP_F64;
 * every state starts uniform. */
int tasep_batch_create(tasep_batch **batch, int count, const int32_t *n,
                       int precision);
int tasep_batch_destroy(tasep_batch *batch);
/* Rates of instance i (n[i]+1 values, as tasep_set_rates) or its rate
 * functions (as tasep_set_time_variant). */
int tasep_batch_set_rates(tasep_batch *batch, int i, const double *r);
int tasep_batch_set_time_variant(tasep_batch *batch, int i,
                                 const int32_t *type, const double *amp,
                                 const double *period, const double *offset,
                                 const double *c);
int tasep_batch_set_x0_uniform(tasep_batch *batch);
int tasep_batch_set_x(tasep_batch *batch, int i, const double *x);
int tasep_batch_get_x(tasep_batch *batch, int i, double *x);
/* Run every instance `steps` steps from t0; info (may be NULL) receives
 * `count` entries.  thr/patience as tasep_run, per instance. */
int tasep_batch_run(tasep_batch *batch, double t0, double h, uint64_t steps,
                    double thr, int patience, tasep_run_info *info);
/* Moments and window marginals of instance i, as tasep_moments and
 * tasep_window_marginals. */
int tasep_batch_moments(tasep_batch *batch, int i, double *y, double *cov,
                        double *mul, double *mass);
int tasep_batch_window_marginals(tasep_batch *batch, int i, double *W);
/* Batch options (defaults in brackets):
 *   TASEP_BATCH_OPT_GRAPH_STEPS [32] steps per captured HIP graph of the
 *                                    shared passes; 0 = direct launches
 *   TASEP_BATCH_OPT_LDS         [1]  instances with n <= 13 run in LDS;
 *                                    0 = only those with n < 10 do
 *   TASEP_BATCH_OPT_CONCURRENT  [1]  the LDS-resident instances run on a
 *                                    second stream beside the shared passes
 *   TASEP_BATCH_OPT_SCHEDULE    [0]  shared passes: 1 one lane per state,
 *                                    2 tiles of 2^10 (FP32) / 2^9 (FP64)
 *                                    states in LDS, 3 the single-lattice
 *                                    tile pass (tiles of 2^13 / 2^12 states
 *                                    with the top three bits, a range of
 *                                    tiles per workgroup; every streamed
 *                                    instance then has exactly a single
 *                                    context's tiled results; needs
 *                                    n >= 14 / 13), 0 = timed on the first
 *                                    run and the fastest that fits kept
 *                                    (get_option then reports the choice;
 *                                    the schedules agree to rounding, so
 *                                    pin one for bitwise run-to-run
 *                                    reproducibility)
 *   TASEP_BATCH_OPT_RANGE_TILES [2]  schedule 3: tiles per workgroup
 *   TASEP_BATCH_OPT_STORE       [0]  schedule 3 output stores: 0 by the
 *                                    batch vector's size (ordinary for
 *                                    16-128 MiB, non-temporal otherwise),
 *                                    1 non-temporal, 2 ordinary */
#define TASEP_BATCH_OPT_GRAPH_STEPS 1
#define TASEP_BATCH_OPT_LDS         2
#define TASEP_BATCH_OPT_CONCURRENT  3
#define TASEP_BATCH_OPT_SCHEDULE    4
#define TASEP_BATCH_OPT_RANGE_TILES 5
#define TASEP_BATCH_OPT_STORE       6
int tasep_batch_set_option(tasep_batch *batch, int key, int64_t value);
int tasep_batch_get_option(tasep_batch *batch, int key, int64_t *value);
int tasep_batch_get_stream(tasep_batch *batch, void **stream);

const char *tasep_strerror(int code);

#ifdef __cplusplus
}
#endif
#endif /* TASEP_HIP_H */
