/*
 * tasep_oracle.h -- CPU restatement of KilianPiochUBT/TASEP's masterEquation
 * hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path: it
 * may be loaded by tests/, by __graft_entry__.smoke() and by the cpu_baseline
 * leg of bench.py, and by nothing else.  The product (libtasep_hip.so and the
 * C host) never links or calls it.
 *
 * Conventions (reference src/masterEquation.c, src/lib/matgen.c):
 *   state s in [0, 2^n); bit k of s is physical site n-k (bit n-1 = entry
 *   site, bit 0 = exit site); r[0] = alpha, r[n] = beta, r[j] = hop j -> j+1;
 *   bond k (particle at bit k+1 hops to empty bit k) has rate r[n-1-k].
 */
#ifndef TASEP_ORACLE_H
#define TASEP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* nnz of the generator in CSR: 2*2^n + (n-1)*2^(n-2)
 * (reference src/masterEquation.c:593). */
uint64_t tor_nnz(int n);

/* Row length of state s: 1 + b_{n-1} + !b_0 + #{k<n-1 : b_k && !b_{k+1}}
 * (follows the entry/diag/exit/bulk emission of src/lib/matgen.c:29-67). */
uint64_t tor_rowlen(int n, uint64_t s);

/* Serial CSR build, same row/term order and same diagonal subtraction order
 * as gencrs (src/lib/matgen.c:7-74).  Arrays sized dim+1 / tor_nnz(n). */
void tor_gencrs(int n, const double *r, uint64_t *rowPtr, uint64_t *colInd,
                double *val);

/* One matrix-free row, same order as gen_meq_row (src/lib/solv.c:81-135).
 * Returns the number of (col, val) pairs written (<= n+2). */
int tor_row(int n, uint64_t s, const double *r, uint64_t *col, double *val);

/* k[i] = h * sum_j val_ij * x[col_ij], sum from 0 in row order, no FMA
 * (rk4_step, src/lib/solv.c:3-53).  nthreads <= 0: OpenMP default. */
void tor_apply_csr(uint64_t dim, const uint64_t *rowPtr, const uint64_t *colInd,
                   const double *val, const double *x, double h, double *k,
                   int nthreads);

/* Same as tor_apply_csr but regenerating each row (gen_meq_row +
 * rk4_step_reduced, src/lib/solv.c:56-135). Serial like the reference. */
void tor_apply_mf(int n, const double *r, const double *x, double h, double *k);

typedef struct {
    uint64_t steps_done;     /* loop iterations executed                     */
    uint64_t steps_at_stop;  /* last sub-threshold iteration index           */
    int      early_stop;     /* consecutive sub-threshold count at the end   */
    int      stopped;        /* 1 if the run broke out on early_stop>patience */
    double   t;              /* t0 + h + h + ... (sequential)                */
    double   last_checksum;  /* sum |dx| of the last executed step           */
} tor_run_info;

/* h = (double)((long double)(T - t) / steps)  (src/masterEquation.c:586). */
double tor_step_h(double T, double t, unsigned int steps);

/* Reference "RK4" loop with constant rates (src/masterEquation.c:628-682 for
 * use_csr=1, :929-982 for use_csr=0).  All four stages evaluate the generator
 * at the unmodified x, so the update is
 *     d = (((k + 2k) + 2k) + k) / 6,  x += d,  checksum += |d|
 * with the early exit after more than `patience` consecutive steps whose
 * checksum is below thr*h.  patience < 0 disables the early exit. */
int tor_run_ref(int n, const double *r, double *x, double t0, double h,
                uint64_t steps, double thr, int patience, int use_csr,
                int nthreads, tor_run_info *info);

/* Time-variant rates (src/lib/matvec.c:44-68): type 0 sin, 1 cos, 2 linear;
 * clamped at 0. */
void tor_tv_rates(double *r, int len, const double *period,
                  const double *offset, const double *amp, double t,
                  const int *type, const double *c);

/* Reference loop with time-variant rates (src/masterEquation.c:883-921):
 * rates evaluated once per step at the step's start time, no early exit. */
int tor_run_ref_tv(int n, double *x, double t0, double h, uint64_t steps,
                   const int *type, const double *amp, const double *period,
                   const double *offset, const double *c, tor_run_info *info);

/* Classical RK4 (the build's "rk4" integrator; no reference loop):
 *   k1 = hL(t)x, k2 = hL(t+h/2)(x+k1/2), k3 = hL(t+h/2)(x+k2/2),
 *   k4 = hL(t+h)(x+k3), x += (k1 + 2k2 + 2k3 + k4)/6.
 * tv == NULL: constant rates r.  Otherwise tv_* arrays give the rate
 * functions and r is ignored. */
int tor_run_rk4(int n, const double *r, double *x, double t0, double h,
                uint64_t steps, const int *tv_type, const double *tv_amp,
                const double *tv_period, const double *tv_offset,
                const double *tv_c, tor_run_info *info);

/* y[i] = sum_s b_i(s) x_s, sequential in s (condense / exp_val_eval,
 * src/lib/matvec.c:10-42,70-91).  Index i is the bit (i = 0: exit site). */
void tor_site_moments(int n, const double *x, double *y);

/* Centred covariance and multiplicative expectation in the orientation the
 * reference writes (src/masterEquation.c:1127-1139, 1291-1303):
 *   cov[(n-1-i)*n + (n-1-k)] = sum_s (b_k - y_k)(b_i - y_i) x_s
 *   mul[(n-1-i)*n + (n-1-k)] = y_k*y_i + cov
 * y from tor_site_moments. */
void tor_cov(int n, const double *x, const double *y, double *cov, double *mul);

/* key_fit sum over all states (src/masterEquation.c:32-45): sum of x_s
 * over the states whose bits off..off+size-1 equal key. */
double tor_key_sum(int n, const double *x, uint64_t key, int off, int size);
/* Window marginals of `width` consecutive sites (4 for tor_windows):
 * W[o*2^width + p] = sum_s x_s [((s >> o) & (2^width - 1)) == p]. */
void tor_windows(int n, const double *x, double *W);
void tor_windows_w(int n, int width, const double *x, double *W);
/* Sequential sum of x (src/masterEquation.c:1009-1011). */
double tor_mass(uint64_t dim, const double *x);

/* The reference's update with four stage arrays, as written at
 * src/masterEquation.c:639-642; returns the checksum sum |d|.  (CPU-baseline
 * timing pairs it with the reference's own rk4_step.) */
double tor_ref_update4(uint64_t dim, double *x, const double *k1,
                       const double *k2, const double *k3, const double *k4);

/* Dense generator L[s*dim + c] for small n (tests: column sums, brute force).*/
void tor_dense(int n, const double *r, double *L);

#ifdef __cplusplus
}
#endif
#endif
