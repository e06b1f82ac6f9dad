/*
 * tasep_oracle.c -- CPU restatement of the reference masterEquation hot path.
 *
 * TEST INFRASTRUCTURE ONLY (checker for the HIP path; see tasep_oracle.h).
 * Built with the reference's own flags (gcc -std=c99 -O2 -march=native
 * -fopenmp, reference Makefile:22): ISO C mode keeps a*b+c uncontracted, so
 * every FP64 operation below rounds exactly as the reference's do.
 */
#include "tasep_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define BIT(s, k) (((s) >> (k)) & 1u)

uint64_t tor_nnz(int n)
{
    /* src/masterEquation.c:593 */
    return 2u * ((uint64_t)1 << n) + ((uint64_t)1 << (n - 2)) * (uint64_t)(n - 1);
}

uint64_t tor_rowlen(int n, uint64_t s)
{
    uint64_t len = 1;                     /* diagonal                     */
    if (BIT(s, n - 1)) len++;             /* entry: alpha from s-2^(n-1)  */
    if (!BIT(s, 0)) len++;                /* exit: beta from s+1          */
    for (int k = 0; k < n - 1; k++)       /* bulk hop into s along bond k */
        if (BIT(s, k) && !BIT(s, k + 1)) len++;
    return len;
}

/* Emit one row in the reference order (src/lib/matgen.c:29-67):
 * [entry] diag [exit] bulk(k ascending). */
static int emit_row(int n, uint64_t s, const double *r, uint64_t *col,
                    double *val)
{
    int c = 0;
    if (BIT(s, n - 1)) {                  /* matgen.c:29-34 */
        col[c] = s - ((uint64_t)1 << (n - 1));
        val[c] = r[0];
        c++;
    }
    double d = 0.;                        /* matgen.c:36-51 */
    if (BIT(s, 0)) d -= r[n];
    if (!BIT(s, n - 1)) d -= r[0];
    for (int k = 0; k < n - 1; k++)
        if (!BIT(s, k) && BIT(s, k + 1)) d -= r[n - 1 - k];
    col[c] = s;
    val[c] = d;
    c++;
    if (!BIT(s, 0)) {                     /* matgen.c:55-59 */
        col[c] = s + 1;
        val[c] = r[n];
        c++;
    }
    for (int k = 0; k < n - 1; k++) {     /* matgen.c:61-67 */
        if (BIT(s, k) && !BIT(s, k + 1)) {
            col[c] = s + ((uint64_t)1 << k);
            val[c] = r[n - 1 - k];
            c++;
        }
    }
    return c;
}

void tor_gencrs(int n, const double *r, uint64_t *rowPtr, uint64_t *colInd,
                double *val)
{
    uint64_t dim = (uint64_t)1 << n;
    rowPtr[0] = 0;
    for (uint64_t s = 0; s < dim; s++) {
        int c = emit_row(n, s, r, colInd + rowPtr[s], val + rowPtr[s]);
        rowPtr[s + 1] = rowPtr[s] + (uint64_t)c;
    }
}

int tor_row(int n, uint64_t s, const double *r, uint64_t *col, double *val)
{
    return emit_row(n, s, r, col, val);
}

void tor_apply_csr(uint64_t dim, const uint64_t *rowPtr, const uint64_t *colInd,
                   const double *val, const double *x, double h, double *k,
                   int nthreads)
{
#ifdef _OPENMP
    int nt = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel for schedule(static) num_threads(nt)
#endif
    for (uint64_t i = 0; i < dim; i++) {   /* src/lib/solv.c:8-27 */
        double T = 0;
        for (uint64_t j = rowPtr[i]; j < rowPtr[i + 1]; j++)
            T += val[j] * x[colInd[j]];
        k[i] = h * T;
    }
    (void)nthreads;
}

void tor_apply_mf(int n, const double *r, const double *x, double h, double *k)
{
    uint64_t dim = (uint64_t)1 << n;
    uint64_t col[66];
    double val[66];
    for (uint64_t s = 0; s < dim; s++) {   /* src/lib/solv.c:56-79 */
        int c = emit_row(n, s, r, col, val);
        double T = 0;
        for (int j = 0; j < c; j++)
            T += val[j] * x[col[j]];
        k[s] = h * T;
    }
}

double tor_step_h(double T, double t, unsigned int steps)
{
    return (double)((T - t) / ((long double)steps));  /* masterEquation.c:586 */
}

/* One reference update over the whole vector; returns sum |d| in index order
 * (src/masterEquation.c:639-642). */
static double ref_update(uint64_t dim, double *x, const double *k)
{
    double chk = 0;
    for (uint64_t s = 0; s < dim; s++) {
        double d = (k[s] + (2. * k[s]) + (2. * k[s]) + k[s]) / 6.;
        x[s] += d;
        chk += fabs(d);
    }
    return chk;
}

int tor_run_ref(int n, const double *r, double *x, double t0, double h,
                uint64_t steps, double thr, int patience, int use_csr,
                int nthreads, tor_run_info *info)
{
    uint64_t dim = (uint64_t)1 << n;
    double *k = (double *)malloc(dim * sizeof(double));
    uint64_t *rowPtr = NULL, *colInd = NULL;
    double *val = NULL;
    if (!k) return -1;
    if (use_csr) {
        uint64_t nnz = tor_nnz(n);
        rowPtr = (uint64_t *)malloc((dim + 1) * sizeof(uint64_t));
        colInd = (uint64_t *)malloc(nnz * sizeof(uint64_t));
        val = (double *)malloc(nnz * sizeof(double));
        if (!rowPtr || !colInd || !val) {
            free(k); free(rowPtr); free(colInd); free(val);
            return -1;
        }
        tor_gencrs(n, r, rowPtr, colInd, val);
    }
    tor_run_info ri;
    memset(&ri, 0, sizeof(ri));
    double t = t0;
    for (uint64_t i = 0; i < steps; i++) {
        if (use_csr)
            tor_apply_csr(dim, rowPtr, colInd, val, x, h, k, nthreads);
        else
            tor_apply_mf(n, r, x, h, k);
        double chk = ref_update(dim, x, k);
        t += h;                                     /* masterEquation.c:644 */
        ri.steps_done = i + 1;
        ri.last_checksum = chk;
        if (patience >= 0) {                        /* :670-679 */
            if (chk < thr * h) {
                ri.early_stop++;
                ri.steps_at_stop = i;
            } else {
                ri.early_stop = 0;
            }
            if (ri.early_stop > patience) {
                ri.stopped = 1;
                break;
            }
        }
    }
    ri.t = t;
    if (info) *info = ri;
    free(k); free(rowPtr); free(colInd); free(val);
    return 0;
}

void tor_tv_rates(double *r, int len, const double *period,
                  const double *offset, const double *amp, double t,
                  const int *type, const double *c)
{
    for (int i = 0; i < len; i++) {        /* src/lib/matvec.c:44-68 */
        double v;
        switch (type[i]) {
        case 0: v = amp[i] * sin(t * period[i] + offset[i]) + c[i]; break;
        case 1: v = amp[i] * cos(t * period[i] + offset[i]) + c[i]; break;
        case 2: v = amp[i] * t + c[i]; break;
        default: continue;                 /* reference leaves r[i] as is */
        }
        r[i] = v < 0 ? 0. : v;
    }
}

int tor_run_ref_tv(int n, double *x, double t0, double h, uint64_t steps,
                   const int *type, const double *amp, const double *period,
                   const double *offset, const double *c, tor_run_info *info)
{
    uint64_t dim = (uint64_t)1 << n;
    double *k = (double *)malloc(dim * sizeof(double));
    double *r = (double *)calloc((size_t)n + 1, sizeof(double));
    if (!k || !r) { free(k); free(r); return -1; }
    tor_run_info ri;
    memset(&ri, 0, sizeof(ri));
    double t = t0;
    for (uint64_t i = 0; i < steps; i++) {          /* masterEquation.c:883-921 */
        tor_tv_rates(r, n + 1, period, offset, amp, t, type, c);
        tor_apply_mf(n, r, x, h, k);
        double last = 0;
        for (uint64_t s = 0; s < dim; s++) {
            double d = (k[s] + (2. * k[s]) + (2. * k[s]) + k[s]) / 6.;
            x[s] += d;
            last = fabs(d);                          /* checksum '=' (:896) */
        }
        ri.last_checksum = last;
        t += h;
        ri.steps_done = i + 1;
    }
    ri.t = t;
    if (info) *info = ri;
    free(k); free(r);
    return 0;
}

/* y = h * L(r) * v, matrix-free, row order as the reference. */
static void apply_h(int n, const double *r, const double *v, double h, double *y)
{
    uint64_t dim = (uint64_t)1 << n;
    uint64_t col[66];
    double val[66];
#ifdef _OPENMP
#pragma omp parallel for schedule(static) private(col, val)
#endif
    for (uint64_t s = 0; s < dim; s++) {
        int c = emit_row(n, s, r, col, val);
        double T = 0;
        for (int j = 0; j < c; j++)
            T += val[j] * v[col[j]];
        y[s] = h * T;
    }
}

int tor_run_rk4(int n, const double *r, double *x, double t0, double h,
                uint64_t steps, const int *tv_type, const double *tv_amp,
                const double *tv_period, const double *tv_offset,
                const double *tv_c, tor_run_info *info)
{
    uint64_t dim = (uint64_t)1 << n;
    double *k1 = (double *)malloc(dim * sizeof(double));
    double *k2 = (double *)malloc(dim * sizeof(double));
    double *k3 = (double *)malloc(dim * sizeof(double));
    double *k4 = (double *)malloc(dim * sizeof(double));
    double *y = (double *)malloc(dim * sizeof(double));
    double *ra = (double *)malloc(((size_t)n + 1) * sizeof(double));
    double *rb = (double *)malloc(((size_t)n + 1) * sizeof(double));
    double *rc = (double *)malloc(((size_t)n + 1) * sizeof(double));
    int rc_ = 0;
    if (!k1 || !k2 || !k3 || !k4 || !y || !ra || !rb || !rc) { rc_ = -1; goto out; }
    tor_run_info ri;
    memset(&ri, 0, sizeof(ri));
    double t = t0;
    for (uint64_t i = 0; i < steps; i++) {
        if (tv_type) {
            tor_tv_rates(ra, n + 1, tv_period, tv_offset, tv_amp, t, tv_type, tv_c);
            tor_tv_rates(rb, n + 1, tv_period, tv_offset, tv_amp, t + 0.5 * h, tv_type, tv_c);
            tor_tv_rates(rc, n + 1, tv_period, tv_offset, tv_amp, t + h, tv_type, tv_c);
        } else {
            memcpy(ra, r, ((size_t)n + 1) * sizeof(double));
            memcpy(rb, r, ((size_t)n + 1) * sizeof(double));
            memcpy(rc, r, ((size_t)n + 1) * sizeof(double));
        }
        apply_h(n, ra, x, h, k1);
        for (uint64_t s = 0; s < dim; s++) y[s] = x[s] + 0.5 * k1[s];
        apply_h(n, rb, y, h, k2);
        for (uint64_t s = 0; s < dim; s++) y[s] = x[s] + 0.5 * k2[s];
        apply_h(n, rb, y, h, k3);
        for (uint64_t s = 0; s < dim; s++) y[s] = x[s] + k3[s];
        apply_h(n, rc, y, h, k4);
        double chk = 0;
        for (uint64_t s = 0; s < dim; s++) {
            double d = (k1[s] + 2. * k2[s] + 2. * k3[s] + k4[s]) / 6.;
            x[s] += d;
            chk += fabs(d);
        }
        t += h;
        ri.steps_done = i + 1;
        ri.last_checksum = chk;
    }
    ri.t = t;
    if (info) *info = ri;
out:
    free(k1); free(k2); free(k3); free(k4); free(y); free(ra); free(rb); free(rc);
    return rc_;
}

void tor_site_moments(int n, const double *x, double *y)
{
    uint64_t dim = (uint64_t)1 << n;
    for (int i = 0; i < n; i++) y[i] = 0.;
    for (uint64_t s = 0; s < dim; s++)     /* src/lib/matvec.c:36-41 */
        for (int i = 0; i < n; i++)
            if (BIT(s, i)) y[i] += x[s];
}

void tor_cov(int n, const double *x, const double *y, double *cov, double *mul)
{
    uint64_t dim = (uint64_t)1 << n;
    for (int i = 0; i < n; i++) {
        for (int k = 0; k < n; k++) {      /* masterEquation.c:1127-1139 */
            double acc = 0;
            for (uint64_t s = 0; s < dim; s++) {
                double bi = (double)BIT(s, i), bk = (double)BIT(s, k);
                acc += (bk - y[k]) * (bi - y[i]) * x[s];
            }
            cov[(n - 1 - i) * n + (n - 1 - k)] = acc;
            if (mul) mul[(n - 1 - i) * n + (n - 1 - k)] = y[k] * y[i] + acc;
        }
    }
}

/* The reference's key_fit sum (masterEquation.c:32-45, summed over all
 * states as in the loops of :61-68, :125-128): sum of x_s over the states
 * whose bits off..off+size-1 equal the low `size` bits of key. */
double tor_key_sum(int n, const double *x, uint64_t key, int off, int size)
{
    uint64_t dim = (uint64_t)1 << n;
    double acc = 0.;
    for (uint64_t s = 0; s < dim; s++) {
        int fit = 1;
        for (int i = 0; i < size; i++)
            if (((key >> i) & 1) != ((s >> (off + i)) & 1)) fit = 0;
        acc += fit ? x[s] : 0.;
    }
    return acc;
}

/* All window marginals of `width` consecutive sites:
 * W[o*2^width + p] = sum_s x_s [((s >> o) & (2^width - 1)) == p] for
 * o = 0..n-width (no reference counterpart: the reference recomputes each
 * key_fit sum over all states; each of its sums is a sum of W entries). */
void tor_windows_w(int n, int width, const double *x, double *W)
{
    uint64_t dim = (uint64_t)1 << n;
    const int bins = 1 << width;
    for (int q = 0; q < (n - width + 1) * bins; q++) W[q] = 0.;
    for (uint64_t s = 0; s < dim; s++)
        for (int o = 0; o + width <= n; o++) W[o * bins + (int)((s >> o) & (uint64_t)(bins - 1))] += x[s];
}

void tor_windows(int n, const double *x, double *W)
{
    tor_windows_w(n, 4, x, W);
}

double tor_mass(uint64_t dim, const double *x)
{
    double m = 0;
    for (uint64_t s = 0; s < dim; s++) m += x[s];
    return m;
}

double tor_ref_update4(uint64_t dim, double *x, const double *k1,
                       const double *k2, const double *k3, const double *k4)
{
    double chk = 0;
    for (uint64_t s = 0; s < dim; s++) {   /* masterEquation.c:639-642 */
        x[s] += (k1[s] + (2. * k2[s]) + (2. * k3[s]) + k4[s]) / 6.;
        chk += fabs((k1[s] + (2. * k2[s]) + (2. * k3[s]) + k4[s]) / 6.);
    }
    return chk;
}

void tor_dense(int n, const double *r, double *L)
{
    uint64_t dim = (uint64_t)1 << n;
    uint64_t col[66];
    double val[66];
    memset(L, 0, dim * dim * sizeof(double));
    for (uint64_t s = 0; s < dim; s++) {
        int c = emit_row(n, s, r, col, val);
        for (int j = 0; j < c; j++) L[s * dim + col[j]] = val[j];
    }
}
