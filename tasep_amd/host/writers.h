/* writers.h -- output files of the masterEquation host (formats as the
 * reference writes them, src/masterEquation.c:1019-1539). */
#ifndef TASEP_WRITERS_H
#define TASEP_WRITERS_H

#include <stdint.h>

#include "../../include/tasep_hip.h"

/* Snapshots taken with -a: densities, covariance and multiplicative
 * matrices (bit-indexed, as tasep_moments returns them) and the state. */
typedef struct {
    int n;
    uint64_t dim;
    int enabled;
    uint64_t count, cap;
    double *y;    /* count * n        */
    double *cov;  /* count * n * n    */
    double *mul;  /* count * n * n    */
    double *x;    /* count * dim, only when the probabilities are written */
    double *w6;   /* count * (n-5) * 64 six-site window marginals (n >= 6) */
} snapshots_t;

void snapshots_init(snapshots_t *s, int n, uint64_t dim, int res, unsigned int steps, int keep_x);
void snapshots_take(snapshots_t *s, tasep_ctx *ctx);
void snapshots_free(snapshots_t *s);

typedef struct {
    int n;
    uint64_t dim;
    double h;
    int anim_res;
    const char *o;
    int save_e, save_p, save_c, save_m, have_a;
} out_ctx_t;

/* Writes every output file of the reference for the final state: x (read
 * only for the probabilities file, may be NULL when it is not selected),
 * moments y (by bit), cov and mul (by bit, [i*n + k]) and the 4-site window
 * marginals win[o*16 + p] of tasep_window_marginals (NULL for n < 4). */
void write_outputs(const out_ctx_t *oc, const double *x, const double *y, const double *cov,
                   const double *mul, const double *win, const snapshots_t *snaps);

#endif
