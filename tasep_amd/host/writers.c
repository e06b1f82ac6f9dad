/*
 * writers.c -- the reference's output files, from device results.
 *
 * File names, formats and the order of writing follow
 * src/masterEquation.c:1019-1539 of the reference, including the name
 * lengths it passes to snprintf (so "<o>_mulExpVa" and
 * "<o>_high_order_point_fun" come out truncated exactly as the reference's).
 * Matrices arrive indexed by bit ([i*n + k]); the reference writes them with
 * both axes reversed (entry site first).
 *
 * The post-processing at the end (error guess, closures, point functions,
 * the high-order correlation, src/masterEquation.c:32-143, 1422-1539) sums
 * 4-site window marginals computed on the device instead of looping over
 * all 2^n states per value; the timed correlation of -a snapshots (4-site
 * keys whose sums span 6 sites) sums the 6-site window marginals each
 * snapshot keeps.
 */
#include "writers.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <sys/stat.h>

#include <omp.h>

/* ------------------------------------------------------------ snapshots */
/* Snapshots keep the moments and the 6-site window marginals (all the
 * timed correlation file needs); the whole state only for the snapshot
 * probabilities file. */
void snapshots_init(snapshots_t *s, int n, uint64_t dim, int res, unsigned int steps, int keep_x)
{
    memset(s, 0, sizeof *s);
    s->n = n;
    s->dim = dim;
    if (res <= 0) return;
    s->enabled = 1;
    s->cap = (uint64_t)steps / (uint64_t)res + 2;
    s->y = (double *)malloc(s->cap * (size_t)n * sizeof(double));
    s->cov = (double *)malloc(s->cap * (size_t)n * n * sizeof(double));
    s->mul = (double *)malloc(s->cap * (size_t)n * n * sizeof(double));
    if (keep_x) s->x = (double *)malloc(s->cap * dim * sizeof(double));
    if (n >= 6) s->w6 = (double *)malloc(s->cap * (size_t)(n - 5) * 64 * sizeof(double));
    if (!s->y || !s->cov || !s->mul || (keep_x && !s->x) || (n >= 6 && !s->w6)) {
        printf("ERROR: Not enough memory to allocate animation data array\n");
        exit(-1);
    }
}

void snapshots_take(snapshots_t *s, tasep_ctx *ctx)
{
    if (!s->enabled || s->count >= s->cap) return;
    const int n = s->n;
    const uint64_t c = s->count;
    int rc = tasep_moments(ctx, s->y + c * n, s->cov + c * n * n, s->mul + c * n * n, NULL);
    if (!rc && s->x) rc = tasep_get_x(ctx, s->x + c * s->dim);
    if (!rc && s->w6) rc = tasep_window_marginals_w(ctx, 6, s->w6 + c * (uint64_t)(n - 5) * 64);
    if (rc) {
        fprintf(stderr, "ERROR: snapshot: %s\n", tasep_strerror(rc));
        exit(-1);
    }
    s->count++;
}

void snapshots_free(snapshots_t *s)
{
    free(s->y);
    free(s->cov);
    free(s->mul);
    free(s->x);
    free(s->w6);
    memset(s, 0, sizeof *s);
}

/* ------------------------------------------------------------ PPM image */
typedef struct {
    unsigned char r, g, b;
} px_t;

static void ppm_write(const px_t *img, int w, int hgt, const char *path)
{
    FILE *f = fopen(path, "w");
    if (!f) {
        fprintf(stderr, "ERROR: cannot write %s\n", path);
        exit(-1);
    }
    fprintf(f, "P6\n%d %d\n255\n", w, hgt);
    for (int i = 0; i < w * hgt; i++) fwrite(&img[i], 1, 3, f);
    fclose(f);
}

/* Colour of one covariance entry (src/masterEquation.c:1169-1191): the
 * doubles are converted to int and then to an 8-bit channel. */
static px_t cov_pixel(double v)
{
    int r = 0, g = 0, b = 0;
    if (v < 0 && fabs(v) > 0.0001) {
        b = (int)((-1) * pow(fabs(v), 0.25) * 255);
    } else if (v > 0 && fabs(v) > 0.0001) {
        r = (int)(pow(v, 0.25) * 255);
    } else {
        g = (int)fabs(pow(fabs(v), 0.25) * 255);
    }
    px_t p = {(unsigned char)r, (unsigned char)g, (unsigned char)b};
    return p;
}

static px_t mul_pixel(double v)
{
    const int c = (int)((-1) * pow(fabs(v), 0.25) * 255);   /* :1334-1336 */
    px_t p = {(unsigned char)c, (unsigned char)c, (unsigned char)c};
    return p;
}

/* image of an n x n matrix given in written orientation M[i][k]; pixel
 * (i, k) lands at row k, column i (set_pixel(S, i, k, .)). */
static void matrix_image(px_t *img, int n, const double *M, px_t (*colour)(double))
{
    for (int i = 0; i < n; i++)
        for (int k = 0; k < n; k++) img[k * n + i] = colour(M[i * n + k]);
}

/* written orientation: W[i][k] = B[(n-1-i)*n + (n-1-k)] */
static void flip(int n, const double *B, double *W)
{
    for (int i = 0; i < n; i++)
        for (int k = 0; k < n; k++) W[i * n + k] = B[(n - 1 - i) * n + (n - 1 - k)];
}

static void write_matrix_rows(FILE *f, int n, const double *W)
{
    for (int i = 0; i < n; i++)
        for (int k = 0; k < n; k++) fprintf(f, k < n - 1 ? "%lf, " : "%lf\n", W[i * n + k]);
}

/* Directory for animation data; asks before reusing one
 * (src/masterEquation.c:1204-1214). */
static void make_anim_dir(const char *dir)
{
    if (mkdir(dir, 0777) == -1) {
        printf("Cauton: Directory <%s> already exists, Data may be overriten or lost! Continue? [Y/N]  ", dir);
        char sel = 'N';
        if (scanf("%c", &sel) == EOF) {
            fprintf(stderr, "ERROR: Issue while reading from stdin\n");
            exit(-1);
        }
        if (sel == 'N' || sel == 'n') exit(-1);
    }
}

/* Animation directory with the final image, one image per snapshot and all
 * snapshot matrices as text (src/masterEquation.c:1197-1279). */
static void write_anim_dir(const out_ctx_t *oc, const char *suffix, const px_t *final_img,
                           const double *mats, uint64_t count, px_t (*colour)(double), const char *msg)
{
    const int n = oc->n;
    const size_t L = strlen(oc->o);
    int size = (int)L + 17;
    char *dir = (char *)malloc((size_t)size);
    snprintf(dir, (size_t)size, "%s%s", oc->o, suffix);
    make_anim_dir(dir);
    size = (int)(strlen(dir) + L + 20);
    char *name = (char *)malloc((size_t)size);
    snprintf(name, (size_t)size, "%s/%s_fin.ppm", dir, oc->o);
    ppm_write(final_img, n, n, name);
    printf(msg, name);
    snprintf(name, (size_t)size, "%s/%s_all.txt", dir, oc->o);
    FILE *f = fopen(name, "w");
    if (f) {
        px_t *img = (px_t *)malloc((size_t)n * n * sizeof(px_t));
        double *W = (double *)malloc((size_t)n * n * sizeof(double));
        for (uint64_t j = 0; j < count; j++) {
            flip(n, mats + j * n * n, W);
            matrix_image(img, n, W, colour);
            const double time = oc->h * oc->anim_res * (double)j;
            snprintf(name, (size_t)size, "%s/%s_%lf.ppm", dir, oc->o, time);
            ppm_write(img, n, n, name);
            fprintf(f, "%d Matrix:\n", (int)j);
            write_matrix_rows(f, n, W);
            fprintf(f, "\n\n");
        }
        free(img);
        free(W);
        fclose(f);
    }
    free(name);
    free(dir);
}

/* ------------------------------------------------- high-order functions */
typedef struct {
    double lr, no, left, right;
} keysums_t;

/* key_fit sum (src/masterEquation.c:32-45) over states whose bits
 * off..off+size-1 equal key, from the window marginals
 * W[o*2^width + p] = P(bits o..o+width-1 = p) (size <= width) */
static double win_sum_w(const double *W, int n, int width, uint64_t key, int off, int size)
{
    const int w = off < n - width ? off : n - width;
    const int sh = off - w;
    const unsigned m = (1u << size) - 1, bins = 1u << width;
    double acc = 0;
    for (unsigned p = 0; p < bins; p++)
        if (((p >> sh) & m) == (unsigned)key) acc += W[(size_t)w * bins + p];
    return acc;
}

static double win_sum(const double *W, int n, uint64_t key, int off, int size)
{
    return win_sum_w(W, n, 4, key, off, size);
}

/* the four pattern marginals of src/masterEquation.c:61-68 from W; the
 * widest one spans ks + 2 sites */
static keysums_t win_sums_w(const double *W, int n, int width, uint64_t key, int ks, int off)
{
    keysums_t r;
    r.lr = win_sum_w(W, n, width, (key << 1) + 1 + ((uint64_t)1 << (ks + 1)), off - 1, ks + 2);
    r.no = win_sum_w(W, n, width, key, off, ks);
    r.left = win_sum_w(W, n, width, key + ((uint64_t)1 << ks), off, ks + 1);
    r.right = win_sum_w(W, n, width, (key << 1) + 1, off - 1, ks + 1);
    return r;
}

static keysums_t win_sums(const double *W, int n, uint64_t key, int ks, int off)
{
    return win_sums_w(W, n, 4, key, ks, off);
}

/* high_order_corr (src/masterEquation.c:47-75) as the normalised
 * correlation it computes */
static double corr_of(keysums_t k)
{
    const double ret = k.lr * k.no - k.left * k.right;
    return ret / sqrt(k.left * (k.no - k.left) * k.right * (k.no - k.right));
}

static double win_corr(const double *W, int n, uint64_t key, int ks, int off)
{
    return corr_of(win_sums(W, n, key, ks, off));
}

static double error_guess(const double *W, int n, uint64_t key, int ks, int off)
{
    const keysums_t k = win_sums(W, n, key, ks, off);
    return (k.lr / k.no - (k.left / k.no) * (k.right / k.no)) * k.no;
}

static double closure_value(const double *W, int n, uint64_t key, int ks, int off)
{
    const keysums_t k = win_sums(W, n, key, ks, off);
    return (k.left * k.right) / k.no;
}

static double point_func(const double *W, int n, uint64_t key, int ks, int off)
{
    return win_sum(W, n, key, off, ks);
}

/* ------------------------------------------------------------- writers */
/* The probabilities file (src/masterEquation.c:1021-1040): one line
 * "<bits, entry site first>\t\t%.20g" per state.  Lines are formatted in
 * parallel, a block of lines per thread, and written in order, so the file
 * is the same byte for byte (about 13 GB at n = 28). */
static void write_probabilities(FILE *f, int n, uint64_t dim, const double *x)
{
    enum { PER = 1 << 16 };                       /* lines per thread and round */
    const int nt = omp_get_max_threads() > 0 ? omp_get_max_threads() : 1;
    const size_t cap = (size_t)PER * (size_t)(n + 32);  /* n bits + 2 tabs + %.20g + \n */
    char **buf = (char **)malloc((size_t)nt * sizeof(char *));
    size_t *len = (size_t *)calloc((size_t)nt, sizeof(size_t));
    if (!buf || !len) {
        printf("ERROR: Not enough memory\n");
        exit(-1);
    }
    for (int t = 0; t < nt; t++)
        if (!(buf[t] = (char *)malloc(cap))) {
            printf("ERROR: Not enough memory\n");
            exit(-1);
        }
    for (uint64_t base = 0; base < dim; base += (uint64_t)nt * PER) {
#pragma omp parallel num_threads(nt)
        {
            const int t = omp_get_thread_num();
            char *p = buf[t];
            const uint64_t lo = base + (uint64_t)t * PER;
            const uint64_t hi = lo + PER < dim ? lo + PER : dim;
            for (uint64_t s = lo; s < hi; s++) {
                for (int j = 0; j < n; j++) *p++ = ((s >> (n - j - 1)) & 1) ? '1' : '0';
                p += sprintf(p, "\t\t%.20g\n", x[s]);
            }
            len[t] = lo < hi ? (size_t)(p - buf[t]) : 0;
        }
        for (int t = 0; t < nt; t++) fwrite(buf[t], 1, len[t], f);
    }
    for (int t = 0; t < nt; t++) free(buf[t]);
    free(buf);
    free(len);
}

void write_outputs(const out_ctx_t *oc, const double *x, const double *y, const double *cov,
                   const double *mul, const double *win, const snapshots_t *snaps)
{
    const int n = oc->n;
    const uint64_t dim = oc->dim;
    const char *o = oc->o;
    const size_t L = strlen(o);
    /* the reference sets this length only in the probabilities block; with
     * 'p' unselected it is uninitialised there, here it keeps that value */
    int size = (int)L + 14;
    FILE *f;

    if (oc->save_p) {
        f = fopen(o, "w");
        if (f) {
            printf("Writing result vector to file '%s'\n", o);
            write_probabilities(f, n, dim, x);
            printf("\n");
            fclose(f);
        }
        char name[size];
        snprintf(name, (size_t)size, "%s_prb.txt", o);
        f = fopen(name, "w");
        if (f) {
            for (uint64_t i = 0; snaps->enabled && i < snaps->count; i++) {
                fprintf(f, "%f, ", oc->h * oc->anim_res * (double)i);
                const double *xs = snaps->x + i * dim;
                for (uint64_t k = 0; k + 1 < dim; k++) fprintf(f, "%f, ", xs[k]);
                fprintf(f, "%f\n", xs[dim - 1]);
            }
            fclose(f);
        } else {
            printf("ERROR: Cant write to %s\n", name);
        }
    }
    if (oc->save_e) {
        const int sz = (int)L + 20;
        char name[sz];
        snprintf(name, (size_t)sz, "%s_exp_values_fin.vec", o);
        if ((f = fopen(name, "w"))) {
            for (int i = 0; i < n - 1; i++) fprintf(f, "%lf, ", y[n - 1 - i]);
            fprintf(f, "%lf", y[0]);
            fclose(f);
        }
        snprintf(name, (size_t)sz, "%s_exp_values_fin.dat", o);
        if ((f = fopen(name, "w"))) {
            for (int i = 0; i < n; i++) fprintf(f, "%d Pos%d %lf\n", i, i, y[n - 1 - i]);
            fclose(f);
        }
        if (oc->have_a) {
            snprintf(name, (size_t)sz, "%s_exp_timed.txt", o);
            if ((f = fopen(name, "w"))) {
                for (uint64_t i = 0; i < snaps->count; i++) {
                    const double *ys = snaps->y + i * n;
                    fprintf(f, "%f, ", oc->h * oc->anim_res * (double)i);
                    for (int k = 0; k < n - 1; k++) fprintf(f, "%lf, ", ys[n - 1 - k]);
                    fprintf(f, "%lf\n", ys[0]);
                }
                fclose(f);
            }
        }
    }
    double *W = (double *)malloc((size_t)n * n * sizeof(double));
    px_t *img = (px_t *)malloc((size_t)n * n * sizeof(px_t));
    if (oc->save_c) {
        size -= 5;
        char name[size];
        snprintf(name, (size_t)size, "%s_cov.mat", o);
        flip(n, cov, W);
        if ((f = fopen(name, "w"))) {
            write_matrix_rows(f, n, W);
            fclose(f);
        }
        size = (int)L + 9;
        char pname[size];
        snprintf(pname, (size_t)size, "%s_cov.ppm", o);
        matrix_image(img, n, W, cov_pixel);
        ppm_write(img, n, n, pname);
        if (oc->have_a) {
            printf("Creating directory for covariance animation data\n");
            write_anim_dir(oc, "_covariance_data", img, snaps->cov, snaps->count, cov_pixel,
                           "Wrote final covariance matrix ppm file to <%s>\n");
        }
    }
    if (oc->save_m) {
        size += 1;
        char name[size];
        snprintf(name, (size_t)size, "%s_mulExpVal.mat", o);
        flip(n, mul, W);
        if ((f = fopen(name, "w"))) {
            write_matrix_rows(f, n, W);
            fclose(f);
        }
        char pname[size];
        snprintf(pname, (size_t)size, "%s_mulExpVal.ppm", o);
        matrix_image(img, n, W, mul_pixel);
        ppm_write(img, n, n, pname);
        if (oc->have_a) {
            printf("Creating directory for mulitplicative expected values animation data\n");
            write_anim_dir(oc, "_multExpval_data", img, snaps->mul, snaps->count, mul_pixel,
                           "Wrote final multiplicative expected values matrix ppm file to <%s>\n");
        }
    }
    free(W);
    free(img);

    /* ---- high-order post-processing (src/masterEquation.c:1422-1539) */
    int ms = 2;
    printf("Calculating Higher Order Correlations\n");
    size = (int)L + 11;
    {
        char name[size];
        snprintf(name, (size_t)size, "%s_highcorr", o);
        if ((f = fopen(name, "w"))) {
            for (int i = 1; i < n - ms; i++) fprintf(f, "\t\t\t%d", i);
            fprintf(f, "\n");
            for (uint64_t mask = 0; mask < ((uint64_t)1 << ms); mask++) {
                for (int i = 0; i < ms; i++) fprintf(f, ((mask >> (ms - 1 - i)) & 1) ? "1" : "0");
                fprintf(f, "\t");
                for (int i = 1; i < n - ms; i++)
                    fprintf(f, "%1.4g\t", win_corr(win, n, mask, ms, n - ms - i));
                fprintf(f, "\n");
            }
            fclose(f);
        }
    }
    uint64_t mask = ((uint64_t)1 << ms) - 1;
    printf("Calculating error guess\n");
    size = (int)L + 13;
    {
        char name[size];
        snprintf(name, (size_t)size, "%s_error_guess", o);
        if ((f = fopen(name, "w"))) {
            for (int i = 1; i < n - ms; i++)
                fprintf(f, "Error guess for t_%d...t_%d:\t %lf\n", i - 1, i + ms,
                        error_guess(win, n, mask, ms, n - ms - i));
            fclose(f);
        }
    }
    printf("Calculating closures\n");
    size = (int)L + 10;
    {
        char name[size];
        snprintf(name, (size_t)size, "%s_closures", o);
        if ((f = fopen(name, "w"))) {
            for (int i = 1; i < n - ms; i++)
                fprintf(f, "Value for closure of t_%d...t_%d:\t %lf\n", i - 1, i + ms,
                        closure_value(win, n, mask, ms, n - ms - i));
            fclose(f);
        }
    }
    ms += 2;
    mask = ((uint64_t)1 << ms) - 1;
    size = (int)L + 22;
    {
        char name[size];
        snprintf(name, (size_t)size, "%s_high_order_point_func", o);
        if ((f = fopen(name, "w"))) {
            for (int i = 0; i < n - ms + 1; i++)
                fprintf(f, "t_%d...t_%d = %lf\n", i, ms + i - 1, point_func(win, n, mask, ms, n - ms - i));
            fclose(f);
        }
    }
    if (oc->have_a) {
        size = (int)L + 16;
        char name[size];
        snprintf(name, (size_t)size, "%s_highcorr_timed.txt", o);
        if ((f = fopen(name, "w"))) {
            for (uint64_t i = 0; i < snaps->count; i++) {
                fprintf(f, "%f ", oc->h * oc->anim_res * (double)i);
                /* 4-site key: sums over 6 sites, from the snapshot's
                 * 6-site window marginals */
                for (int k = 1; k < n - ms; k++)
                    fprintf(f, "%1.7g ", corr_of(win_sums_w(snaps->w6 + i * (uint64_t)(n - 5) * 64, n, 6, mask,
                                                            ms, n - ms - k)));
                fprintf(f, "\n");
            }
            fclose(f);
        }
    }
}
