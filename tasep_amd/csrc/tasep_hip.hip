// tasep_hip.hip -- C ABI (include/tasep_hip.h) over the gfx950 kernels.
//
// The context owns the state vector and work vectors in HBM, a run-state
// block, per-block partial buffers and one default-priority stream.  A run
// enqueues the step kernels in chunks; the early-exit decision is made on
// the device by the end-of-step launch (k_finish_step), so the host only
// reads one flag per chunk.  Lattices with n <= 12 run whole inside one
// workgroup's LDS (k_lds_run), one launch per up to 65536 steps.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <mutex>
#include <vector>

#include "../../include/tasep_hip.h"
#include "host_common.hpp"
#include "kernels.hpp"
#include "lds_launch.hpp"
#include "tile_kernels.hpp"

using namespace tasep;
using namespace tasep_host;

struct tasep_ctx {
    int n = 0, prec = 0, integ = 0, gen = 0;
    uint64_t dim = 0;
    size_t esize = 8;               // bytes per state entry
    hipStream_t stream = nullptr;
    void *buf[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    int cur = 0;                    // which of buf[0], buf[1] holds x
    double *rates = nullptr;        // 3 sets of kRateStride doubles
    bool have_rates = false;
    bool tv = false;
    int32_t *tv_type = nullptr;
    double *tv_par = nullptr;       // amp | period | offset | c, each kRateStride
    uint64_t *rowPtr = nullptr, *colInd = nullptr;
    double *val = nullptr;
    bool csr_built = false;
    RunState *st = nullptr;
    RunState *st_host = nullptr;    // pinned
    double *pchk = nullptr, *pmass = nullptr;
    double *mom_part = nullptr, *mom_out = nullptr;
    double host_rates[kRateStride] = {0};
    // options (tasep_set_option)
    int tiled = 1;              // tiled matrix-free passes: 0 never, 1 n >= 18 (19 REF), 2 where they fit
    bool graphs = false;        // long runs replay a captured graph of steps
    int xcd_order = 4;          // tile order of the tiled passes (tile_walk; 4: level walk)
    uint32_t *walk = nullptr;   // level-walk table of the context's tiles (built on first use)
    int walk_tb = -1;           // tile-index bits the table was built for
    bool force_staged = false;  // staged RK4 even for constant rates
    int tile_blocks = 0;        // persistent grid of the tiled passes (0: by occupancy)
    int lds = 1;                // whole runs in one workgroup's LDS: 0 never, 1 n <= 12, 2 where it fits
    int lds_threads = 0;        // workgroup size of the LDS-resident kernel (0: automatic)
    int store = 0;              // output stores of the tiled passes: 0 by size, 1 non-temporal, 2 ordinary
    uint64_t steps_before = 0;  // cumulative step count when the current run call began
    bool window = false;        // exit test armed: reductions run per window of steps
    uint32_t slot = 0;          // partial-sum slot of the step being enqueued (window mode)
    uint32_t pending = 0;       // steps whose end-of-step reduction is deferred
    uint32_t pending_blocks = 0;
    double *wtot = nullptr;     // per-step totals of a window (2 * kWindow doubles)
    LdsInstance *lds_inst = nullptr;       // device: the instance table of one entry
    LdsInstance *lds_inst_host = nullptr;  // pinned
};

// --------------------------------------------------------------- utilities
static int ensure_buf(tasep_ctx *c, int i) {
    if (c->buf[i]) return TASEP_OK;
    HIP_OK(hipMalloc(&c->buf[i], c->dim * c->esize));
    return TASEP_OK;
}

static void *xcur(tasep_ctx *c) { return c->buf[c->cur]; }

template <class V>
static void launch_fill(tasep_ctx *c, void *p, double v) {
    hipLaunchKernelGGL(k_fill<V>, dim3(grid_for(c->dim)), dim3(kBlock), 0, c->stream,
                       (V *)p, c->dim, v);
}

extern "C" {

const char *tasep_strerror(int code) {
    switch (code) {
    case TASEP_OK: return "ok";
    case TASEP_ERR_ARG: return g_errmsg[0] ? g_errmsg : "invalid argument";
    case TASEP_ERR_NOMEM: return g_errmsg[0] ? g_errmsg : "out of memory";
    case TASEP_ERR_HIP: return g_errmsg[0] ? g_errmsg : "HIP runtime error";
    case TASEP_ERR_STATE: return g_errmsg[0] ? g_errmsg : "call out of order";
    case TASEP_ERR_NODEV: return g_errmsg[0] ? g_errmsg : "no gfx950 device";
    default: return "unknown error";
    }
}

uint64_t tasep_csr_nnz(int n) {
    if (n < 2 || n > 40) return 0;
    return 2 * ((uint64_t)1 << n) + ((uint64_t)1 << (n - 2)) * (uint64_t)(n - 1);
}

int tasep_ctx_destroy(tasep_ctx *c) {
    if (!c) return TASEP_OK;
    if (c->stream) hipStreamSynchronize(c->stream);
    for (void *&p : c->buf) if (p) hipFree(p);
    hipFree(c->rates);
    hipFree(c->tv_type);
    hipFree(c->tv_par);
    hipFree(c->rowPtr);
    hipFree(c->colInd);
    hipFree(c->val);
    hipFree(c->st);
    if (c->st_host) hipHostFree(c->st_host);
    hipFree(c->pchk);
    hipFree(c->pmass);
    hipFree(c->wtot);
    hipFree(c->mom_part);
    hipFree(c->mom_out);
    hipFree(c->lds_inst);
    hipFree(c->walk);
    if (c->lds_inst_host) hipHostFree(c->lds_inst_host);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
    return TASEP_OK;
}

int tasep_ctx_create(tasep_ctx **out, int n, int precision, int integrator, int generator) {
    if (!out) return fail(TASEP_ERR_ARG, "ctx pointer is NULL");
    *out = nullptr;
    if (n < 2 || n > kMaxN) return fail(TASEP_ERR_ARG, "n must be in [2, 34]");
    if (precision != TASEP_F64 && precision != TASEP_F32)
        return fail(TASEP_ERR_ARG, "unknown precision");
    if (integrator != TASEP_REF && integrator != TASEP_RK4)
        return fail(TASEP_ERR_ARG, "unknown integrator");
    if (generator != TASEP_MF && generator != TASEP_CSR)
        return fail(TASEP_ERR_ARG, "unknown generator");
    if (integrator == TASEP_REF && precision != TASEP_F64)
        return fail(TASEP_ERR_ARG, "the reference integrator runs in FP64 only");
    if (generator == TASEP_CSR && (precision != TASEP_F64 || n > 30))
        return fail(TASEP_ERR_ARG, "CSR generator needs FP64 and n <= 30");
    int rc = check_device();
    if (rc) return rc;
    tasep_ctx *c = new tasep_ctx;
    c->n = n;
    c->prec = precision;
    c->integ = integrator;
    c->gen = generator;
    c->dim = (uint64_t)1 << n;
    c->esize = precision == TASEP_F32 ? 4 : 8;
#define CK(expr)                       \
    do {                               \
        int r_ = (expr);               \
        if (r_) {                      \
            tasep_ctx_destroy(c);      \
            return r_;                 \
        }                              \
    } while (0)
    auto hipck = [&](hipError_t e, const char *what) {
        return e == hipSuccess ? TASEP_OK
                               : fail(e == hipErrorOutOfMemory ? TASEP_ERR_NOMEM : TASEP_ERR_HIP, what, e);
    };
    CK(hipck(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking), "hipStreamCreate"));
    CK(hipck(hipMalloc(&c->rates, 3 * kRateStride * sizeof(double)), "hipMalloc(rates)"));
    CK(hipck(hipMalloc(&c->st, sizeof(RunState)), "hipMalloc(state)"));
    CK(hipck(hipHostMalloc((void **)&c->st_host, sizeof(RunState), hipHostMallocDefault),
             "hipHostMalloc(state)"));
    CK(hipck(hipMalloc(&c->pchk, (size_t)kWindow * kMaxBlocks * sizeof(double)), "hipMalloc(partials)"));
    CK(hipck(hipMalloc(&c->pmass, (size_t)kWindow * kMaxBlocks * sizeof(double)), "hipMalloc(partials)"));
    CK(hipck(hipMalloc(&c->wtot, (size_t)2 * kWindow * sizeof(double)), "hipMalloc(window)"));
    CK(hipck(hipMalloc(&c->lds_inst, sizeof(LdsInstance)), "hipMalloc(instance)"));
    CK(hipck(hipHostMalloc((void **)&c->lds_inst_host, sizeof(LdsInstance), hipHostMallocDefault),
             "hipHostMalloc(instance)"));
    CK(ensure_buf(c, 0));
    memset(c->st_host, 0, sizeof(RunState));
    CK(hipck(hipMemcpyAsync(c->st, c->st_host, sizeof(RunState), hipMemcpyHostToDevice, c->stream),
             "hipMemcpy(state)"));
    if (c->prec == TASEP_F32) launch_fill<float>(c, c->buf[0], 1.0 / (double)c->dim);
    else launch_fill<double>(c, c->buf[0], 1.0 / (double)c->dim);
    CK(hipck(hipStreamSynchronize(c->stream), "init"));
#undef CK
    *out = c;
    return TASEP_OK;
}

int tasep_set_rates(tasep_ctx *c, const double *r) {
    if (!c || !r) return fail(TASEP_ERR_ARG, "NULL argument");
    for (int i = 0; i <= c->n; i++) {
        if (!(r[i] >= 0.0) || !isfinite(r[i])) return fail(TASEP_ERR_ARG, "rates must be finite and >= 0");
        c->host_rates[i] = r[i];
    }
    for (int s = 0; s < 3; s++)
        HIP_OK(hipMemcpyAsync(c->rates + s * kRateStride, c->host_rates, kRateStride * sizeof(double),
                              hipMemcpyHostToDevice, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    c->have_rates = true;
    c->tv = false;
    c->csr_built = false;
    return TASEP_OK;
}

int tasep_set_time_variant(tasep_ctx *c, const int32_t *type, const double *amp,
                           const double *period, const double *offset, const double *cst) {
    if (!c || !type || !amp || !period || !offset || !cst) return fail(TASEP_ERR_ARG, "NULL argument");
    if (c->gen != TASEP_MF) return fail(TASEP_ERR_ARG, "time-variant rates need the matrix-free generator");
    for (int i = 0; i <= c->n; i++)
        if (type[i] < 0 || type[i] > 2) return fail(TASEP_ERR_ARG, "function type must be 0, 1 or 2");
    if (!c->tv_type) {
        HIP_OK(hipMalloc(&c->tv_type, kRateStride * sizeof(int32_t)));
        HIP_OK(hipMalloc(&c->tv_par, 4 * kRateStride * sizeof(double)));
    }
    const size_t nb = (size_t)(c->n + 1) * sizeof(double);
    HIP_OK(hipMemcpyAsync(c->tv_type, type, (size_t)(c->n + 1) * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
    HIP_OK(hipMemcpyAsync(c->tv_par + 0 * kRateStride, amp, nb, hipMemcpyHostToDevice, c->stream));
    HIP_OK(hipMemcpyAsync(c->tv_par + 1 * kRateStride, period, nb, hipMemcpyHostToDevice, c->stream));
    HIP_OK(hipMemcpyAsync(c->tv_par + 2 * kRateStride, offset, nb, hipMemcpyHostToDevice, c->stream));
    HIP_OK(hipMemcpyAsync(c->tv_par + 3 * kRateStride, cst, nb, hipMemcpyHostToDevice, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    c->tv = true;
    c->have_rates = true;
    return TASEP_OK;
}

int tasep_set_x0_uniform(tasep_ctx *c) {
    if (!c) return fail(TASEP_ERR_ARG, "NULL context");
    c->cur = 0;
    if (c->prec == TASEP_F32) launch_fill<float>(c, xcur(c), 1.0 / (double)c->dim);
    else launch_fill<double>(c, xcur(c), 1.0 / (double)c->dim);
    HIP_OK(hipStreamSynchronize(c->stream));
    return TASEP_OK;
}

int tasep_set_x0_state(tasep_ctx *c, uint64_t s) {
    if (!c) return fail(TASEP_ERR_ARG, "NULL context");
    if (s >= c->dim) return fail(TASEP_ERR_ARG, "state index out of range");
    c->cur = 0;
    HIP_OK(hipMemsetAsync(xcur(c), 0, c->dim * c->esize, c->stream));
    if (c->prec == TASEP_F32) {
        float one = 1.0f;
        HIP_OK(hipMemcpyAsync((float *)xcur(c) + s, &one, 4, hipMemcpyHostToDevice, c->stream));
    } else {
        double one = 1.0;
        HIP_OK(hipMemcpyAsync((double *)xcur(c) + s, &one, 8, hipMemcpyHostToDevice, c->stream));
    }
    HIP_OK(hipStreamSynchronize(c->stream));
    return TASEP_OK;
}

int tasep_set_x(tasep_ctx *c, const double *x) {
    if (!c || !x) return fail(TASEP_ERR_ARG, "NULL argument");
    c->cur = 0;
    if (c->prec == TASEP_F64) {
        HIP_OK(hipMemcpyAsync(xcur(c), x, c->dim * 8, hipMemcpyHostToDevice, c->stream));
    } else {
        double *tmp = nullptr;
        HIP_OK(hipMallocAsync((void **)&tmp, c->dim * 8, c->stream));
        HIP_OK(hipMemcpyAsync(tmp, x, c->dim * 8, hipMemcpyHostToDevice, c->stream));
        hipLaunchKernelGGL((k_convert<double, float>), dim3(grid_for(c->dim)), dim3(kBlock), 0, c->stream,
                           tmp, (float *)xcur(c), c->dim);
        HIP_OK(hipFreeAsync(tmp, c->stream));
    }
    HIP_OK(hipStreamSynchronize(c->stream));
    return TASEP_OK;
}

int tasep_get_x(tasep_ctx *c, double *x) {
    if (!c || !x) return fail(TASEP_ERR_ARG, "NULL argument");
    if (c->prec == TASEP_F64) {
        HIP_OK(hipMemcpyAsync(x, xcur(c), c->dim * 8, hipMemcpyDeviceToHost, c->stream));
    } else {
        double *tmp = nullptr;
        HIP_OK(hipMallocAsync((void **)&tmp, c->dim * 8, c->stream));
        hipLaunchKernelGGL((k_convert<float, double>), dim3(grid_for(c->dim)), dim3(kBlock), 0, c->stream,
                           (const float *)xcur(c), tmp, c->dim);
        HIP_OK(hipMemcpyAsync(x, tmp, c->dim * 8, hipMemcpyDeviceToHost, c->stream));
        HIP_OK(hipFreeAsync(tmp, c->stream));
    }
    HIP_OK(hipStreamSynchronize(c->stream));
    return TASEP_OK;
}

int tasep_set_x0_product(tasep_ctx *c, const double *p) {
    if (!c || !p) return fail(TASEP_ERR_ARG, "NULL argument");
    ProductArgs a;
    memset(&a, 0, sizeof a);
    for (int k = 0; k < c->n; k++) {
        if (!(p[k] >= 0.0 && p[k] <= 1.0)) return fail(TASEP_ERR_ARG, "site densities must lie in [0, 1]");
        a.p[k] = p[k];
    }
    c->cur = 0;
    if (c->prec == TASEP_F32)
        hipLaunchKernelGGL(k_product_state<float>, dim3(grid_for(c->dim)), dim3(kBlock), 0, c->stream,
                           (float *)xcur(c), c->dim, c->n, a);
    else
        hipLaunchKernelGGL(k_product_state<double>, dim3(grid_for(c->dim)), dim3(kBlock), 0, c->stream,
                           (double *)xcur(c), c->dim, c->n, a);
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(c->stream));
    return TASEP_OK;
}

int tasep_get_x_range(tasep_ctx *c, uint64_t offset, uint64_t count, double *x) {
    if (!c || !x) return fail(TASEP_ERR_ARG, "NULL argument");
    if (offset > c->dim || count > c->dim - offset) return fail(TASEP_ERR_ARG, "range outside the state vector");
    if (!count) return TASEP_OK;
    if (c->prec == TASEP_F64) {
        HIP_OK(hipMemcpyAsync(x, (const double *)xcur(c) + offset, count * 8, hipMemcpyDeviceToHost, c->stream));
    } else {
        double *tmp = nullptr;
        HIP_OK(hipMallocAsync((void **)&tmp, count * 8, c->stream));
        hipLaunchKernelGGL((k_convert<float, double>), dim3(grid_for(count)), dim3(kBlock), 0, c->stream,
                           (const float *)xcur(c) + offset, tmp, count);
        HIP_OK(hipMemcpyAsync(x, tmp, count * 8, hipMemcpyDeviceToHost, c->stream));
        HIP_OK(hipFreeAsync(tmp, c->stream));
    }
    HIP_OK(hipStreamSynchronize(c->stream));
    return TASEP_OK;
}

int tasep_build_csr(tasep_ctx *c) {
    if (!c) return fail(TASEP_ERR_ARG, "NULL context");
    if (c->gen != TASEP_CSR) return fail(TASEP_ERR_ARG, "context was not created with TASEP_CSR");
    if (!c->have_rates || c->tv) return fail(TASEP_ERR_STATE, "set constant rates before building the CSR");
    const uint64_t nnz = tasep_csr_nnz(c->n);
    if (!c->rowPtr) {
        HIP_OK(hipMalloc(&c->rowPtr, (c->dim + 1) * sizeof(uint64_t)));
        HIP_OK(hipMalloc(&c->colInd, nnz * sizeof(uint64_t)));
        HIP_OK(hipMalloc(&c->val, nnz * sizeof(double)));
    }
    hipLaunchKernelGGL(k_csr_build, dim3(grid_for(c->dim, 1 << 16)), dim3(kBlock), 0, c->stream, c->n,
                       c->rates, c->rowPtr, c->colInd, c->val, c->dim);
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(c->stream));
    c->csr_built = true;
    return TASEP_OK;
}

int tasep_get_csr(tasep_ctx *c, uint64_t *rowPtr, uint64_t *colInd, double *val) {
    if (!c || !rowPtr || !colInd || !val) return fail(TASEP_ERR_ARG, "NULL argument");
    if (!c->csr_built) return fail(TASEP_ERR_STATE, "CSR not built");
    const uint64_t nnz = tasep_csr_nnz(c->n);
    HIP_OK(hipMemcpyAsync(rowPtr, c->rowPtr, (c->dim + 1) * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipMemcpyAsync(colInd, c->colInd, nnz * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipMemcpyAsync(val, c->val, nnz * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    return TASEP_OK;
}

int tasep_gencrs(uint64_t *rowPtr, double *valArr, uint64_t *colInd, const double *r, uint64_t dim,
                 int n) {
    if (!rowPtr || !valArr || !colInd || !r) return fail(TASEP_ERR_ARG, "NULL argument");
    if (n < 2 || n > 30 || dim != ((uint64_t)1 << n)) return fail(TASEP_ERR_ARG, "dim must be 2^n, 2 <= n <= 30");
    tasep_ctx *c = nullptr;
    int rc = tasep_ctx_create(&c, n, TASEP_F64, TASEP_REF, TASEP_CSR);
    if (rc) return rc;
    if (!(rc = tasep_set_rates(c, r)) && !(rc = tasep_build_csr(c)))
        rc = tasep_get_csr(c, rowPtr, colInd, valArr);
    tasep_ctx_destroy(c);
    return rc;
}

}  // extern "C"

// ----------------------------------------------------------------- the loop
template <class V>
static PassArgs<V> pass_base(tasep_ctx *c) {
    PassArgs<V> a;
    memset(&a, 0, sizeof a);
    a.rates = c->rates;
    a.rowPtr = c->rowPtr;
    a.colInd = c->colInd;
    a.val = c->val;
    a.st = c->st;
    a.pchk = c->pchk + (size_t)c->slot * kMaxBlocks;
    a.pmass = c->pmass + (size_t)c->slot * kMaxBlocks;
    a.dim = c->dim;
    a.g = make_lattice(c->n);
    return a;
}

// Optional per-launch timing (tasep_step_profile): events around each launch.
struct LaunchProf {
    bool on = false;
    int count = 0;
    hipEvent_t ev[2 * 8];
    const char *name[8];
    double bytes[8];
};
static thread_local LaunchProf *g_prof = nullptr;

static void prof_begin(tasep_ctx *c) {
    if (g_prof && g_prof->count < 8) hipEventRecord(g_prof->ev[2 * g_prof->count], c->stream);
}
static void prof_end(tasep_ctx *c, const char *name, double bytes_per_state) {
    if (g_prof && g_prof->count < 8) {
        hipEventRecord(g_prof->ev[2 * g_prof->count + 1], c->stream);
        g_prof->name[g_prof->count] = name;
        g_prof->bytes[g_prof->count] = bytes_per_state;
        g_prof->count++;
    }
}

// End-of-step reduction of the final kernel's nblocks partials.  With the
// exit test armed the reductions run per window of steps (the run loop
// launches them); runs without one defer the reduction to the end of the
// run (one launch then accounts for all pending steps).  The per-launch
// profile reduces every step.
static void launch_finish(tasep_ctx *c, uint32_t nblocks) {
    if (!g_prof) {
        c->pending_blocks = nblocks;
        if (!c->window) c->pending += 1;
        return;
    }
    prof_begin(c);
    hipLaunchKernelGGL(k_finish_step, dim3(1), dim3(kFinishThreads), 0, c->stream, c->st, c->pchk, c->pmass, nblocks,
                       1u);
    prof_end(c, "finish_step", 0.0);
}

static void flush_finish(tasep_ctx *c) {
    if (!c->pending) return;
    hipLaunchKernelGGL(k_finish_step, dim3(1), dim3(kFinishThreads), 0, c->stream, c->st, c->pchk, c->pmass,
                       c->pending_blocks, c->pending);
    c->pending = 0;
}

template <class V, bool CSR, bool TWO, int EPI, bool SELF = false>
static void launch_pass(tasep_ctx *c, const PassArgs<V> &a, const char *name, double bytes) {
    prof_begin(c);
    hipLaunchKernelGGL((k_rk_pass<V, CSR, TWO, EPI, SELF>), dim3(grid_for(c->dim)), dim3(kBlock), 0,
                       c->stream, a);
    prof_end(c, name, bytes);
    if (EPI != EPI_STORE) launch_finish(c, (uint32_t)grid_for(c->dim));
}

// One classical RK4 step in Horner form (constant rates; x updated in place).
template <class V, bool CSR>
static void rk4_horner_step(tasep_ctx *c, double h) {
    V *x = (V *)xcur(c);
    V *u = (V *)c->buf[2];
    V *v = (V *)c->buf[3];
    PassArgs<V> a = pass_base<V>(c);
    a.x = x;
    a.use_self = 1;
    const double e = (double)sizeof(V);
    const double csr = CSR ? 8.0 + 16.0 * (2.0 + (c->n - 1) / 4.0) : 0.0;
    // P1: u = x + h/4 L x          (read x, write u)
    a.gx = x; a.out = u; a.c = 0.25 * h;
    launch_pass<V, CSR, false, EPI_STORE, true>(c, a, "horner_p1", 2 * e + csr);
    // P2: v = x + h/3 L u          (read u, x, write v)
    a.gx = u; a.out = v; a.c = h / 3.0;
    launch_pass<V, CSR, false, EPI_STORE>(c, a, "horner_p23", 3 * e + csr);
    // P3: u = x + h/2 L v
    a.gx = v; a.out = u; a.c = 0.5 * h;
    launch_pass<V, CSR, false, EPI_STORE>(c, a, "horner_p23", 3 * e + csr);
    // P4: x = x + h L u            (read u, x, write x)
    a.gx = u; a.out = x; a.c = h;
    launch_pass<V, CSR, false, EPI_HORNER>(c, a, "horner_p4", 3 * e + csr);
}

// One classical RK4 step, staged (time-variant rates): x -> xn (ping-pong).
template <class V, bool CSR>
static void rk4_staged_step(tasep_ctx *c, double h, uint64_t j) {
    V *x = (V *)c->buf[c->cur ^ (int)(j & 1)];
    V *xn = (V *)c->buf[c->cur ^ (int)((j + 1) & 1)];
    V *k1 = (V *)c->buf[2], *k2 = (V *)c->buf[3], *k3 = (V *)c->buf[4];
    PassArgs<V> a = pass_base<V>(c);
    a.x = x;
    a.gx = x;
    a.c = h;
    a.use_self = 0;
    const double e = (double)sizeof(V);
    a.rset = 0; a.out = k1;
    launch_pass<V, CSR, false, EPI_STORE>(c, a, "rk4_s1", 2 * e);
    a.rset = 1; a.gk = k1; a.ck = 0.5; a.out = k2;
    launch_pass<V, CSR, true, EPI_STORE>(c, a, "rk4_s23", 3 * e);
    a.rset = 1; a.gk = k2; a.ck = 0.5; a.out = k3;
    launch_pass<V, CSR, true, EPI_STORE>(c, a, "rk4_s23", 3 * e);
    a.rset = 2; a.gk = k3; a.ck = 1.0; a.out = xn; a.k1 = k1; a.k2 = k2; a.k3 = k3;
    launch_pass<V, CSR, true, EPI_RK4>(c, a, "rk4_s4", 5 * e);
}

// ---- tiled matrix-free passes (n >= b + 1) ----
// Output stores of the tiled passes.  Ordinary stores when one state
// vector takes 16-128 MiB: the memory-side cache (256 MB) then still holds
// a pass's output when the next pass reads it.  Non-temporal otherwise:
// smaller vectors stay in the L2s anyway (ordinary stores leave them dirty
// there for the end-of-kernel write-back), larger ones would only evict the
// gathered vector.  Measured per step, cached / non-temporal, n = 19..28:
// reference loop FP64 1.08 1.21 | 0.97 0.96 0.97 0.97 | 1.08 1.03 1.03 1.02,
// RK4 FP32 1.01 1.05 1.04 | 0.97 1.00 0.98 0.99 | 1.07 1.04 1.04,
// RK4 FP64 1.05 1.06 | 0.96 0.98 0.99 0.98 | 1.06 1.04 1.03 1.05.
static int nt_out(const tasep_ctx *c) {
    if (c->store) return c->store == 1;
    const uint64_t bytes = c->dim * c->esize;
    return bytes < ((uint64_t)16 << 20) || bytes > ((uint64_t)128 << 20);
}

template <class V>
static TileArgs<V> tile_base(tasep_ctx *c) {
    TileArgs<V> a;
    memset(&a, 0, sizeof a);
    a.rates = c->rates;
    a.st = c->st;
    a.pchk = c->pchk + (size_t)c->slot * kMaxBlocks;
    a.pmass = c->pmass + (size_t)c->slot * kMaxBlocks;
    a.n = c->n;
    a.ntiles = c->dim >> TileGeom<V>::b;
    return a;
}

// Workgroups of a tiled kernel resident on the whole device at once
// (occupancy x CUs), 0 if unknown.
static int resident_slots(const void *kernel) {
    int per_cu = 0, dev = 0, cus = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kTileThreads, 0) != hipSuccess ||
        hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        per_cu < 1 || cus < 1)
        return 0;
    return per_cu * cus;
}

// Compute units of the current device (cached per device).
static int device_cus() {
    static std::atomic<int> cache[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 0;
    int v = cache[dev].load(std::memory_order_relaxed);
    if (!v && hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && v > 0)
        cache[dev].store(v, std::memory_order_relaxed);
    return v > 0 ? v : 0;
}

// Grid and tile order of a tiled pass.  Grid, by the number of tiles (the
// kernels run one 512-lane workgroup per CU at a time, 256 resident slots):
//   up to 2^15 tiles (2^16 for the reference loop)  one workgroup per slot
//   above that                                      4 per slot (1024)
//   RK4 passes (`wide`) over 2^17 tiles             4096
//   RK4 passes over 2^18 tiles or more              kMaxBlocks (16384)
// always at most ntiles/8 but at least one per CU, rounded down to a
// multiple of the 8 XCDs, so that where tiles are few (mid n) each
// workgroup still walks several and the next tile's reads overlap the
// current tile's LDS phase (reference loop FP64 n=22 / 23 / 24 24.3 / 44.9 /
// 89.3 us per step against 30.2 / 59.5 / 116.0 with one tile per
// workgroup).  Measured per step:
//   256 vs 1024      RK4 FP32 n=26 / 27 / 28 0.654 / 1.345 / 2.820 ms against
//                    0.674 / 1.364 / 2.830, n=29 6.08 against 6.03, n=30
//                    13.38 against 13.05; FP64 n=26 1.32 against 1.34, n=28
//                    5.97 against 5.92; reference loop n=26 / 27 / 28 0.319 /
//                    0.670 / 1.391 against 0.329 / 0.678 / 1.395, n=30 6.33
//                    against 6.25
//   1024 / 2048 /    RK4 FP32 n=30 13.14 / 13.06 / 12.99 / 13.08 ms, FP64 n=29
//   4096 / 16384     13.07 / 12.98 / 12.89 / 12.94 ms; at 2^16 tiles 1024 and
//                    2048 tie; FP32 n=32 60.2 ms with 16384 against 61.7 with
//                    4096
// Order (tile_walk): the requested one where the grid allows it, else the
// next simpler one.
// The level-walk table the context's passes will use, built ahead of a
// graph capture: its host-to-device copy cannot be captured.  (Built for
// the tile geometry of the context's precision, or the 64-row groups of the
// stored-matrix loop; a table the grid then cannot use is only memory.)
static void prewarm_walk(tasep_ctx *c);

struct TileLaunch {
    uint64_t grid;
    int order;                  // order the kernel walks its positions in (0..3)
    uint32_t res;
    const uint32_t *walk;       // level-walk table over the positions of order 2, or null
};

// Level-walk table for tiles with tb index bits: walk[p] is the tile at
// position p.  Bit 0 of the tile index varies fastest; the bits above it
// are read as particles on tb - 1 sites and ordered by their count, then by
// the sum of their positions, ascending within odd sums and descending
// within even ones.  A hop across a bond between two of those bits keeps
// the count and raises the position sum by one, so the tile read across
// such a bond (the configuration before the hop) comes in the next group
// of the walk.  In an LRU model of one XCD's L2 replaying a pass's reads
// the gathered vector is fetched 2.78 instead of 3.54 times per pass at
// n = 28 (FP32 tiles), 3.08 instead of 3.81 (FP64), 3.95 instead of 4.62
// at n = 32.
static void level_walk(int tb, std::vector<uint32_t> &walk) {
    const int F = tb - 1;
    const uint32_t np = (uint32_t)1 << F;
    const int nb = (F + 1) * (F * (F - 1) / 2 + 1);               // (count, position sum) groups
    auto group = [F](uint32_t p) {
        int m = 0, sum = 0;
        for (int i = 0; i < F; i++)
            if ((p >> i) & 1) { m++; sum += i; }
        return m * (F * (F - 1) / 2 + 1) + sum;
    };
    std::vector<uint32_t> start((size_t)nb + 1, 0);
    for (uint32_t p = 0; p < np; p++) start[group(p) + 1]++;
    for (int b = 0; b < nb; b++) start[b + 1] += start[b];
    std::vector<uint32_t> fill(start.begin(), start.end() - 1), pats(np);
    for (uint32_t p = 0; p < np; p++) pats[fill[group(p)]++] = p;  // ascending within a group
    const int S = F * (F - 1) / 2 + 1;
    for (int b = 0; b < nb; b++)
        if ((b % S) % 2 == 0) std::reverse(pats.begin() + start[b], pats.begin() + start[b + 1]);
    walk.resize((size_t)np * 2);
    for (uint32_t i = 0; i < np; i++) {
        walk[2 * i] = pats[i] << 1;
        walk[2 * i + 1] = (pats[i] << 1) | 1;
    }
}

static const uint32_t *ensure_walk(tasep_ctx *c, int tb) {
    if (c->walk_tb == tb) return c->walk;
    // a table missing under a graph capture (its allocation failed before
    // the capture) is not built there: the allocation and copy would end
    // the capture; the passes take the order the table would permute
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(c->stream, &cs) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    if (cs != hipStreamCaptureStatusNone) return nullptr;
    std::vector<uint32_t> w;
    level_walk(tb, w);
    // without the table the passes take the order it permutes; clear the
    // failed call's error so that it does not fail the run that follows
    uint32_t *d = nullptr;
    if (hipMalloc(&d, w.size() * sizeof(uint32_t)) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    if (hipMemcpy(d, w.data(), w.size() * sizeof(uint32_t), hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipGetLastError();
        hipFree(d);
        return nullptr;
    }
    hipFree(c->walk);
    c->walk = d;
    c->walk_tb = tb;
    return d;
}
// slots_cache: the kernel's resident workgroups, looked up once (contexts on
// other threads may launch the same kernel at the same time)
static TileLaunch tile_launch(tasep_ctx *c, const void *kernel, std::atomic<int> &slots_cache, uint64_t ntiles,
                              bool wide = false) {
    int slots = slots_cache.load(std::memory_order_relaxed);
    if (!slots) {
        slots = resident_slots(kernel);
        slots_cache.store(slots, std::memory_order_relaxed);
    }
    uint64_t cap = slots > 0 ? ((uint64_t)slots * 4) & ~(uint64_t)7 : 1024;
    if (slots >= 8 && ntiles <= ((uint64_t)1 << (wide ? 15 : 16))) cap = (uint64_t)slots & ~(uint64_t)7;
    if (cap < 8) cap = 8;
    if (wide && ntiles >= ((uint64_t)1 << 17)) cap = ntiles >= ((uint64_t)1 << 18) ? kMaxBlocks : 4096;
    if (cap > (uint64_t)kMaxBlocks) cap = kMaxBlocks;
    const int cus = device_cus();
    uint64_t few = ntiles / 8;
    if (few < (uint64_t)cus) few = (uint64_t)cus;
    few &= ~(uint64_t)7;
    if (few >= 8 && cap > few) cap = few;
    if (c->tile_blocks > 0) cap = (uint64_t)c->tile_blocks;
    TileLaunch L;
    L.grid = ntiles < cap ? ntiles : cap;
    L.order = c->xcd_order;
    L.walk = nullptr;
    if (L.order == 4) {
        // the table permutes the positions of order 2
        L.res = slots >= 8 ? (uint32_t)(slots / 8) : 1;
        if ((L.grid % 8) == 0 && ((L.grid / 8) % L.res) == 0 && (ntiles % L.grid) == 0 && ntiles >= 2) {
            L.walk = ensure_walk(c, 63 - __builtin_clzll(ntiles));
            if (L.walk) {
                L.order = 2;
                return L;
            }
        }
        L.order = 2;
    }
    if (L.order == 3) {
        L.res = slots > 0 ? (uint32_t)slots : 1;
        if ((L.grid % L.res) == 0 && (ntiles % L.grid) == 0) return L;
        L.order = 2;
    }
    L.res = slots >= 8 ? (uint32_t)(slots / 8) : 1;
    if (L.order == 2 && !((L.grid % 8) == 0 && ((L.grid / 8) % L.res) == 0 && (ntiles % L.grid) == 0)) L.order = 1;
    if (L.order == 1 && !((ntiles % 8) == 0 && (L.grid % 8) == 0)) L.order = 0;
    return L;
}

template <class V, int EPI, bool SELF, bool NT>
static uint32_t launch_tile_nt(tasep_ctx *c, TileArgs<V> a) {
    static std::atomic<int> slots{0};
    const TileLaunch L = tile_launch(c, (const void *)k_tile_pass<V, EPI, SELF, NT>, slots, a.ntiles, true);
    a.xcd_order = L.order;
    a.xcd_res = L.res;
    a.walk = L.walk;
    hipLaunchKernelGGL((k_tile_pass<V, EPI, SELF, NT>), dim3(L.grid), dim3(kTileThreads), 0, c->stream, a);
    return (uint32_t)L.grid;
}

template <class V, int EPI, bool SELF = false>
static void launch_tile(tasep_ctx *c, TileArgs<V> a, const char *name, double bytes) {
    prof_begin(c);
    const uint32_t g = nt_out(c) ? launch_tile_nt<V, EPI, SELF, true>(c, a) : launch_tile_nt<V, EPI, SELF, false>(c, a);
    prof_end(c, name, bytes);
    if (EPI == EPI_HORNER || EPI == EPI_RK4ACC) launch_finish(c, g);
}

template <class V>
static void rk4_horner_tile(tasep_ctx *c, double h) {
    V *x = (V *)xcur(c);
    V *u = (V *)c->buf[2];
    V *v = (V *)c->buf[3];
    TileArgs<V> a = tile_base<V>(c);
    const double e = (double)sizeof(V);
    a.x = x;
    a.use_self = 1;
    a.gx = x; a.out = u; a.c = (V)(0.25 * h);
    launch_tile<V, EPI_STORE, true>(c, a, "horner_p1", 2 * e);
    a.gx = u; a.out = v; a.c = (V)(h / 3.0);
    launch_tile<V, EPI_STORE>(c, a, "horner_p23", 3 * e);
    a.gx = v; a.out = u; a.c = (V)(0.5 * h);
    launch_tile<V, EPI_STORE>(c, a, "horner_p23", 3 * e);
    a.gx = u; a.out = x; a.c = (V)h;
    launch_tile<V, EPI_HORNER>(c, a, "horner_p4", 3 * e);
}

// Staged RK4 on tiles (time-variant rates) with materialised stage inputs:
// each pass gathers one vector, y_i, and writes y_{i+1} = x + c_i k_i and
// the running sum acc = k1 + 2 k2 + 2 k3 (same roundings as stored k's).
//   s1: k1 = h L(t) x          y2 = x + k1/2   acc = k1
//   s2: k2 = h L(t+h/2) y2     y3 = x + k2/2   acc += 2 k2
//   s3: k3 = h L(t+h/2) y3     y4 = x + k3     acc += 2 k3
//   s4: k4 = h L(t+h) y4       x' = x + (acc + k4) / 6
template <class V>
static void rk4_staged_tile(tasep_ctx *c, double h, uint64_t j) {
    V *x = (V *)c->buf[c->cur ^ (int)(j & 1)];
    V *xn = (V *)c->buf[c->cur ^ (int)((j + 1) & 1)];
    V *ya = (V *)c->buf[2], *yb = (V *)c->buf[3], *acc = (V *)c->buf[4];
    TileArgs<V> a = tile_base<V>(c);
    const double e = (double)sizeof(V);
    a.x = x;
    a.c = (V)h;
    a.use_self = 1;
    a.acc_in = acc;
    a.acc_out = acc;
    a.rset = 0; a.gx = x; a.out = ya; a.cy = (V)0.5; a.stage = 0;
    launch_tile<V, EPI_STAGE, true>(c, a, "rk4_s1", 3 * e);
    a.rset = 1; a.gx = ya; a.out = yb; a.cy = (V)0.5; a.stage = 1;
    launch_tile<V, EPI_STAGE>(c, a, "rk4_s23", 5 * e);
    a.rset = 1; a.gx = yb; a.out = ya; a.cy = (V)1.0; a.stage = 2;
    launch_tile<V, EPI_STAGE>(c, a, "rk4_s23", 5 * e);
    a.rset = 2; a.gx = ya; a.out = xn;
    launch_tile<V, EPI_RK4ACC>(c, a, "rk4_s4", 4 * e);
}

static void prewarm_walk(tasep_ctx *c) {
    if (c->xcd_order != 4) return;
    if (c->gen == TASEP_CSR) {
        if (c->integ == TASEP_REF && c->n >= 10) ensure_walk(c, c->n - 6);
        return;
    }
    const int b = c->prec == TASEP_F32 ? TileGeom<float>::b : TileGeom<double>::b;
    if (c->n > b) ensure_walk(c, c->n - b);
}

// Tiles need n >= b + 1 (two tiles); below n = 18 the one-lane-per-state
// passes are faster (too few tiles to fill 256 CUs).
constexpr int kTileAutoMinN = 18, kTileRefAutoMinN = 19;

template <class V>
static bool use_tile(const tasep_ctx *c) {
    if (!c->tiled || c->gen != TASEP_MF || c->n < TileGeom<V>::b + 1) return false;
    return c->tiled == 2 || c->n >= kTileAutoMinN;
}

static bool use_tile_ref(const tasep_ctx *c) {
    if (!c->tiled || c->gen != TASEP_MF || c->n < TileGeom<double>::b + 1) return false;
    return c->tiled == 2 || c->n >= kTileRefAutoMinN;
}

static void ref_step(tasep_ctx *c, double h, uint64_t j) {
    if (use_tile_ref(c)) {
        TileRefArgs t;
        t.x = (const double *)c->buf[c->cur ^ (int)(j & 1)];
        t.out = (double *)c->buf[c->cur ^ (int)((j + 1) & 1)];
        t.rates = c->rates;
        t.st = c->st;
        t.pchk = c->pchk + (size_t)c->slot * kMaxBlocks;
        t.pmass = c->pmass + (size_t)c->slot * kMaxBlocks;
        t.h = h;
        t.n = c->n;
        t.ntiles = c->dim >> TileGeom<double>::b;
        static std::atomic<int> slots[2];
        const bool nt = nt_out(c);
        const void *kernel = nt ? (const void *)k_tile_ref<true> : (const void *)k_tile_ref<false>;
        const TileLaunch L = tile_launch(c, kernel, slots[nt], t.ntiles);
        const uint64_t g = L.grid;
        t.xcd_order = L.order;
        t.xcd_res = L.res;
        t.walk = L.walk;
        prof_begin(c);
        if (nt) hipLaunchKernelGGL(k_tile_ref<true>, dim3(g), dim3(kTileThreads), 0, c->stream, t);
        else hipLaunchKernelGGL(k_tile_ref<false>, dim3(g), dim3(kTileThreads), 0, c->stream, t);
        prof_end(c, "ref_tile", 16.0);
        launch_finish(c, (uint32_t)g);
        return;
    }
    RefArgs a;
    memset(&a, 0, sizeof a);
    a.x = (const double *)c->buf[c->cur ^ (int)(j & 1)];
    a.xn = (double *)c->buf[c->cur ^ (int)((j + 1) & 1)];
    a.rates = c->rates;
    a.rowPtr = c->rowPtr;
    a.colInd = c->colInd;
    a.val = c->val;
    a.st = c->st;
    a.pchk = c->pchk + (size_t)c->slot * kMaxBlocks;
    a.pmass = c->pmass + (size_t)c->slot * kMaxBlocks;
    a.dim = c->dim;
    a.h = h;
    a.g = make_lattice(c->n);
    prof_begin(c);
    if (c->gen == TASEP_CSR) {
        const uint64_t groups = (c->dim + kWave - 1) / kWave;
        const uint32_t g = (uint32_t)(groups < (uint64_t)kMaxBlocks ? groups : kMaxBlocks);
        // 64-row groups in the level walk over the group index (tile order 4;
        // groups is a power of two >= 128 here, the grid a multiple of 8)
        a.walk = c->xcd_order == 4 && groups >= 16 && (g % 8) == 0 ? ensure_walk(c, 63 - __builtin_clzll(groups))
                                                                   : nullptr;
        hipLaunchKernelGGL(k_ref_csr_wave, dim3(g), dim3(kWave), 0, c->stream, a);
        prof_end(c, "ref_step_csr", 24.0 + 16.0 * (2.0 + (c->n - 1) / 4.0));
        launch_finish(c, g);
        return;
    } else {
        hipLaunchKernelGGL(k_ref_step, dim3(grid_for(c->dim)), dim3(kBlock), 0, c->stream, a);
        prof_end(c, "ref_step_mf", 16.0);
    }
    launch_finish(c, (uint32_t)grid_for(c->dim));
}

static int lds_threads(const tasep_ctx *c) {
    return c->lds_threads ? c->lds_threads : lds_auto_threads(c->n);
}

// A single instance runs LDS-resident by default up to n = 12: beyond that
// one CU is slower than the streaming passes spread over the whole chip.
constexpr int kLdsAutoMaxN = 12;

static bool use_lds(const tasep_ctx *c) {
    const bool csr_ref = c->gen == TASEP_CSR && c->integ == TASEP_REF;
    if ((c->gen != TASEP_MF && !csr_ref) || !c->lds || !lds_fits(c->n, (int)c->esize, c->integ)) return false;
    return c->lds == 2 || c->n <= kLdsAutoMaxN;
}

// Allocate what a run needs and reset the device loop state (resume: keep
// t, the step count and the early-exit counters of the previous run).
static int prepare_run(tasep_ctx *c, double t0, double h, double thr, int patience, int resume = 0) {
    if (!c->have_rates) return fail(TASEP_ERR_STATE, "set rates before running");
    if (!isfinite(h) || !isfinite(t0)) return fail(TASEP_ERR_ARG, "t0 and h must be finite");
    c->pending = 0;
    c->slot = 0;
    c->window = !c->tv && patience >= 0;
    if (c->gen == TASEP_CSR && !c->csr_built) {
        int rc = tasep_build_csr(c);
        if (rc) return rc;
    }
    const bool staged = c->integ == TASEP_RK4 && (c->tv || c->force_staged);
    if (c->integ == TASEP_REF || staged) {
        int rc = ensure_buf(c, 1);
        if (rc) return rc;
    }
    if (c->integ == TASEP_RK4) {
        for (int i = 2; i < (staged ? 5 : 4); i++) {
            int rc = ensure_buf(c, i);
            if (rc) return rc;
        }
    }
    // fresh loop state, or the previous run's with new step parameters
    HIP_OK(hipMemcpyAsync(c->st_host, c->st, sizeof(RunState), hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    c->steps_before = resume ? c->st_host->steps_done : 0;
    if (resume) {
        c->st_host->h = h;
        c->st_host->thr_h = thr * h;
        c->st_host->patience = c->tv ? -1 : patience;
        c->st_host->t_rates = c->st_host->t;
        HIP_OK(hipMemcpyAsync(c->st, c->st_host, sizeof(RunState), hipMemcpyHostToDevice, c->stream));
        return TASEP_OK;
    }
    memset(c->st_host, 0, sizeof(RunState));
    c->st_host->t = t0;
    c->st_host->t_rates = t0;
    c->st_host->h = h;
    c->st_host->thr_h = thr * h;
    c->st_host->patience = c->tv ? -1 : patience;
    HIP_OK(hipMemcpyAsync(c->st, c->st_host, sizeof(RunState), hipMemcpyHostToDevice, c->stream));
    return TASEP_OK;
}

// Enqueue step j of the current run.
static void enqueue_step(tasep_ctx *c, double h, uint64_t j) {
    const bool staged = c->integ == TASEP_RK4 && (c->tv || c->force_staged);
    if (c->tv) {
        TvArgs tva;
        tva.type = c->tv_type;
        tva.amp = c->tv_par;
        tva.period = c->tv_par + kRateStride;
        tva.offset = c->tv_par + 2 * kRateStride;
        tva.c = c->tv_par + 3 * kRateStride;
        hipLaunchKernelGGL(k_tv_rates, dim3(1), dim3(64), 0, c->stream, c->st, tva, c->rates, c->n,
                           c->integ == TASEP_RK4 ? 3 : 1);
    }
    if (c->integ == TASEP_REF) {
        ref_step(c, h, j);
    } else if (c->prec == TASEP_F32) {
        if (use_tile<float>(c)) {
            if (staged) rk4_staged_tile<float>(c, h, j);
            else rk4_horner_tile<float>(c, h);
        } else {
            if (staged) rk4_staged_step<float, false>(c, h, j);
            else rk4_horner_step<float, false>(c, h);
        }
    } else if (c->gen == TASEP_CSR) {
        rk4_horner_step<double, true>(c, h);
    } else if (use_tile<double>(c)) {
        if (staged) rk4_staged_tile<double>(c, h, j);
        else rk4_horner_tile<double>(c, h);
    } else {
        if (staged) rk4_staged_step<double, false>(c, h, j);
        else rk4_horner_step<double, false>(c, h);
    }
}

// Read the loop state back and move the ping-pong index.
static int finish_run(tasep_ctx *c, tasep_run_info *info, bool inplace = false) {
    flush_finish(c);
    HIP_OK(hipMemcpyAsync(c->st_host, c->st, sizeof(RunState), hipMemcpyDeviceToHost, c->stream));
    HIP_OK(hipStreamSynchronize(c->stream));
    const RunState &s = *c->st_host;
    const bool staged = c->integ == TASEP_RK4 && (c->tv || c->force_staged);
    if (!inplace && (c->integ == TASEP_REF || staged))
        c->cur ^= (int)((s.steps_done - c->steps_before) & 1);
    c->steps_before = s.steps_done;
    if (info) {
        info->steps_done = s.steps_done;
        info->steps_at_stop = s.steps_at_stop;
        info->early_stop = s.early_stop;
        info->stopped = s.stopped;
        info->t = s.t;
        info->last_checksum = s.checksum;
        info->mass = s.mass;
    }
    return TASEP_OK;
}

extern "C" {

int tasep_run(tasep_ctx *c, double t0, double h, uint64_t steps, double thr, int patience,
              tasep_run_info *info) {
    return tasep_run_resume(c, t0, h, steps, thr, patience, 0, info);
}

int tasep_run_resume(tasep_ctx *c, double t0, double h, uint64_t steps, double thr, int patience,
                     int resume, tasep_run_info *info) {
    if (!c) return fail(TASEP_ERR_ARG, "NULL context");
    int rc = prepare_run(c, t0, h, thr, patience, resume);
    if (rc) return rc;
    if (use_lds(c)) {
        // the whole run in one workgroup, in launches of about 2^27
        // state-updates and at most 65536 steps
        const int nt = lds_threads(c);
        LdsInstance &in = *c->lds_inst_host;
        in.off = 0;
        in.n = c->n;
        in.tv = c->tv ? 1 : 0;
        in.rate_off = 0;
        in.tv_off = 0;
        in.id = 0;
        HIP_OK(hipMemcpyAsync(c->lds_inst, c->lds_inst_host, sizeof(LdsInstance), hipMemcpyHostToDevice,
                              c->stream));
        LdsArgs a;
        a.inst = c->lds_inst;
        a.x = xcur(c);
        a.rates = c->rates;
        a.tv_type = c->tv_type;
        a.tv_amp = c->tv_par;
        a.tv_period = c->tv_par ? c->tv_par + kRateStride : nullptr;
        a.tv_offset = c->tv_par ? c->tv_par + 2 * kRateStride : nullptr;
        a.tv_c = c->tv_par ? c->tv_par + 3 * kRateStride : nullptr;
        a.st = c->st;
        a.rowPtr = c->rowPtr;
        a.colInd = c->colInd;
        a.val = c->val;
        const bool csr = c->gen == TASEP_CSR;
        const bool tv = c->tv || (c->integ == TASEP_RK4 && c->force_staged);
        uint64_t chunk = ((uint64_t)1 << 27) >> c->n;
        if (chunk > 65536) chunk = 65536;
        for (uint64_t j = 0; j < steps; j += chunk) {
            a.steps = steps - j < chunk ? steps - j : chunk;
            launch_lds_any(c->prec, c->integ, tv, a, 1, nt, lds_bytes(c->n, (int)c->esize, c->integ), c->stream,
                           csr);
        }
        HIP_OK(hipGetLastError());
        return finish_run(c, info, true);
    }
    const bool watch = c->window;
    uint64_t j = 0;
    if (watch) {
        // Exit test armed: after e consecutive sub-threshold steps no exit is
        // possible for the next patience + 1 - e steps, so those steps run
        // back to back with their partials in ring slots, and one pair of
        // launches reduces them and replays the exit logic step by step.
        const int patience_ = c->st_host->patience;
        int early = c->st_host->early_stop;
        while (j < steps) {
            uint64_t m = (uint64_t)(patience_ + 1 - early);
            if (m < 1) m = 1;
            if (m > (uint64_t)kWindow) m = kWindow;
            if (m > steps - j) m = steps - j;
            for (uint64_t k = 0; k < m; k++) {
                c->slot = (uint32_t)k;
                enqueue_step(c, h, j + k);
            }
            c->slot = 0;
            hipLaunchKernelGGL(k_window_sums, dim3((uint32_t)m), dim3(kFinishThreads), 0, c->stream, c->pchk,
                               c->pmass, c->pending_blocks, (uint32_t)kMaxBlocks, c->wtot);
            hipLaunchKernelGGL(k_window_apply, dim3(1), dim3(64), 0, c->stream, c->st, c->wtot, (uint32_t)m);
            HIP_OK(hipGetLastError());
            j += m;
            HIP_OK(hipMemcpyAsync(c->st_host, c->st, sizeof(RunState), hipMemcpyDeviceToHost, c->stream));
            HIP_OK(hipStreamSynchronize(c->stream));
            if (c->st_host->stopped) break;
            early = c->st_host->early_stop;
        }
        return finish_run(c, info);
    }
    // steps per enqueued chunk: about 2^26 state-updates
    uint64_t chunk = ((uint64_t)1 << 26) >> c->n;
    if (chunk < 4) chunk = 4;
    if (chunk > 4096) chunk = 4096;
    // Long runs replay a captured HIP graph of `chunk` steps (an even number,
    // so the ping-pong buffers keep their parity); the remainder and short
    // runs are launched directly.
    if (c->graphs && chunk % 2 == 0 && steps >= 2 * chunk) {
        hipGraph_t g = nullptr;
        hipGraphExec_t exec = nullptr;
        const uint32_t p0 = c->pending;
        prewarm_walk(c);
        HIP_OK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        for (uint64_t k = 0; k < chunk; k++) enqueue_step(c, h, k);
        hipError_t e = hipStreamEndCapture(c->stream, &g);
        const uint32_t per = c->pending - p0;   // deferred reductions per replay
        c->pending = p0;
        if (e != hipSuccess) return fail(TASEP_ERR_HIP, "hipStreamEndCapture", e);
        e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
        hipGraphDestroy(g);
        if (e != hipSuccess) return fail(TASEP_ERR_HIP, "hipGraphInstantiate", e);
        for (; j + chunk <= steps; j += chunk) {
            e = hipGraphLaunch(exec, c->stream);
            if (e != hipSuccess) {
                hipGraphExecDestroy(exec);
                return fail(TASEP_ERR_HIP, "hipGraphLaunch", e);
            }
            c->pending += per;
        }
        HIP_OK(hipStreamSynchronize(c->stream));
        hipGraphExecDestroy(exec);
    }
    for (; j < steps; j++) enqueue_step(c, h, j);
    HIP_OK(hipGetLastError());
    return finish_run(c, info);
}

int tasep_set_option(tasep_ctx *c, int key, int64_t value) {
    if (!c) return fail(TASEP_ERR_ARG, "NULL context");
    switch (key) {
    case TASEP_OPT_TILED:
        if (value < 0 || value > 2) return fail(TASEP_ERR_ARG, "tiled: 0, 1 or 2");
        c->tiled = (int)value;
        break;
    case TASEP_OPT_XCD_ORDER:
        if (value < 0 || value > 4) return fail(TASEP_ERR_ARG, "tile order: 0 to 4");
        c->xcd_order = (int)value;
        break;
    case TASEP_OPT_RK4_STAGED: c->force_staged = value != 0; break;
    case TASEP_OPT_GRAPH: c->graphs = value != 0; break;
    case TASEP_OPT_LDS:
        if (value < 0 || value > 2) return fail(TASEP_ERR_ARG, "LDS mode: 0, 1 or 2");
        c->lds = (int)value;
        break;
    case TASEP_OPT_LDS_THREADS:
        if (value != 0 && (value < 64 || value > 1024 || (value & (value - 1))))
            return fail(TASEP_ERR_ARG, "LDS threads: 0 or a power of two in [64, 1024]");
        c->lds_threads = (int)value;
        break;
    case TASEP_OPT_STORE:
        if (value < 0 || value > 2) return fail(TASEP_ERR_ARG, "store policy: 0, 1 or 2");
        c->store = (int)value;
        break;
    case TASEP_OPT_TILE_BLOCKS:
        if (value != 0 && (value < 8 || value > kMaxBlocks || value % 8))
            return fail(TASEP_ERR_ARG, "tile blocks: 0 or a multiple of 8 in [8, 16384]");
        c->tile_blocks = (int)value;
        break;
    default: return fail(TASEP_ERR_ARG, "unknown option");
    }
    return TASEP_OK;
}

int tasep_step_profile(tasep_ctx *c, double t0, double h, int reps, tasep_kernel_time *out, int max,
                       int *count) {
    if (!c || !out || !count || reps < 1 || max < 1) return fail(TASEP_ERR_ARG, "bad argument");
    int rc = prepare_run(c, t0, h, 0.0, -1);
    if (rc) return rc;
    LaunchProf prof;
    for (int i = 0; i < 16; i++) HIP_OK(hipEventCreate(&prof.ev[i]));
    std::vector<double> acc(8, 0.0);
    int nk = 0;
    for (int r = 0; r < reps; r++) {
        prof.count = 0;
        g_prof = &prof;
        enqueue_step(c, h, (uint64_t)r);
        g_prof = nullptr;
        HIP_OK(hipGetLastError());
        HIP_OK(hipStreamSynchronize(c->stream));
        nk = prof.count;
        for (int k = 0; k < nk; k++) {
            float ms = 0.f;
            HIP_OK(hipEventElapsedTime(&ms, prof.ev[2 * k], prof.ev[2 * k + 1]));
            acc[k] += ms;
        }
    }
    for (int i = 0; i < 16; i++) hipEventDestroy(prof.ev[i]);
    *count = nk < max ? nk : max;
    for (int k = 0; k < *count; k++) {
        snprintf(out[k].name, sizeof out[k].name, "%s", prof.name[k]);
        out[k].ms = acc[k] / reps;
        out[k].bytes_per_state = prof.bytes[k];
    }
    return finish_run(c, nullptr);
}

int tasep_moments(tasep_ctx *c, double *y, double *cov, double *mul, double *mass) {
    if (!c) return fail(TASEP_ERR_ARG, "NULL context");
    return device_moments(c->stream, xcur(c), c->n, c->prec, &c->mom_part, &c->mom_out, y, cov, mul, mass);
}

int tasep_get_stream(tasep_ctx *c, void **stream) {
    if (!c || !stream) return fail(TASEP_ERR_ARG, "NULL argument");
    *stream = (void *)c->stream;
    return TASEP_OK;
}

int tasep_get_x_device(tasep_ctx *c, void **ptr) {
    if (!c || !ptr) return fail(TASEP_ERR_ARG, "NULL argument");
    *ptr = xcur(c);
    return TASEP_OK;
}

}  // extern "C"

// The two moment launches on a state vector x (2^n entries of the given
// precision): block partials into part (nb x nq doubles, every one written),
// their fixed-order sums into out (nq doubles), and the sums copied to E.
static int moment_sums(hipStream_t stream, const void *x, int n, int prec, int max_blocks, double *part,
                       double *out, double *E) {
    const uint64_t dim = (uint64_t)1 << n;
    const int nq = n * (n + 1) / 2 + 1;
    const uint64_t rows = dim >> kMomRowBits ? dim >> kMomRowBits : 1;
    const int nb = (int)(rows < (uint64_t)max_blocks ? rows : max_blocks);
    if (prec == TASEP_F32)
        hipLaunchKernelGGL(k_moments_partial<float>, dim3(nb), dim3(kBlock), 0, stream, (const float *)x, n, dim,
                           rows, part);
    else
        hipLaunchKernelGGL(k_moments_partial<double>, dim3(nb), dim3(kBlock), 0, stream, (const double *)x, n, dim,
                           rows, part);
    hipLaunchKernelGGL(k_moments_final, dim3(nq), dim3(kBlock), 0, stream, part, nb, nq, out);
    HIP_OK(hipGetLastError());
    HIP_OK(hipMemcpyAsync(E, out, nq * sizeof(double), hipMemcpyDeviceToHost, stream));
    HIP_OK(hipStreamSynchronize(stream));
    return TASEP_OK;
}

// y, cov, mul, mass from the sums E (the n(n+1)/2 pair sums by tri_index and
// the mass last); FP32 vectors are divided by their FP64 mass.
static void moments_from_sums(int n, int prec, double *E, double *y, double *cov, double *mul, double *mass);

// Moments of a state vector x (2^n entries of the given precision) on the
// device; *part and *out are allocated on first use and kept by the caller.
int tasep_host::device_moments(hipStream_t stream, const void *x, int n, int prec, double **part, double **out,
                               double *y, double *cov, double *mul, double *mass) {
    if (!*part) {
        HIP_OK(hipMalloc(part, (size_t)kMomMaxBlocks * kMomMaxQ * sizeof(double)));
        HIP_OK(hipMalloc(out, kWinMaxQ * sizeof(double)));
    }
    double E[kMomMaxQ];
    int rc = moment_sums(stream, x, n, prec, kMomMaxBlocks, *part, *out, E);
    if (rc) return rc;
    moments_from_sums(n, prec, E, y, cov, mul, mass);
    return TASEP_OK;
}

static void moments_from_sums(int n, int prec, double *E, double *y, double *cov, double *mul, double *mass) {
    const int nq = n * (n + 1) / 2 + 1;
    double M = E[nq - 1];
    if (prec == TASEP_F32 && M > 0.0) {
        for (int q = 0; q < nq - 1; q++) E[q] /= M;
        M = 1.0;
    }
    if (mass) *mass = M;
    double yy[kMaxN];
    for (int i = 0; i < n; i++) yy[i] = E[tri_index(i, i)];
    if (y) for (int i = 0; i < n; i++) y[i] = yy[i];
    for (int i = 0; i < n; i++) {
        for (int k = 0; k < n; k++) {
            const double e = E[i <= k ? tri_index(i, k) : tri_index(k, i)];
            // sum_s (b_k - y_k)(b_i - y_i) x_s = E_ik - y_i y_k (2 - M)
            const double cv = e - yy[i] * yy[k] * (2.0 - M);
            if (cov) cov[i * n + k] = cv;
            if (mul) mul[i * n + k] = yy[k] * yy[i] + cv;
        }
    }
}

template <int WW>
static void launch_windows(hipStream_t stream, const void *x, int n, int prec, uint64_t dim, uint64_t rows, int nb,
                           double *part) {
    if (prec == TASEP_F32)
        hipLaunchKernelGGL((k_windows_partial<float, WW>), dim3(nb), dim3(kBlock), 0, stream, (const float *)x, n,
                           dim, rows, part);
    else
        hipLaunchKernelGGL((k_windows_partial<double, WW>), dim3(nb), dim3(kBlock), 0, stream, (const double *)x,
                           n, dim, rows, part);
}

// Window marginals of a state vector (2^n entries, n >= width) on the
// device: W[o*2^width + p] for o = 0..n-width, width 4 or 6; FP32 vectors are
// divided by their FP64 mass.  The block count is capped so that the
// partials fit the moment buffer (kMomMaxBlocks x kMomMaxQ doubles).
int tasep_host::device_windows(hipStream_t stream, const void *x, int n, int prec, double **part, double **out,
                               int width, double *W) {
    const uint64_t dim = (uint64_t)1 << n;
    const int nq = (n - width + 1) * (1 << width) + 1;
    const uint64_t rows = dim >> kMomRowBits ? dim >> kMomRowBits : 1;
    const uint64_t cap = (uint64_t)kMomMaxBlocks * kMomMaxQ / (uint64_t)nq;
    const uint64_t maxb = cap < (uint64_t)kMomMaxBlocks ? cap : (uint64_t)kMomMaxBlocks;
    const int nb = (int)(rows < maxb ? rows : maxb);
    if (!*part) {
        HIP_OK(hipMalloc(part, (size_t)kMomMaxBlocks * kMomMaxQ * sizeof(double)));
        HIP_OK(hipMalloc(out, kWinMaxQ * sizeof(double)));
    }
    if (width == kWin6) launch_windows<kWin6>(stream, x, n, prec, dim, rows, nb, *part);
    else launch_windows<kWin>(stream, x, n, prec, dim, rows, nb, *part);
    hipLaunchKernelGGL(k_moments_final, dim3(nq), dim3(kBlock), 0, stream, *part, nb, nq, *out);
    HIP_OK(hipGetLastError());
    std::vector<double> E(nq);
    HIP_OK(hipMemcpyAsync(E.data(), *out, nq * sizeof(double), hipMemcpyDeviceToHost, stream));
    HIP_OK(hipStreamSynchronize(stream));
    const double M = E[nq - 1];
    for (int q = 0; q < nq - 1; q++) W[q] = (prec == TASEP_F32 && M > 0.0) ? E[q] / M : E[q];
    return TASEP_OK;
}

extern "C" int tasep_window_marginals(tasep_ctx *c, double *W) {
    return tasep_window_marginals_w(c, kWin, W);
}

extern "C" int tasep_window_marginals_w(tasep_ctx *c, int width, double *W) {
    if (!c || !W) return fail(TASEP_ERR_ARG, "NULL argument");
    if (width != kWin && width != kWin6) return fail(TASEP_ERR_ARG, "window width: 4 or 6");
    if (c->n < width) return fail(TASEP_ERR_ARG, "window marginals need n >= width");
    return device_windows(c->stream, xcur(c), c->n, c->prec, &c->mom_part, &c->mom_out, width, W);
}

// ---------------------------------------------------- allocation-free moments
// tasep_moments_noalloc keeps its block partials and sums in this file-scope
// device array instead of allocating (each GPU the library runs on holds its
// own copy until the process ends).  Host code never uses the array's name
// as an address: the per-device pointer from hipGetSymbolAddress is what the
// kernels and copies receive.
constexpr int kNoAllocBlocks = 256;                                     // partial-sum blocks
constexpr size_t kNoAllocDoubles = (size_t)kNoAllocBlocks * kMomMaxQ + kMomMaxQ;
static __device__ __align__(256) double scratch[kNoAllocDoubles];        // 1.17 MiB

constexpr int kMaxDevices = 64;
static double *g_scratch_dev[kMaxDevices];                              // per device, from hipGetSymbolAddress
static std::mutex g_scratch_mu;

// The current device's copy of `scratch`, looked up on the first call on
// that device.
static int scratch_on_device(double **ptr) {
    int dev = 0;
    HIP_OK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(TASEP_ERR_NODEV, "device index beyond the scratch table");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_dev[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return fail(TASEP_ERR_HIP, "hipGetSymbolAddress(scratch)", e);
        g_scratch_dev[dev] = (double *)p;
    }
    *ptr = g_scratch_dev[dev];
    return TASEP_OK;
}

extern "C" int tasep_moments_noalloc(const void *x, int n, int precision, double *y, double *cov, double *mul,
                                     double *mass, void *stream) {
    if (!x) return fail(TASEP_ERR_ARG, "NULL state vector");
    if (n < 2 || n > kMaxN) return fail(TASEP_ERR_ARG, "n must be in [2, 34] (the scratch holds 34-site moments)");
    if (precision != TASEP_F64 && precision != TASEP_F32) return fail(TASEP_ERR_ARG, "unknown precision");
    hipPointerAttribute_t attr;
    const hipError_t pe = hipPointerGetAttributes(&attr, x);
    if (pe != hipSuccess || attr.type != hipMemoryTypeDevice) {
        (void)hipGetLastError();  // clear the lookup's error state
        return fail(TASEP_ERR_ARG, "the state vector is not device memory");
    }
    int cur = 0;
    HIP_OK(hipGetDevice(&cur));
    if (attr.device != cur) return fail(TASEP_ERR_ARG, "the state vector is on another device than the current one");
    double *part = nullptr;
    int rc = scratch_on_device(&part);
    if (rc) return rc;
    double E[kMomMaxQ];
    rc = moment_sums((hipStream_t)stream, x, n, precision, kNoAllocBlocks, part,
                     part + (size_t)kNoAllocBlocks * kMomMaxQ, E);
    if (rc) return rc;
    moments_from_sums(n, precision, E, y, cov, mul, mass);
    return TASEP_OK;
}
