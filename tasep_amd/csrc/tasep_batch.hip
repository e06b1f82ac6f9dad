// tasep_batch.hip -- C ABI of the batched engine (tasep_batch_*).
//
// A batch holds `count` independent instances (own n, rates or rate
// functions, state and run state) in one set of device vectors.  A run
// drives all of them at once:
//   - instances whose RK4 working set fits one workgroup's LDS (n <= 13)
//     run every step inside one workgroup each, all in one launch
//     (k_lds_run), on a second stream;
//   - the others share each RK4 pass (k_batch_pass, one launch per pass for
//     all of them, XCD-aware block table), replayed from a captured HIP
//     graph of several steps.
// The reference runs one process per instance (src/masterEquation.c); the
// sequential per-instance loop over tasep_ctx is the comparison point.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "../../include/tasep_hip.h"
#include "batch_kernels.hpp"
#include "host_common.hpp"
#include "lds_launch.hpp"

using namespace tasep;
using namespace tasep_host;

struct tasep_batch {
    int count = 0, prec = TASEP_F32;
    size_t esize = 4;
    std::vector<int> n;
    std::vector<uint64_t> off;  // element offset of each instance
    uint64_t total = 0;         // elements per vector
    void *x = nullptr, *u = nullptr, *v = nullptr, *acc = nullptr;
    double *rates = nullptr;    // count * kRateStride
    int32_t *tv_type = nullptr; // count * kRateStride
    double *tv_par = nullptr;   // amp | period | offset | c, each count * kRateStride
    std::vector<double> h_rates, h_tv_par;
    std::vector<int32_t> h_tv_type;
    std::vector<char> have_rates, tv;
    RunState *st = nullptr, *st_host = nullptr;
    // LDS-resident instances: constant-rate ones first, then time-variant
    std::vector<int> lds_ids;
    int lds_const = 0;          // how many of lds_ids have constant rates
    LdsInstance *lds_tab = nullptr;
    // streamed instances
    std::vector<int> str_ids;
    BatchInst *btab = nullptr;
    BatchBlock *blocks = nullptr;
    BatchRange *ranges = nullptr;  // schedule 3: one tile range per workgroup
    uint32_t nblocks = 0;       // workgroups per pass (with padding)
    double *pchk = nullptr, *pmass = nullptr;
    hipStream_t stream = nullptr, stream2 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    bool tables_dirty = true, rates_dirty = true;
    // options
    int graph_steps = 32;       // steps per captured graph (0: direct launches)
    int use_lds = 1;
    int schedule = 0;           // shared passes: 0 autotuned, 1 one lane per state, 2 tiles, 3 big tiles
    int chosen = 1;             // schedule the current block table is built for
    int range_tiles = 2;        // schedule 3: tiles per workgroup (config-4 batch, per step: 1 356 us, 2 331, 4 344, 8 353)
    int store = 0;              // schedule 3 output stores: 0 by vector size, 1 non-temporal, 2 ordinary
    int concurrent = 1;
    hipGraphExec_t graph = nullptr;
    int graph_built_steps = 0;
    double *mom_part = nullptr, *mom_out = nullptr;
};

namespace {

constexpr uint32_t kNoInst = 0xffffffffu;

// Smallest instance the streamed passes take: one workgroup's states.
constexpr int kStreamMinN = 10;

bool lds_class(const tasep_batch *b, int i) {
    return b->n[i] < kStreamMinN || (b->use_lds && lds_fits(b->n[i], (int)b->esize, TASEP_RK4));
}

void drop_graph(tasep_batch *b) {
    if (b->graph) hipGraphExecDestroy(b->graph);
    b->graph = nullptr;
    b->graph_built_steps = 0;
}

// States per workgroup of the shared passes under schedule `sched`.
uint32_t states_per_block(const tasep_batch *b, int sched) {
    if (sched == 2) return 1u << (b->prec == TASEP_F32 ? BatchTileGeom<float>::TL : BatchTileGeom<double>::TL);
    return kBatchThreads * (b->prec == TASEP_F32 ? BatchGeom<float>::E : BatchGeom<double>::E);
}

// The tile schedule needs at least two tiles per streamed instance.
bool tiles_fit(const tasep_batch *b) {
    const uint32_t per = states_per_block(b, 2);
    for (int i = 0; i < b->count; i++)
        if (!lds_class(b, i) && ((uint64_t)1 << b->n[i]) < 2 * (uint64_t)per) return false;
    return true;
}

// log2 of the big tile (schedule 3): the single-lattice tile pass's.
int bigtile_bits(const tasep_batch *b) {
    return b->prec == TASEP_F32 ? TileGeom<float>::b : TileGeom<double>::b;
}

// The big-tile schedule needs two of its tiles per streamed instance.
bool bigtiles_fit(const tasep_batch *b) {
    for (int i = 0; i < b->count; i++)
        if (!lds_class(b, i) && b->n[i] < bigtile_bits(b) + 1) return false;
    return true;
}

// Work units of one pass over instance i: blocks (schedules 1, 2) or tile
// ranges (schedule 3).
uint32_t units_of(const tasep_batch *b, int i, int sched) {
    if (sched == 3) {
        const uint64_t nt = (uint64_t)1 << (b->n[i] - bigtile_bits(b));
        return (uint32_t)((nt + b->range_tiles - 1) / b->range_tiles);
    }
    return (uint32_t)(((uint64_t)1 << b->n[i]) / states_per_block(b, sched));
}

// Instance tables: LDS class and streamed class, XCD-aware block table for
// schedule `sched`.
int build_tables(tasep_batch *b, int sched) {
    b->lds_ids.clear();
    b->str_ids.clear();
    std::vector<int> lds_tv;
    for (int i = 0; i < b->count; i++) {
        if (lds_class(b, i)) (b->tv[i] ? lds_tv : b->lds_ids).push_back(i);
        else b->str_ids.push_back(i);
    }
    b->lds_const = (int)b->lds_ids.size();
    b->lds_ids.insert(b->lds_ids.end(), lds_tv.begin(), lds_tv.end());
    std::vector<LdsInstance> lt(b->lds_ids.size());
    for (size_t j = 0; j < lt.size(); j++) {
        const int i = b->lds_ids[j];
        lt[j].off = b->off[i];
        lt[j].n = b->n[i];
        lt[j].tv = b->tv[i];
        lt[j].rate_off = i * kRateStride;
        lt[j].tv_off = i * kRateStride;
        lt[j].id = i;
    }
    hipFree(b->lds_tab);
    b->lds_tab = nullptr;
    if (!lt.empty()) {
        HIP_OK(hipMalloc(&b->lds_tab, lt.size() * sizeof(LdsInstance)));
        HIP_OK(hipMemcpy(b->lds_tab, lt.data(), lt.size() * sizeof(LdsInstance), hipMemcpyHostToDevice));
    }
    // streamed instances: partial slots, then the block table.  Workgroup w
    // runs on XCD w % 8, so instance j's blocks go to the XCD with the least
    // work so far, at positions 8 * k + xcd: each instance's vectors stay in
    // one XCD's L2.  Work is weighed by the bytes per state and step of the
    // instance's RK4 form (stored stages 68, Horner 44 in FP32), and the
    // time-variant instances come first (larger first within each kind), so
    // every XCD ends its pass with the cheap blocks (config-4 batch: the
    // XCDs' work within 0.4% of each other instead of 23% apart; 329 -> 322
    // us per step).
    b->chosen = sched;
    std::vector<BatchInst> bt(b->str_ids.size());
    uint32_t slots = 0;
    for (size_t j = 0; j < bt.size(); j++) {
        const int i = b->str_ids[j];
        bt[j].off = b->off[i];
        bt[j].n = b->n[i];
        bt[j].tv = b->tv[i];
        bt[j].rate_off = i * kRateStride;
        bt[j].tv_off = i * kRateStride;
        bt[j].part_off = slots;
        bt[j].nblocks = units_of(b, i, sched);
        bt[j].id = i;
        slots += bt[j].nblocks;
    }
    std::vector<size_t> order(bt.size());
    for (size_t j = 0; j < order.size(); j++) order[j] = j;
    std::stable_sort(order.begin(), order.end(), [&](size_t p, size_t q) {
        if (bt[p].tv != bt[q].tv) return bt[p].tv > bt[q].tv;
        return bt[p].nblocks > bt[q].nblocks;
    });
    auto unit_cost = [&](size_t j) -> uint64_t { return bt[j].tv ? 68 : 44; };
    // An instance larger than an eighth of all blocks is split into eight
    // contiguous runs, one per XCD (a lone large lattice still fills the chip).
    uint64_t total_blocks = 0;
    for (const BatchInst &bi : bt) total_blocks += bi.nblocks;
    std::vector<std::vector<BatchBlock>> lane(8);
    uint64_t work[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (size_t j : order) {
        const uint32_t nb = bt[j].nblocks;
        const uint64_t uc = unit_cost(j);
        if (nb >= 8 && (uint64_t)nb * 8 > total_blocks) {
            for (uint32_t x = 0; x < 8; x++)
                for (uint32_t k = x * (nb / 8); k < (x + 1) * (nb / 8); k++) {
                    lane[x].push_back(BatchBlock{(uint32_t)j, k});
                    work[x] += uc;
                }
            for (uint32_t k = 8 * (nb / 8); k < nb; k++) {
                lane[k % 8].push_back(BatchBlock{(uint32_t)j, k});
                work[k % 8] += uc;
            }
            continue;
        }
        size_t best = 0;
        for (size_t x = 1; x < 8; x++)
            if (work[x] < work[best]) best = x;
        for (uint32_t k = 0; k < nb; k++) lane[best].push_back(BatchBlock{(uint32_t)j, k});
        work[best] += uc * nb;
    }
    size_t longest = 0;
    for (auto &l : lane) longest = std::max(longest, l.size());
    std::vector<BatchBlock> tab(longest * 8, BatchBlock{kNoInst, 0});
    for (size_t x = 0; x < 8; x++)
        for (size_t k = 0; k < lane[x].size(); k++) tab[8 * k + x] = lane[x][k];
    b->nblocks = (uint32_t)tab.size();
    // schedule 3: unit k of instance j is the tile range [k K, min((k+1) K, tiles))
    std::vector<BatchRange> rt;
    if (sched == 3) {
        rt.resize(tab.size(), BatchRange{kNoInst, 0, 0, 0});
        for (size_t w = 0; w < tab.size(); w++) {
            if (tab[w].inst == kNoInst) continue;
            const BatchInst &bi = bt[tab[w].inst];
            const uint32_t nt = 1u << (bi.n - bigtile_bits(b)), K = (uint32_t)b->range_tiles;
            const uint32_t t0 = tab[w].blk * K;
            rt[w] = BatchRange{tab[w].inst, t0, std::min(K, nt - t0), bi.part_off + tab[w].blk};
        }
    }
    hipFree(b->btab);
    hipFree(b->blocks);
    hipFree(b->ranges);
    hipFree(b->pchk);
    hipFree(b->pmass);
    b->btab = nullptr;
    b->blocks = nullptr;
    b->ranges = nullptr;
    b->pchk = b->pmass = nullptr;
    if (!bt.empty()) {
        HIP_OK(hipMalloc(&b->btab, bt.size() * sizeof(BatchInst)));
        HIP_OK(hipMemcpy(b->btab, bt.data(), bt.size() * sizeof(BatchInst), hipMemcpyHostToDevice));
        HIP_OK(hipMalloc(&b->blocks, tab.size() * sizeof(BatchBlock)));
        HIP_OK(hipMemcpy(b->blocks, tab.data(), tab.size() * sizeof(BatchBlock), hipMemcpyHostToDevice));
        if (!rt.empty()) {
            HIP_OK(hipMalloc(&b->ranges, rt.size() * sizeof(BatchRange)));
            HIP_OK(hipMemcpy(b->ranges, rt.data(), rt.size() * sizeof(BatchRange), hipMemcpyHostToDevice));
        }
        HIP_OK(hipMalloc(&b->pchk, slots * sizeof(double)));
        HIP_OK(hipMalloc(&b->pmass, slots * sizeof(double)));
    }
    b->tables_dirty = false;
    drop_graph(b);
    return TASEP_OK;
}

template <class V>
BatchArgs<V> batch_args(tasep_batch *b, int pass) {
    BatchArgs<V> a;
    a.inst = b->btab;
    a.blocks = b->blocks;
    a.x = (V *)b->x;
    a.u = (V *)b->u;
    a.v = (V *)b->v;
    a.acc = (V *)b->acc;
    a.rates = b->rates;
    a.tv_type = b->tv_type;
    const size_t stride = (size_t)b->count * kRateStride;
    a.tv_amp = b->tv_par;
    a.tv_period = b->tv_par + stride;
    a.tv_offset = b->tv_par + 2 * stride;
    a.tv_c = b->tv_par + 3 * stride;
    a.st = b->st;
    a.pchk = b->pchk;
    a.pmass = b->pmass;
    a.pass = pass;
    return a;
}

template <class V>
BatchBigArgs<V> bigtile_args(tasep_batch *b) {
    BatchArgs<V> c = batch_args<V>(b, 0);
    BatchBigArgs<V> a;
    a.inst = c.inst;
    a.ranges = b->ranges;
    a.x = c.x; a.u = c.u; a.v = c.v; a.acc = c.acc;
    a.rates = c.rates;
    a.tv_type = c.tv_type;
    a.tv_amp = c.tv_amp; a.tv_period = c.tv_period; a.tv_offset = c.tv_offset; a.tv_c = c.tv_c;
    a.st = c.st;
    a.pchk = c.pchk;
    a.pmass = c.pmass;
    return a;
}

// Output stores of the big-tile passes, as a single lattice's tiled passes
// choose them from one vector's size (here the batch's concatenated vector):
// ordinary stores for 16-128 MiB, non-temporal otherwise.
bool bigtile_nt(const tasep_batch *b) {
    if (b->store) return b->store == 1;
    const uint64_t bytes = b->total * b->esize;
    return bytes < ((uint64_t)16 << 20) || bytes > ((uint64_t)128 << 20);
}

template <class V, bool NT>
void launch_bigtile(tasep_batch *b, int p, hipStream_t s) {
    const BatchBigArgs<V> a = bigtile_args<V>(b);
    const dim3 g(b->nblocks), t(kTileThreads);
    switch (p) {
    case 0: hipLaunchKernelGGL((k_batch_bigtile<V, 0, NT>), g, t, 0, s, a); break;
    case 1: hipLaunchKernelGGL((k_batch_bigtile<V, 1, NT>), g, t, 0, s, a); break;
    case 2: hipLaunchKernelGGL((k_batch_bigtile<V, 2, NT>), g, t, 0, s, a); break;
    default: hipLaunchKernelGGL((k_batch_bigtile<V, 3, NT>), g, t, 0, s, a); break;
    }
}

void launch_pass(tasep_batch *b, int p, hipStream_t s) {
    if (b->chosen == 3) {
        const bool nt = bigtile_nt(b);
        if (b->prec == TASEP_F32) nt ? launch_bigtile<float, true>(b, p, s) : launch_bigtile<float, false>(b, p, s);
        else nt ? launch_bigtile<double, true>(b, p, s) : launch_bigtile<double, false>(b, p, s);
    } else if (b->chosen == 2) {
        if (b->prec == TASEP_F32)
            hipLaunchKernelGGL(k_batch_tile<float>, dim3(b->nblocks), dim3(kBatchTileThreads), 0, s,
                               batch_args<float>(b, p));
        else
            hipLaunchKernelGGL(k_batch_tile<double>, dim3(b->nblocks), dim3(kBatchTileThreads), 0, s,
                               batch_args<double>(b, p));
    } else {
        if (b->prec == TASEP_F32)
            hipLaunchKernelGGL(k_batch_pass<float>, dim3(b->nblocks), dim3(kBatchThreads), 0, s,
                               batch_args<float>(b, p));
        else
            hipLaunchKernelGGL(k_batch_pass<double>, dim3(b->nblocks), dim3(kBatchThreads), 0, s,
                               batch_args<double>(b, p));
    }
}

// Autotuner of the shared passes: time a side-effect-free pass (pass 2 of
// the step writes only the work vectors, which the next step's first pass
// rewrites) under every schedule that fits and keep the fastest.  Needs
// valid run states (h, stopped) on the device.
int autotune(tasep_batch *b) {
    hipEvent_t e0, e1;
    HIP_OK(hipEventCreate(&e0));
    HIP_OK(hipEventCreate(&e1));
    float ms[4] = {0, 0, 0, 0};
    int best = 1;
    for (int sc = 1; sc <= 3; sc++) {
        if ((sc == 2 && !tiles_fit(b)) || (sc == 3 && !bigtiles_fit(b))) continue;
        int rc = build_tables(b, sc);
        if (rc) return rc;
        launch_pass(b, 1, b->stream);
        HIP_OK(hipEventRecord(e0, b->stream));
        for (int r = 0; r < 3; r++) launch_pass(b, 1, b->stream);
        HIP_OK(hipEventRecord(e1, b->stream));
        HIP_OK(hipEventSynchronize(e1));
        HIP_OK(hipEventElapsedTime(&ms[sc], e0, e1));
        if (ms[sc] < ms[best]) best = sc;
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    return build_tables(b, best);
}

// The streamed instances' passes of one step and their run-state update.
void enqueue_stream_step(tasep_batch *b, hipStream_t s) {
    for (int p = 0; p < 4; p++) launch_pass(b, p, s);
    hipLaunchKernelGGL(k_batch_finish, dim3((unsigned)b->str_ids.size()), dim3(kBatchThreads), 0, s, b->btab,
                       b->st, b->pchk, b->pmass);
}

int build_graph(tasep_batch *b, int steps) {
    drop_graph(b);
    hipGraph_t g = nullptr;
    HIP_OK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < steps; k++) enqueue_stream_step(b, b->stream);
    HIP_OK(hipStreamEndCapture(b->stream, &g));
    hipError_t e = hipGraphInstantiate(&b->graph, g, nullptr, nullptr, 0);
    hipGraphDestroy(g);
    if (e != hipSuccess) return fail(TASEP_ERR_HIP, "hipGraphInstantiate", e);
    b->graph_built_steps = steps;
    return TASEP_OK;
}

int check_index(const tasep_batch *b, int i) {
    if (!b) return fail(TASEP_ERR_ARG, "NULL batch");
    if (i < 0 || i >= b->count) return fail(TASEP_ERR_ARG, "instance index out of range");
    return TASEP_OK;
}

template <class V>
__global__ void k_batch_fill(V *x, uint64_t len, double v) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < len; i += (uint64_t)gridDim.x * blockDim.x)
        x[i] = (V)v;
}

}  // namespace

extern "C" {

int tasep_batch_destroy(tasep_batch *b) {
    if (!b) return TASEP_OK;
    if (b->stream) hipStreamSynchronize(b->stream);
    if (b->stream2) hipStreamSynchronize(b->stream2);
    drop_graph(b);
    hipFree(b->x);
    hipFree(b->u);
    hipFree(b->v);
    hipFree(b->acc);
    hipFree(b->rates);
    hipFree(b->tv_type);
    hipFree(b->tv_par);
    hipFree(b->st);
    if (b->st_host) hipHostFree(b->st_host);
    hipFree(b->lds_tab);
    hipFree(b->btab);
    hipFree(b->blocks);
    hipFree(b->ranges);
    hipFree(b->pchk);
    hipFree(b->pmass);
    hipFree(b->mom_part);
    hipFree(b->mom_out);
    if (b->ev_fork) hipEventDestroy(b->ev_fork);
    if (b->ev_join) hipEventDestroy(b->ev_join);
    if (b->stream) hipStreamDestroy(b->stream);
    if (b->stream2) hipStreamDestroy(b->stream2);
    delete b;
    return TASEP_OK;
}

int tasep_batch_create(tasep_batch **out, int count, const int32_t *n, int precision) {
    if (!out) return fail(TASEP_ERR_ARG, "batch pointer is NULL");
    *out = nullptr;
    if (count < 1 || !n) return fail(TASEP_ERR_ARG, "a batch needs at least one instance");
    if (precision != TASEP_F32 && precision != TASEP_F64) return fail(TASEP_ERR_ARG, "unknown precision");
    for (int i = 0; i < count; i++)
        if (n[i] < 2 || n[i] > kBatchMaxN) return fail(TASEP_ERR_ARG, "batch instances need 2 <= n <= 30");
    int rc = check_device();
    if (rc) return rc;
    tasep_batch *b = new tasep_batch;
    b->count = count;
    b->prec = precision;
    b->esize = precision == TASEP_F32 ? 4 : 8;
    b->n.assign(n, n + count);
    b->off.resize(count);
    uint64_t o = 0;
    for (int i = 0; i < count; i++) {
        b->off[i] = o;
        o += ((((uint64_t)1 << n[i]) + 63) / 64) * 64;  // 256-byte aligned (FP32) instances
    }
    b->total = o;
    b->h_rates.assign((size_t)count * kRateStride, 0.0);
    b->h_tv_par.assign((size_t)4 * count * kRateStride, 0.0);
    b->h_tv_type.assign((size_t)count * kRateStride, 0);
    b->have_rates.assign(count, 0);
    b->tv.assign(count, 0);
#define CK(expr)                   \
    do {                           \
        int r_ = (expr);           \
        if (r_) {                  \
            tasep_batch_destroy(b);\
            return r_;             \
        }                          \
    } while (0)
    auto hipck = [&](hipError_t e, const char *what) {
        return e == hipSuccess ? TASEP_OK
                               : fail(e == hipErrorOutOfMemory ? TASEP_ERR_NOMEM : TASEP_ERR_HIP, what, e);
    };
    CK(hipck(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking), "hipStreamCreate"));
    CK(hipck(hipStreamCreateWithFlags(&b->stream2, hipStreamNonBlocking), "hipStreamCreate"));
    CK(hipck(hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming), "hipEventCreate"));
    CK(hipck(hipEventCreateWithFlags(&b->ev_join, hipEventDisableTiming), "hipEventCreate"));
    const size_t bytes = b->total * b->esize;
    CK(hipck(hipMalloc(&b->x, bytes), "hipMalloc(x)"));
    CK(hipck(hipMalloc(&b->u, bytes), "hipMalloc(u)"));
    CK(hipck(hipMalloc(&b->v, bytes), "hipMalloc(v)"));
    CK(hipck(hipMalloc(&b->acc, bytes), "hipMalloc(acc)"));
    CK(hipck(hipMalloc(&b->rates, (size_t)count * kRateStride * sizeof(double)), "hipMalloc(rates)"));
    CK(hipck(hipMalloc(&b->tv_type, (size_t)count * kRateStride * sizeof(int32_t)), "hipMalloc(tv)"));
    CK(hipck(hipMalloc(&b->tv_par, (size_t)4 * count * kRateStride * sizeof(double)), "hipMalloc(tv)"));
    CK(hipck(hipMalloc(&b->st, (size_t)count * sizeof(RunState)), "hipMalloc(state)"));
    CK(hipck(hipHostMalloc((void **)&b->st_host, (size_t)count * sizeof(RunState), hipHostMallocDefault),
             "hipHostMalloc(state)"));
    memset(b->st_host, 0, (size_t)count * sizeof(RunState));
    CK(hipck(hipMemcpy(b->st, b->st_host, (size_t)count * sizeof(RunState), hipMemcpyHostToDevice), "init"));
    CK(tasep_batch_set_x0_uniform(b));
#undef CK
    *out = b;
    return TASEP_OK;
}

int tasep_batch_set_rates(tasep_batch *b, int i, const double *r) {
    int rc = check_index(b, i);
    if (rc) return rc;
    if (!r) return fail(TASEP_ERR_ARG, "NULL rates");
    for (int k = 0; k <= b->n[i]; k++)
        if (!(r[k] >= 0.0) || !isfinite(r[k])) return fail(TASEP_ERR_ARG, "rates must be finite and >= 0");
    for (int k = 0; k <= b->n[i]; k++) b->h_rates[(size_t)i * kRateStride + k] = r[k];
    if (b->tv[i]) b->tables_dirty = true;
    b->tv[i] = 0;
    b->have_rates[i] = 1;
    b->rates_dirty = true;
    return TASEP_OK;
}

int tasep_batch_set_time_variant(tasep_batch *b, int i, const int32_t *type, const double *amp,
                                 const double *period, const double *offset, const double *c) {
    int rc = check_index(b, i);
    if (rc) return rc;
    if (!type || !amp || !period || !offset || !c) return fail(TASEP_ERR_ARG, "NULL argument");
    const size_t stride = (size_t)b->count * kRateStride;
    for (int k = 0; k <= b->n[i]; k++) {
        if (type[k] < 0 || type[k] > 2) return fail(TASEP_ERR_ARG, "rate function type must be 0, 1 or 2");
        const size_t o = (size_t)i * kRateStride + k;
        b->h_tv_type[o] = type[k];
        b->h_tv_par[o] = amp[k];
        b->h_tv_par[stride + o] = period[k];
        b->h_tv_par[2 * stride + o] = offset[k];
        b->h_tv_par[3 * stride + o] = c[k];
    }
    if (!b->tv[i]) b->tables_dirty = true;
    b->tv[i] = 1;
    b->have_rates[i] = 1;
    b->rates_dirty = true;
    return TASEP_OK;
}

int tasep_batch_set_x0_uniform(tasep_batch *b) {
    if (!b) return fail(TASEP_ERR_ARG, "NULL batch");
    for (int i = 0; i < b->count; i++) {
        const uint64_t d = (uint64_t)1 << b->n[i];
        const int grid = (int)std::min<uint64_t>((d + 255) / 256, 1024);
        if (b->prec == TASEP_F32)
            hipLaunchKernelGGL(k_batch_fill<float>, dim3(grid), dim3(256), 0, b->stream, (float *)b->x + b->off[i], d,
                               1.0 / (double)d);
        else
            hipLaunchKernelGGL(k_batch_fill<double>, dim3(grid), dim3(256), 0, b->stream, (double *)b->x + b->off[i],
                               d, 1.0 / (double)d);
    }
    HIP_OK(hipGetLastError());
    HIP_OK(hipStreamSynchronize(b->stream));
    return TASEP_OK;
}

int tasep_batch_set_x(tasep_batch *b, int i, const double *x) {
    int rc = check_index(b, i);
    if (rc) return rc;
    if (!x) return fail(TASEP_ERR_ARG, "NULL vector");
    const uint64_t d = (uint64_t)1 << b->n[i];
    if (b->prec == TASEP_F64) {
        HIP_OK(hipMemcpy((double *)b->x + b->off[i], x, d * sizeof(double), hipMemcpyHostToDevice));
    } else {
        std::vector<float> f(d);
        for (uint64_t s = 0; s < d; s++) f[s] = (float)x[s];
        HIP_OK(hipMemcpy((float *)b->x + b->off[i], f.data(), d * sizeof(float), hipMemcpyHostToDevice));
    }
    return TASEP_OK;
}

int tasep_batch_get_x(tasep_batch *b, int i, double *x) {
    int rc = check_index(b, i);
    if (rc) return rc;
    if (!x) return fail(TASEP_ERR_ARG, "NULL vector");
    const uint64_t d = (uint64_t)1 << b->n[i];
    HIP_OK(hipStreamSynchronize(b->stream));
    if (b->prec == TASEP_F64) {
        HIP_OK(hipMemcpy(x, (double *)b->x + b->off[i], d * sizeof(double), hipMemcpyDeviceToHost));
    } else {
        std::vector<float> f(d);
        HIP_OK(hipMemcpy(f.data(), (float *)b->x + b->off[i], d * sizeof(float), hipMemcpyDeviceToHost));
        for (uint64_t s = 0; s < d; s++) x[s] = f[s];
    }
    return TASEP_OK;
}

int tasep_batch_moments(tasep_batch *b, int i, double *y, double *cov, double *mul, double *mass) {
    int rc = check_index(b, i);
    if (rc) return rc;
    const void *x = (const char *)b->x + b->off[i] * b->esize;
    return device_moments(b->stream, x, b->n[i], b->prec, &b->mom_part, &b->mom_out, y, cov, mul, mass);
}

int tasep_batch_window_marginals(tasep_batch *b, int i, double *W) {
    int rc = check_index(b, i);
    if (rc) return rc;
    if (!W) return fail(TASEP_ERR_ARG, "NULL argument");
    if (b->n[i] < 4) return fail(TASEP_ERR_ARG, "window marginals need n >= 4");
    const void *x = (const char *)b->x + b->off[i] * b->esize;
    return device_windows(b->stream, x, b->n[i], b->prec, &b->mom_part, &b->mom_out, 4, W);
}

int tasep_batch_get_option(tasep_batch *b, int key, int64_t *value) {
    if (!b || !value) return fail(TASEP_ERR_ARG, "NULL argument");
    switch (key) {
    case TASEP_BATCH_OPT_GRAPH_STEPS: *value = b->graph_steps; break;
    case TASEP_BATCH_OPT_LDS: *value = b->use_lds; break;
    case TASEP_BATCH_OPT_CONCURRENT: *value = b->concurrent; break;
    case TASEP_BATCH_OPT_SCHEDULE: *value = b->tables_dirty ? b->schedule : b->chosen; break;
    case TASEP_BATCH_OPT_RANGE_TILES: *value = b->range_tiles; break;
    case TASEP_BATCH_OPT_STORE: *value = b->store; break;
    default: return fail(TASEP_ERR_ARG, "unknown batch option");
    }
    return TASEP_OK;
}

int tasep_batch_set_option(tasep_batch *b, int key, int64_t value) {
    if (!b) return fail(TASEP_ERR_ARG, "NULL batch");
    switch (key) {
    case TASEP_BATCH_OPT_GRAPH_STEPS:
        if (value < 0 || value > 4096) return fail(TASEP_ERR_ARG, "graph steps: 0..4096");
        if ((int)value != b->graph_steps) drop_graph(b);
        b->graph_steps = (int)value;
        break;
    case TASEP_BATCH_OPT_LDS:
        b->use_lds = value != 0;
        b->tables_dirty = true;
        break;
    case TASEP_BATCH_OPT_CONCURRENT: b->concurrent = value != 0; break;
    case TASEP_BATCH_OPT_SCHEDULE:
        if (value < 0 || value > 3) return fail(TASEP_ERR_ARG, "schedule: 0 (autotuned), 1, 2 or 3");
        b->schedule = (int)value;
        b->tables_dirty = true;
        break;
    case TASEP_BATCH_OPT_RANGE_TILES:
        if (value < 1 || value > 64) return fail(TASEP_ERR_ARG, "tiles per workgroup: 1..64");
        b->range_tiles = (int)value;
        b->tables_dirty = true;
        break;
    case TASEP_BATCH_OPT_STORE:
        if (value < 0 || value > 2) return fail(TASEP_ERR_ARG, "store policy: 0, 1 or 2");
        b->store = (int)value;
        drop_graph(b);
        break;
    default: return fail(TASEP_ERR_ARG, "unknown batch option");
    }
    return TASEP_OK;
}

int tasep_batch_run(tasep_batch *b, double t0, double h, uint64_t steps, double thr, int patience,
                    tasep_run_info *info) {
    if (!b) return fail(TASEP_ERR_ARG, "NULL batch");
    if (!isfinite(h) || !isfinite(t0)) return fail(TASEP_ERR_ARG, "t0 and h must be finite");
    for (int i = 0; i < b->count; i++)
        if (!b->have_rates[i]) return fail(TASEP_ERR_STATE, "set rates of every instance before running");
    HIP_OK(hipStreamSynchronize(b->stream));
    if (b->rates_dirty) {
        const size_t m = (size_t)b->count * kRateStride;
        HIP_OK(hipMemcpy(b->rates, b->h_rates.data(), m * sizeof(double), hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(b->tv_type, b->h_tv_type.data(), m * sizeof(int32_t), hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(b->tv_par, b->h_tv_par.data(), 4 * m * sizeof(double), hipMemcpyHostToDevice));
        b->rates_dirty = false;
    }
    for (int i = 0; i < b->count; i++) {
        RunState &s = b->st_host[i];
        memset(&s, 0, sizeof s);
        s.t = t0;
        s.h = h;
        s.thr_h = thr * h;
        s.patience = b->tv[i] ? -1 : patience;
    }
    HIP_OK(hipMemcpyAsync(b->st, b->st_host, (size_t)b->count * sizeof(RunState), hipMemcpyHostToDevice,
                          b->stream));
    if (b->tables_dirty) {
        int sched = b->schedule;
        // the streamed set decides whether tiles fit (build once to know it)
        int rc = build_tables(b, 1);
        if (rc) return rc;
        if (sched == 3 && !bigtiles_fit(b)) sched = 2;
        if (sched == 2 && !tiles_fit(b)) sched = 1;
        if (sched == 0 && !b->str_ids.empty()) rc = autotune(b);
        else rc = build_tables(b, sched == 0 ? 1 : sched);
        if (rc) return rc;
    }
    // LDS-resident instances: whole run, on the second stream when concurrent
    hipStream_t ls = b->concurrent ? b->stream2 : b->stream;
    if (!b->lds_ids.empty()) {
        HIP_OK(hipEventRecord(b->ev_fork, b->stream));
        HIP_OK(hipStreamWaitEvent(ls, b->ev_fork, 0));
        LdsArgs a;
        a.x = b->x;
        a.rates = b->rates;
        a.tv_type = b->tv_type;
        const size_t stride = (size_t)b->count * kRateStride;
        a.tv_amp = b->tv_par;
        a.tv_period = b->tv_par + stride;
        a.tv_offset = b->tv_par + 2 * stride;
        a.tv_c = b->tv_par + 3 * stride;
        a.st = b->st;
        a.rowPtr = a.colInd = nullptr;
        a.val = nullptr;
        const int groups[2][2] = {{0, b->lds_const}, {b->lds_const, (int)b->lds_ids.size()}};
        for (int gi = 0; gi < 2; gi++) {
            const int lo = groups[gi][0], hi = groups[gi][1];
            if (hi <= lo) continue;
            int nmax = 0;
            for (int j = lo; j < hi; j++) nmax = std::max(nmax, b->n[b->lds_ids[j]]);
            a.inst = b->lds_tab + lo;
            for (uint64_t k = 0; k < steps; k += 65536) {
                a.steps = std::min<uint64_t>(65536, steps - k);
                launch_lds_any(b->prec, TASEP_RK4, gi == 1, a, hi - lo, lds_auto_threads(nmax),
                               lds_bytes(nmax, (int)b->esize, TASEP_RK4), ls);
            }
        }
        HIP_OK(hipGetLastError());
    }
    // streamed instances: graph replays of graph_steps steps, then the rest
    if (!b->str_ids.empty() && steps) {
        uint64_t done = 0;
        const int gs = b->graph_steps;
        if (gs > 0 && steps >= (uint64_t)gs) {
            if (b->graph_built_steps != gs) {
                int rc = build_graph(b, gs);
                if (rc) return rc;
            }
            for (; done + gs <= steps; done += gs) HIP_OK(hipGraphLaunch(b->graph, b->stream));
        }
        for (; done < steps; done++) enqueue_stream_step(b, b->stream);
        HIP_OK(hipGetLastError());
    }
    if (!b->lds_ids.empty() && ls != b->stream) {
        HIP_OK(hipEventRecord(b->ev_join, ls));
        HIP_OK(hipStreamWaitEvent(b->stream, b->ev_join, 0));
    }
    HIP_OK(hipMemcpyAsync(b->st_host, b->st, (size_t)b->count * sizeof(RunState), hipMemcpyDeviceToHost,
                          b->stream));
    HIP_OK(hipStreamSynchronize(b->stream));
    if (info) {
        for (int i = 0; i < b->count; i++) {
            const RunState &s = b->st_host[i];
            info[i].steps_done = s.steps_done;
            info[i].steps_at_stop = s.steps_at_stop;
            info[i].early_stop = s.early_stop;
            info[i].stopped = s.stopped;
            info[i].t = s.t;
            info[i].last_checksum = s.checksum;
            info[i].mass = s.mass;
        }
    }
    return TASEP_OK;
}

int tasep_batch_get_stream(tasep_batch *b, void **stream) {
    if (!b || !stream) return fail(TASEP_ERR_ARG, "NULL argument");
    *stream = (void *)b->stream;
    return TASEP_OK;
}

}  // extern "C"
