// batch_kernels.hpp -- one RK4 pass over many instances in one launch.
//
// The instances of a batch that do not fit one workgroup's LDS (n >= 14)
// share every launch: their vectors are concatenated in HBM and a block
// table maps each workgroup to (instance, block of that instance).  A
// workgroup of kBatchThreads lanes covers kBatchThreads * E consecutive
// states of one instance; lane t holds states base + j*kBatchThreads + t, so
// a wave holds 64 consecutive states and
//   - exit and bonds k < 5 pull from inside the wave (cross-lane swizzles),
//   - bond 5 pulls from the neighbouring 64-state window (one coalesced load),
//   - bonds k >= 6 and the entry bit are the same for the whole wave, so
//     they are scalar branches and only real in-bonds load anything.
// Constant-rate instances take RK4 in Horner form (passes x->u, u->v, v->u,
// u->x), time-variant ones RK4 with stored stages (accumulator vector acc,
// rates evaluated in FP64 per block at the pass's time).  The last pass
// leaves per-block sums of |dx| and x; k_batch_finish reduces them per
// instance and advances each run state (t, early exit) on the device, so a
// captured sequence of steps replays without host work.
#pragma once

#include "device_common.hpp"
#include "lds_kernels.hpp"
#include "tile_kernels.hpp"

namespace tasep {

constexpr int kBatchThreads = 256;
constexpr int kBatchMaxN = 30;

template <class V> struct BatchGeom;
template <> struct BatchGeom<float> { static constexpr int E = 4; };
template <> struct BatchGeom<double> { static constexpr int E = 2; };

struct BatchInst {
    uint64_t off;       // element offset of the instance in x, u, v, acc
    int32_t n, tv;
    int32_t rate_off;   // rates[rate_off .. rate_off+n]
    int32_t tv_off;     // tv_*[tv_off .. tv_off+n]
    uint32_t part_off;  // first per-block partial slot of the instance
    uint32_t nblocks;   // workgroups of one pass over the instance
    int32_t id;         // index of the instance's run state in BatchArgs::st
};

struct BatchBlock {
    uint32_t inst;  // 0xffffffff: padding, no work
    uint32_t blk;
};

template <class V>
struct BatchArgs {
    const BatchInst *inst;
    const BatchBlock *blocks;  // one entry per workgroup
    V *x, *u, *v, *acc;
    const double *rates;
    const int32_t *tv_type;
    const double *tv_amp, *tv_period, *tv_offset, *tv_c;
    RunState *st;              // run states, indexed by BatchInst::id
    double *pchk, *pmass;      // per-block partials
    int pass;                  // 0..3
};

// Rates of one pass in registers: the entry/exit rates and bonds k < 6 per
// lane; bonds k >= 6 are read from LDS inside their (uniform) branch.
template <class A>
struct BatchRates {
    A alpha, beta, bk[kLaneBits];
};

template <class V, class A>
__device__ __forceinline__ A batch_diag(int n, uint32_t s, uint32_t hi, const BatchRates<A> &q, const V *rs) {
    A d = (s & 1) ? q.beta : (A)0;
    if (!(hi & (1u << (n - 1)))) d += q.alpha;
#pragma unroll
    for (int k = 0; k < kLaneBits; k++)
        if (((s >> k) & 3u) == 2u) d += q.bk[k];
    for (int k = kLaneBits; k < n - 1; k++)
        if (((hi >> k) & 3u) == 2u) d += (A)rs[n - 1 - k];
    return d;
}

// (L g)_s: own = g[s], g = the instance's gathered vector.
template <class V, class A>
__device__ __forceinline__ A batch_row(int n, uint32_t s, uint32_t hi, const V *__restrict__ g, A own,
                                       const BatchRates<A> &q, const V *rs, A dpos) {
    A acc = -dpos * own;
    const uint32_t top = 1u << (n - 1);
    if (hi & top) acc = fma(q.alpha, (A)g[s - top], acc);
    acc = fma((s & 1) ? (A)0 : q.beta, (A)__shfl_xor(own, 1, kWave), acc);
#pragma unroll
    for (int k = 0; k < kLaneBits - 1; k++)
        acc = fma(((s >> k) & 3u) == 1u ? q.bk[k] : (A)0, (A)__shfl_xor(own, 3 << k, kWave), acc);
    // bond 5: bits 5 (per lane) and 6 (per wave); s ^ (3 << 5) is in the
    // neighbouring window and inside [0, 2^n) for every lane
    acc = fma(((s >> 5) & 3u) == 1u ? q.bk[5] : (A)0, (A)g[s ^ (3u << 5)], acc);
    for (int k = kLaneBits; k < n - 1; k++)
        if (((hi >> k) & 3u) == 1u) acc = fma((A)rs[n - 1 - k], (A)g[s + (1u << k)], acc);
    return acc;
}

template <class V>
__global__ __launch_bounds__(kBatchThreads) void k_batch_pass(const BatchArgs<V> a) {
    using A = V;  // FP32 instances compute in FP32, FP64 ones in FP64
    constexpr int E = BatchGeom<V>::E;
    constexpr uint32_t B = kBatchThreads * E;
    __shared__ V rs[kRateStride];
    __shared__ double red[2 * kBatchThreads / kWave];
    const BatchBlock bb = a.blocks[blockIdx.x];
    if (bb.inst == 0xffffffffu) return;
    const BatchInst in = a.inst[bb.inst];
    RunState *st = a.st + in.id;
    if (st->stopped) return;
    const int n = in.n, p = a.pass, t = threadIdx.x;
    const double h = st->h;
    // rates of this pass: constant, or r(t), r(t+h/2), r(t+h) for stages 1, 2-3, 4
    if (t <= n) {
        double r;
        if (in.tv) {
            const int o = in.tv_off + t;
            const double t0 = st->t;
            const double tq = p == 0 ? t0 : (p < 3 ? t0 + 0.5 * h : t0 + h);
            r = tv_rate(a.tv_type[o], a.tv_amp[o], a.tv_period[o], a.tv_offset[o], a.tv_c[o], tq);
        } else {
            r = a.rates[in.rate_off + t];
        }
        rs[t] = (V)r;
    }
    __syncthreads();
    BatchRates<A> q;
    q.alpha = (A)rs[0];
    q.beta = (A)rs[n];
#pragma unroll
    for (int k = 0; k < kLaneBits; k++) q.bk[k] = (A)rs[n - 1 - k];
    V *x = a.x + in.off, *u = a.u + in.off, *v = a.v + in.off, *kacc = a.acc + in.off;
    const V *g = p == 0 ? x : (p == 2 ? v : u);
    V *dst = p == 1 ? v : u;
    const uint32_t base = bb.blk * B;
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(base + (t & ~(kWave - 1))));
    Kahan ck, km;
    if (!in.tv) {
        const A cs = p == 0 ? (A)(0.25 * h) : (p == 1 ? (A)(h / 3.0) : (p == 2 ? (A)(0.5 * h) : (A)h));
#pragma unroll
        for (int j = 0; j < E; j++) {
            const uint32_t s = base + j * kBatchThreads + t;
            const uint32_t hj = hi + j * kBatchThreads;
            const A own = (A)g[s];
            const A lg = batch_row<V, A>(n, s, hj, g, own, q, rs, batch_diag<V, A>(n, s, hj, q, rs));
            const A xs = p == 0 ? own : (A)x[s];
            if (p < 3) {
                dst[s] = (V)fma(cs, lg, xs);
            } else {
                const A d = cs * lg;
                const V xn = (V)(xs + d);
                x[s] = xn;
                ck.add(fabs((double)d));
                km.add((double)xn);
            }
        }
    } else {
        const A hv = (A)h;
        const A cy = p == 2 ? (A)1 : (A)0.5;
#pragma unroll
        for (int j = 0; j < E; j++) {
            const uint32_t s = base + j * kBatchThreads + t;
            const uint32_t hj = hi + j * kBatchThreads;
            const A own = (A)g[s];
            const A kk = hv * batch_row<V, A>(n, s, hj, g, own, q, rs, batch_diag<V, A>(n, s, hj, q, rs));
            const A xs = p == 0 ? own : (A)x[s];
            if (p == 0) {
                kacc[s] = (V)kk;
                dst[s] = (V)fma(cy, kk, xs);
            } else if (p < 3) {
                kacc[s] = (V)((A)kacc[s] + (A)2 * kk);
                dst[s] = (V)fma(cy, kk, xs);
            } else {
                const A d = ((A)kacc[s] + kk) / (A)6;
                const V xn = (V)(xs + d);
                x[s] = xn;
                ck.add(fabs((double)d));
                km.add((double)xn);
            }
        }
    }
    if (p == 3) {
        double bc = ck.value(), bm = km.value();
        block_sum2<kBatchThreads>(bc, bm, red);
        if (t == 0) {
            a.pchk[in.part_off + bb.blk] = bc;
            a.pmass[in.part_off + bb.blk] = bm;
        }
    }
}

// After the last pass of a step: one workgroup per streamed instance sums
// its block partials in a fixed order and advances its run state.  (A
// separate launch instead of a last-block-done count: on gfx950 each
// block's device-scope release would write back its XCD's L2.)
__global__ __launch_bounds__(kBatchThreads) void k_batch_finish(const BatchInst *__restrict__ inst,
                                                                RunState *__restrict__ st_all,
                                                                const double *__restrict__ pchk,
                                                                const double *__restrict__ pmass) {
    const BatchInst in = inst[blockIdx.x];
    RunState *st = st_all + in.id;
    if (st->stopped) return;
    double chk, mass;
    reduce_partials<kBatchThreads>(pchk + in.part_off, pmass + in.part_off, in.nblocks, chk, mass);
    if (threadIdx.x == 0) advance_state(st, chk, mass);
}

// ===========================================================================
// Tile form of the batched pass.  A 128-lane workgroup takes one tile of
// 2^TL consecutive states of one instance (TL = 10 FP32, 9 FP64; 4 KiB of
// LDS) and lane t holds local states t + 128 j.  Small workgroups keep many
// tiles in flight per CU, which hides the latency of each tile's chain of
// loads (512 instances, n in 8..18, half with time-variant rates, 1000
// FP32 steps: 1024 lanes 494 ms, 512 510, 256 445,
// 128 431, 64 478).  Bonds inside the tile come
// from LDS (bonds < 6 branch-free per lane, bonds 6..TL-2 wave-uniform);
// the bonds that cross tiles are uniform over the workgroup: the seam bond
// TL-1 (local bit TL-1, tile bit 0: half a tile of T+1), the tile bonds
// (tile T + 2^(k-TL), two streams in flight) and the entry (tile T - 2^(n-1-TL)),
// each a coalesced stream read only when it pulls into the tile.
// ===========================================================================
constexpr int kBatchTileThreads = 128;
template <class V> struct BatchTileGeom;
template <> struct BatchTileGeom<float> { static constexpr int TL = 10, J = 8; };
template <> struct BatchTileGeom<double> { static constexpr int TL = 9, J = 4; };

template <class V>
__global__ __launch_bounds__(kBatchTileThreads, 8) void k_batch_tile(const BatchArgs<V> a) {
    using A = V;
    constexpr int TL = BatchTileGeom<V>::TL, J = BatchTileGeom<V>::J;
    constexpr uint32_t TS = 1u << TL;
    constexpr int NT = kBatchTileThreads;
    __shared__ V tile[TS];
    __shared__ V rs[kRateStride];
    __shared__ double red[2 * NT / kWave];
    const BatchBlock bb = a.blocks[blockIdx.x];
    if (bb.inst == 0xffffffffu) return;
    const BatchInst in = a.inst[bb.inst];
    RunState *st = a.st + in.id;
    if (st->stopped) return;
    const int n = in.n, p = a.pass, t = threadIdx.x;
    const double h = st->h;
    if (t <= n) {
        double r;
        if (in.tv) {
            const int o = in.tv_off + t;
            const double t0 = st->t;
            const double tq = p == 0 ? t0 : (p < 3 ? t0 + 0.5 * h : t0 + h);
            r = tv_rate(a.tv_type[o], a.tv_amp[o], a.tv_period[o], a.tv_offset[o], a.tv_c[o], tq);
        } else {
            r = a.rates[in.rate_off + t];
        }
        rs[t] = (V)r;
    }
    V *x = a.x + in.off, *u = a.u + in.off, *v = a.v + in.off, *kacc = a.acc + in.off;
    const V *g = p == 0 ? x : (p == 2 ? v : u);
    V *dst = p == 1 ? v : u;
    const uint32_t T = bb.blk;
    const uint32_t base = T << TL;
    const int ntb = n - TL;                                  // tile-index bits (>= 1)
    // ---- own values, self term and the first uniform streams
    A own[J], xs[J];
#pragma unroll
    for (int j = 0; j < J; j++) {
        own[j] = (A)g[base + t + j * NT];
        xs[j] = p == 0 ? own[j] : (A)x[base + t + j * NT];
    }
    const bool seam_on = !(T & 1u);                          // bond TL-1 pulls from tile T+1
    A seam[J / 2];
    if (seam_on) {
#pragma unroll
        for (int j = 0; j < J / 2; j++) seam[j] = (A)g[base + TS + t + j * NT];
    }
    const uint32_t tmask = ntb > 1 ? ((1u << (ntb - 1)) - 1) : 0u;  // tile bonds TL..n-2
    uint32_t mm = T & ~(T >> 1) & tmask;
    int kA = mm ? TL + __builtin_ctz(mm) : -1;
    if (mm) mm &= mm - 1;
    A sA[J], sB[J];
    if (kA >= 0) {
#pragma unroll
        for (int j = 0; j < J; j++) sA[j] = (A)g[base + (1u << kA) + t + j * NT];
    }
#pragma unroll
    for (int j = 0; j < J; j++) tile[t + j * NT] = (V)own[j];
    __syncthreads();
    BatchRates<A> q;
    q.alpha = (A)rs[0];
    q.beta = (A)rs[n];
#pragma unroll
    for (int k = 0; k < kLaneBits; k++) q.bk[k] = (A)rs[n - 1 - k];
    // ---- uniform parts of the outflow: tile bonds, entry
    const bool t_top = (T >> (ntb - 1)) & 1u;
    const uint32_t hout = ~T & (T >> 1) & tmask;
    A d_tile = t_top ? (A)0 : q.alpha;
    for (uint32_t m = hout; m; m &= m - 1) d_tile += (A)rs[n - 1 - (TL + __builtin_ctz(m))];
    const A r_seam = (A)rs[n - TL];                          // bond TL-1
    // ---- in-tile rows
    A L[J];
#pragma unroll
    for (int j = 0; j < J; j++) {
        const uint32_t loc = t + j * NT;
        const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(loc & ~63u));
        A d = d_tile + ((loc & 1) ? q.beta : (A)0);
#pragma unroll
        for (int k = 0; k < kLaneBits; k++)
            if (((loc >> k) & 3u) == 2u) d += q.bk[k];
        for (int k = kLaneBits; k < TL - 1; k++)
            if (((hi >> k) & 3u) == 2u) d += (A)rs[n - 1 - k];
        if (!(loc & (TS >> 1)) && (T & 1u)) d += r_seam;    // seam out-flow
        A acc = -d * own[j];
        acc = fma((loc & 1) ? (A)0 : q.beta, (A)tile[loc ^ 1u], acc);
#pragma unroll
        for (int k = 0; k < kLaneBits; k++)
            acc = fma(((loc >> k) & 3u) == 1u ? q.bk[k] : (A)0, (A)tile[loc ^ (3u << k)], acc);
        for (int k = kLaneBits; k < TL - 1; k++)
            if (((hi >> k) & 3u) == 1u) acc = fma((A)rs[n - 1 - k], (A)tile[loc + (1u << k)], acc);
        L[j] = acc;
    }
    // ---- cross-tile in-flow: seam, tile bonds (two streams in flight), entry
    if (seam_on) {
#pragma unroll
        for (int j = J / 2; j < J; j++) L[j] = fma(r_seam, seam[j - J / 2], L[j]);
    }
    while (kA >= 0) {
        const int kB = mm ? TL + __builtin_ctz(mm) : -1;
        if (mm) mm &= mm - 1;
        if (kB >= 0) {
#pragma unroll
            for (int j = 0; j < J; j++) sB[j] = (A)g[base + (1u << kB) + t + j * NT];
        }
        {
            const A r = (A)rs[n - 1 - kA];
#pragma unroll
            for (int j = 0; j < J; j++) L[j] = fma(r, sA[j], L[j]);
        }
        if (kB < 0) break;
        kA = mm ? TL + __builtin_ctz(mm) : -1;
        if (mm) mm &= mm - 1;
        if (kA >= 0) {
#pragma unroll
            for (int j = 0; j < J; j++) sA[j] = (A)g[base + (1u << kA) + t + j * NT];
        }
        {
            const A r = (A)rs[n - 1 - kB];
#pragma unroll
            for (int j = 0; j < J; j++) L[j] = fma(r, sB[j], L[j]);
        }
    }
    if (t_top) {
        const uint32_t eo = base - (1u << (n - 1));
#pragma unroll
        for (int j = 0; j < J; j++) L[j] = fma(q.alpha, (A)g[eo + t + j * NT], L[j]);
    }
    // ---- epilogue: Horner or stored stages
    Kahan ck, km;
    if (!in.tv) {
        const A cs = p == 0 ? (A)(0.25 * h) : (p == 1 ? (A)(h / 3.0) : (p == 2 ? (A)(0.5 * h) : (A)h));
#pragma unroll
        for (int j = 0; j < J; j++) {
            const uint32_t s = base + t + j * NT;
            if (p < 3) {
                dst[s] = (V)fma(cs, L[j], xs[j]);
            } else {
                const A d = cs * L[j];
                const V xn = (V)(xs[j] + d);
                x[s] = xn;
                ck.add(fabs((double)d));
                km.add((double)xn);
            }
        }
    } else {
        const A hv = (A)h;
        const A cy = p == 2 ? (A)1 : (A)0.5;
#pragma unroll
        for (int j = 0; j < J; j++) {
            const uint32_t s = base + t + j * NT;
            const A kk = hv * L[j];
            if (p == 0) {
                kacc[s] = (V)kk;
                dst[s] = (V)fma(cy, kk, xs[j]);
            } else if (p < 3) {
                kacc[s] = (V)((A)kacc[s] + (A)2 * kk);
                dst[s] = (V)fma(cy, kk, xs[j]);
            } else {
                const A d = ((A)kacc[s] + kk) / (A)6;
                const V xn = (V)(xs[j] + d);
                x[s] = xn;
                ck.add(fabs((double)d));
                km.add((double)xn);
            }
        }
    }
    if (p == 3) {
        double bc = ck.value(), bm = km.value();
        block_sum2<NT>(bc, bm, red);
        if (t == 0) {
            a.pchk[in.part_off + bb.blk] = bc;
            a.pmass[in.part_off + bb.blk] = bm;
        }
    }
}

// ===========================================================================
// Big-tile form of the batched pass: the single-lattice tile pass
// (tile_pass_from: 512 lanes, tiles of 2^13 FP32 / 2^12 FP64 states with the
// top three state bits in the tile, every global read of a tile issued before
// its LDS phase and the next tile's own vector prefetched) over a range of
// consecutive tiles of one instance per workgroup.  Needs n >= 14 (FP32) or
// n >= 13 (FP64): two tiles.  Constant-rate instances take the Horner passes,
// time-variant ones the staged passes with materialised stage inputs (u, v)
// and the running k-sum in acc, the last of them writing x in place (it
// gathers u, so x is read only at each state's own slot).  The arithmetic is
// that of a single context's tiled run, so every instance's state is the
// same, bit for bit.
// ===========================================================================
struct BatchRange {
    uint32_t inst;   // 0xffffffff: padding, no work
    uint32_t t0;     // first tile of the range
    uint32_t count;  // consecutive tiles
    uint32_t part;   // partial-sum slot of the workgroup
};

template <class V>
struct BatchBigArgs {
    const BatchInst *inst;
    const BatchRange *ranges;  // one entry per workgroup
    V *x, *u, *v, *acc;
    const double *rates;
    const int32_t *tv_type;
    const double *tv_amp, *tv_period, *tv_offset, *tv_c;
    RunState *st;
    double *pchk, *pmass;
};

template <class V, int P, bool NT>
__global__ __launch_bounds__(kTileThreads, TASEP_TILE_WAVES) void k_batch_bigtile(const BatchBigArgs<V> b) {
    using G = TileGeom<V>;
    const BatchRange r = b.ranges[blockIdx.x];
    if (r.inst == 0xffffffffu) return;
    const BatchInst in = b.inst[r.inst];
    __shared__ V tile[kTileThreads * G::C * G::E];
    __shared__ V rr[kRateStride];
    __shared__ double red[2 * kTileWaves];
    const int n = in.n, t = threadIdx.x;
    // Issue order once the instance is known: its rate parameters, the
    // range's first tile (the gathered vector of pass P is x, u, v, u in
    // both RK4 forms), then the run state (scalar loads), so that a
    // workgroup waits on one memory latency before its LDS phase.
    double q = 0.0, amp = 0.0, per = 0.0, ofs = 0.0, cc = 0.0;
    int typ = 0;
    if (t <= n) {
        if (in.tv) {
            const int o = in.tv_off + t;
            typ = b.tv_type[o];
            amp = b.tv_amp[o];
            per = b.tv_period[o];
            ofs = b.tv_offset[o];
            cc = b.tv_c[o];
        } else {
            q = b.rates[in.rate_off + t];
        }
    }
    V *x = b.x + in.off, *u = b.u + in.off, *v = b.v + in.off;
    TileArgs<V> a = {};
    a.gx = P == 0 ? x : (P == 2 ? v : u);
    Cols<V, G::C, G::E> first;
    issue_cols<V, G::E, G::C, G::LB>(a, ((uint64_t)r.t0 << G::TB0) + ((uint64_t)__builtin_amdgcn_readfirstlane(t >> 6) << (n - 3)),
                                      (uint32_t)(t & 63) * G::E, first);
    RunState *st = b.st + in.id;
    if (st->stopped) return;
    const double h = st->h;
    // rates of this pass: constant, or r(t), r(t+h/2), r(t+h) for stages 1, 2-3, 4
    if (t <= n) {
        if (in.tv) {
            const double t0 = st->t;
            const double tq = P == 0 ? t0 : (P < 3 ? t0 + 0.5 * h : t0 + h);
            q = tv_rate(typ, amp, per, ofs, cc, tq);
        }
        rr[t] = (V)q;
    }
    __syncthreads();
    a.n = n;
    a.ntiles = (uint64_t)1 << (n - G::b);
    a.st = st;
    a.pchk = b.pchk;
    a.pmass = b.pmass;
    a.use_self = 1;
    a.x = x;
    if (!in.tv) {
        // Horner: u = x + h/4 Lx, v = x + h/3 Lu, u = x + h/2 Lv, x += h Lu
        if constexpr (P == 0) {
            a.gx = x; a.out = u; a.c = (V)(0.25 * h);
            tile_pass_from<V, EPI_STORE, true, NT>(a, rr, tile, red, WalkTiles(nullptr, r.t0, 1, r.count), r.part, first);
        } else if constexpr (P == 1) {
            a.gx = u; a.out = v; a.c = (V)(h / 3.0);
            tile_pass_from<V, EPI_STORE, false, NT>(a, rr, tile, red, WalkTiles(nullptr, r.t0, 1, r.count), r.part, first);
        } else if constexpr (P == 2) {
            a.gx = v; a.out = u; a.c = (V)(0.5 * h);
            tile_pass_from<V, EPI_STORE, false, NT>(a, rr, tile, red, WalkTiles(nullptr, r.t0, 1, r.count), r.part, first);
        } else {
            a.gx = u; a.out = x; a.c = (V)h;
            tile_pass_from<V, EPI_HORNER, false, NT>(a, rr, tile, red, WalkTiles(nullptr, r.t0, 1, r.count), r.part, first);
        }
    } else {
        V *kacc = b.acc + in.off;
        a.c = (V)h;
        a.acc_in = kacc;
        a.acc_out = kacc;
        if constexpr (P == 0) {
            a.gx = x; a.out = u; a.cy = (V)0.5; a.stage = 0;
            tile_pass_from<V, EPI_STAGE, true, NT>(a, rr, tile, red, WalkTiles(nullptr, r.t0, 1, r.count), r.part, first);
        } else if constexpr (P == 1) {
            a.gx = u; a.out = v; a.cy = (V)0.5; a.stage = 1;
            tile_pass_from<V, EPI_STAGE, false, NT>(a, rr, tile, red, WalkTiles(nullptr, r.t0, 1, r.count), r.part, first);
        } else if constexpr (P == 2) {
            a.gx = v; a.out = u; a.cy = (V)1.0; a.stage = 2;
            tile_pass_from<V, EPI_STAGE, false, NT>(a, rr, tile, red, WalkTiles(nullptr, r.t0, 1, r.count), r.part, first);
        } else {
            a.gx = u; a.out = x;
            tile_pass_from<V, EPI_RK4ACC, false, NT>(a, rr, tile, red, WalkTiles(nullptr, r.t0, 1, r.count), r.part, first);
        }
    }
}

}  // namespace tasep
