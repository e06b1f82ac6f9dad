// kernels.hpp -- the step, stage, CSR-build, rate and moment kernels.
#pragma once

#include "device_common.hpp"
#include "generator.hpp"

namespace tasep {

// ===========================================================================
// Reference loop (TASEP_REF): one kernel per step.  Because the reference's
// four stages all read the unmodified x, one generator pass gives k and the
// update x_new = x + (((k + 2k) + 2k) + k)/6 is fused into its epilogue
// (16 B of HBM per state: read x, write x_new; neighbours hit in cache).
// ===========================================================================
struct RefArgs {
    const double *x;
    double *xn;
    const double *rates;
    const uint64_t *rowPtr;
    const uint64_t *colInd;
    const double *val;
    RunState *st;
    double *pchk;
    double *pmass;
    uint64_t dim;
    double h;
    Lattice g;
    const uint32_t *walk;  // CSR loop: level walk of the 64-row groups, or null
};

__global__ __launch_bounds__(kBlock) void k_ref_step(const RefArgs a) {
    if (a.st->stopped) return;
    __shared__ double r[kRateStride];
    __shared__ double red[kWaves];
    stage_rates(r, a.rates, 0, a.g.n);
    Kahan chk, mass;
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t s = (uint64_t)blockIdx.x * kBlock + threadIdx.x; s < a.dim; s += stride) {
        double T = ref_row_mf(a.g, s, a.x, r);
        double d = ref_delta(a.h, T);
        double xn = a.x[s] + d;
        a.xn[s] = xn;
        chk.add(fabs(d));
        mass.add(xn);
    }
    double bc = block_sum(chk.value(), red);
    double bm = block_sum(mass.value(), red);
    store_partials(a.pchk, a.pmass, bc, bm);
}

// The reference loop through the stored CSR with coalesced matrix reads.
// One wave takes 64 consecutive rows; their entries are one contiguous
// segment of colInd/val, which the wave copies to LDS in chunks of 512
// (every lane loading 8 consecutive entries), then each lane sums its own
// row from LDS in CSR order (src/lib/solv.c:8-13) -- where lanes that each
// walked their own row in global memory touched 64 different cache lines
// per load.  A row that crosses a chunk keeps its partial sum.  Within a
// chunk a lane first gathers the x values of all its entries (independent
// loads, in flight together) and then adds the products in CSR order.
// 6 KiB of LDS per wave (n=20, 20,000 steps: 256-entry chunks 693 ms, 512 637 ms)
constexpr int kCsrChunk = 512;
// Row entries gathered together (a row has 1 + b_top + !b_0 + #in-bonds
// <= 3 + ceil((n-1)/2) entries: 6.75 on average at n = 20).
constexpr int kCsrGather = 8;

__global__ __launch_bounds__(kWave) void k_ref_csr_wave(const RefArgs a) {
    if (a.st->stopped) return;
    __shared__ uint32_t col[kCsrChunk];                           // state indices < 2^32 (n <= 32)
    __shared__ double val[kCsrChunk];
    __shared__ double red[1];
    const int t = threadIdx.x;
    Kahan chk, mass;
    const uint64_t groups = (a.dim + kWave - 1) / kWave;
    // Row groups in the level walk: the workgroups of XCD x (blockIdx % 8)
    // take positions x * groups/8 ... in turn, so a group's far partners
    // (whose x values it gathers) are walked close to it, as for the tiles.
    // Without a table: grid order.
    uint64_t P0 = blockIdx.x, step = gridDim.x;
    if (a.walk) {
        const uint64_t per = groups >> 3;
        P0 = (blockIdx.x & 7) * per + (blockIdx.x >> 3);
        step = gridDim.x >> 3;
    }
    const uint64_t lim = a.walk ? ((blockIdx.x & 7) + 1) * (groups >> 3) : groups;
    const uint64_t count = P0 < lim ? (lim - P0 + step - 1) / step : 0;
    WalkTiles tiles(a.walk, P0, step, count);
    for (uint64_t c0 = 0; c0 < count; c0 += WalkTiles::kChunk) {
      if (c0) tiles.load(c0);
      const uint64_t cend_k = c0 + WalkTiles::kChunk < count ? c0 + WalkTiles::kChunk : count;
      for (uint64_t kt = c0; kt < cend_k; kt++) {
        const uint64_t gi = tiles.at(kt);
        const uint64_t s0 = gi * kWave, s = s0 + t;
        const uint64_t last = s0 + kWave < a.dim ? s0 + kWave : a.dim;
        const uint64_t seg0 = a.rowPtr[s0], seg1 = a.rowPtr[last];
        uint64_t j = s < a.dim ? a.rowPtr[s] : seg1;
        const uint64_t je = s < a.dim ? a.rowPtr[s + 1] : seg1;
        double T = 0.0;
        for (uint64_t base = seg0; base < seg1; base += kCsrChunk) {
            const uint64_t cend = base + kCsrChunk < seg1 ? base + kCsrChunk : seg1;
            __syncthreads();  // the previous chunk is consumed
            // all loads in flight together (lanes past the chunk's end reload
            // its last entry and store nothing)
            constexpr int U = kCsrChunk / kWave;
            uint32_t cu[U];
            double vu[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const uint64_t e = base + t + (uint64_t)u * kWave;
                const uint64_t ec = e < cend ? e : cend - 1;
                cu[u] = (uint32_t)a.colInd[ec];
                vu[u] = a.val[ec];
            }
#pragma unroll
            for (int u = 0; u < U; u++) {
                const uint64_t e = base + t + (uint64_t)u * kWave;
                if (e < cend) {
                    col[e - base] = cu[u];
                    val[e - base] = vu[u];
                }
            }
            __syncthreads();
            // this lane's entries in the chunk: [j, jc) (empty when its row
            // starts in a later chunk or has ended)
            const uint64_t jc = je < cend ? je : cend;
            for (uint64_t q = j; q < jc; q += kCsrGather) {
                double xv[kCsrGather];
#pragma unroll
                for (int r = 0; r < kCsrGather; r++) xv[r] = q + r < jc ? a.x[col[q + r - base]] : 0.0;
#pragma unroll
                for (int r = 0; r < kCsrGather; r++)
                    if (q + r < jc) T += val[q + r - base] * xv[r];
            }
            if (j < jc) j = jc;
        }
        if (s < a.dim) {
            const double d = ref_delta(a.h, T);
            const double xn = a.x[s] + d;
            a.xn[s] = xn;
            chk.add(fabs(d));
            mass.add(xn);
        }
      }
    }
    const double bc = wave_sum(chk.value());
    const double bm = wave_sum(mass.value());
    (void)red;
    store_partials(a.pchk, a.pmass, bc, bm);
}

constexpr int kFinishThreads = 1024;

// A window of m steps with the exit test armed (no exit is possible before
// its last step): workgroup i sums step i's block partials (slot i of a
// ring of `stride`-sized slots); k_window_apply then advances the run state
// through the m steps in order, exactly as m end-of-step reductions would.
__global__ __launch_bounds__(kFinishThreads) void k_window_sums(const double *__restrict__ pchk,
                                                                const double *__restrict__ pmass,
                                                                uint32_t nblocks, uint32_t stride,
                                                                double *__restrict__ tot) {
    double chk, mass;
    reduce_partials<kFinishThreads>(pchk + (uint64_t)blockIdx.x * stride, pmass + (uint64_t)blockIdx.x * stride,
                                    nblocks, chk, mass);
    if (threadIdx.x == 0) {
        tot[2 * blockIdx.x] = chk;
        tot[2 * blockIdx.x + 1] = mass;
    }
}

__global__ void k_window_apply(RunState *st, const double *__restrict__ tot, uint32_t m) {
    if (threadIdx.x != 0) return;
    RunState s = *st;                  // the m steps in registers, one write back
    for (uint32_t i = 0; i < m && !s.stopped; i++) advance_state(&s, tot[2 * i], tot[2 * i + 1]);
    *st = s;
}

// End of `steps` steps: sum the last final kernel's `nblocks` block
// partials in a fixed order and advance the run state (t += h per step,
// early-exit counter and test).  steps > 1 only for runs without an exit
// test, whose intermediate checksums are not needed.
__global__ __launch_bounds__(kFinishThreads) void k_finish_step(RunState *st, const double *__restrict__ pchk,
                                                        const double *__restrict__ pmass, uint32_t nblocks,
                                                        uint32_t steps) {
    if (st->stopped) return;
    double chk, mass;
    reduce_partials<kFinishThreads>(pchk, pmass, nblocks, chk, mass);
    if (threadIdx.x == 0) {
        for (uint32_t i = 1; i < steps; i++) {
            st->t = st->t + st->h;
            st->steps_done += 1;
        }
        advance_state(st, chk, mass);
    }
}

// ===========================================================================
// Classical RK4 passes (TASEP_RK4).  One kernel template covers both forms:
//
//  staged (any rates, incl. time-variant):
//    S1: K1 = h L(t) X            S2: K2 = h L(t+h/2) (X + K1/2)
//    S3: K3 = h L(t+h/2)(X + K2/2) S4: XN = X + (K1 + 2K2 + 2K3 + h L(t+h)(X+K3))/6
//  Horner (constant rates; the same update, since for x' = Lx one RK4 step
//  is x <- (1 + hL + (hL)^2/2 + (hL)^3/6 + (hL)^4/24) x):
//    P1: U = X + h/4 L X   P2: V = X + h/3 L U   P3: U = X + h/2 L V
//    P4: X = X + h L U  (in place: X is read only at s)
//
// The gathered stage input is G(j) = GX[j] + ck*GK[j], the self term X[s].
// Rows, stage combination and update are computed in the storage type V
// (FP32 contexts: FP32, as in the tiled, LDS-resident and batch kernels);
// the checksum and mass are FP64 sums.
// The epilogue (EPI_*) is a template parameter (device_common.hpp).
// ===========================================================================

template <class V>
struct PassArgs {
    const V *gx;      // gathered vector
    const V *gk;      // optional second gathered vector
    double ck;        // weight of gk
    int use_self;     // STORE: include X[s]
    const V *x;       // state (self term)
    V *out;
    const V *k1, *k2, *k3;
    double c;         // multiplier of (L G)(s)
    const double *rates;
    int rset;
    const uint64_t *rowPtr;
    const uint64_t *colInd;
    const double *val;
    RunState *st;
    double *pchk;
    double *pmass;
    uint64_t dim;
    Lattice g;
};

template <class V>
struct Gather1 {
    const V *__restrict__ p;
    __device__ __forceinline__ V operator()(uint64_t j) const { return p[j]; }
};

template <class V>
struct Gather2 {
    const V *__restrict__ p;
    const V *__restrict__ q;
    V ck;
    __device__ __forceinline__ V operator()(uint64_t j) const { return fma(ck, q[j], p[j]); }
};

// SELF: the self term is read from the gathered vector (Horner P1, where the
// gathered vector is the state itself), so the pass streams 8 B/state in FP32.
template <class V, bool CSR, bool TWO, int EPI, bool SELF = false>
__global__ __launch_bounds__(kBlock) void k_rk_pass(const PassArgs<V> a) {
    if (a.st->stopped) return;
    __shared__ double r[kRateStride];
    __shared__ double red[kWaves];
    stage_rates(r, a.rates, a.rset, a.g.n);
    const V cmul = (V)a.c;
    Kahan chk, mass;
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t s = (uint64_t)blockIdx.x * kBlock + threadIdx.x; s < a.dim; s += stride) {
        V acc;
        if constexpr (TWO) {
            Gather2<V> G{a.gx, a.gk, (V)a.ck};
            acc = CSR ? rk_row_csr(s, a.rowPtr, a.colInd, a.val, G) : rk_row_mf<V>(a.g, s, G, r);
        } else {
            Gather1<V> G{a.gx};
            acc = CSR ? rk_row_csr(s, a.rowPtr, a.colInd, a.val, G) : rk_row_mf<V>(a.g, s, G, r);
        }
        if constexpr (EPI == EPI_STORE) {
            V o = cmul * acc;
            if constexpr (SELF) o += a.gx[s];
            else if (a.use_self) o += a.x[s];
            a.out[s] = o;
        } else {
            V d;
            if constexpr (EPI == EPI_HORNER) {
                d = cmul * acc;
            } else {
                const V k4 = cmul * acc;
                d = ((a.k1[s] + (V)2 * a.k2[s]) + (V)2 * a.k3[s] + k4) / (V)6;
            }
            const V o = a.x[s] + d;
            a.out[s] = o;
            chk.add(fabs((double)d));
            mass.add((double)o);
        }
    }
    if constexpr (EPI != EPI_STORE) {
        double bc = block_sum(chk.value(), red);
        double bm = block_sum(mass.value(), red);
        store_partials(a.pchk, a.pmass, bc, bm);
    }
}

// ===========================================================================
// Time-variant rates, FP64 on the device (replaces timeVariantProbs,
// src/lib/matvec.c:44-68): r_i(t) = max(0, A_i f(w_i t + phi_i) + c_i).
// Set 0 at the step's start time, sets 1 and 2 at t + h/2 and t + h.
// ===========================================================================
struct TvArgs {
    const int32_t *type;
    const double *amp, *period, *offset, *c;
};

__global__ void k_tv_rates(RunState *st, TvArgs tv, double *rates, int n, int nsets) {
    if (st->stopped) return;
    const int i = threadIdx.x;
    // the step's start time, on a clock of its own (t_rates += h per step,
    // the same sums as t), so that t itself may be advanced later
    const double t0 = st->t_rates, h = st->h;
    if (i <= n) {
        for (int set = 0; set < nsets; set++) {
            const double t = set == 0 ? t0 : (set == 1 ? t0 + 0.5 * h : t0 + h);
            double v;
            switch (tv.type[i]) {
            case 0: v = tv.amp[i] * sin(t * tv.period[i] + tv.offset[i]) + tv.c[i]; break;
            case 1: v = tv.amp[i] * cos(t * tv.period[i] + tv.offset[i]) + tv.c[i]; break;
            default: v = tv.amp[i] * t + tv.c[i]; break;
            }
            rates[set * kRateStride + i] = v < 0.0 ? 0.0 : v;
        }
    }
    __syncthreads();  // every lane has read t_rates
    if (i == 0) st->t_rates = t0 + h;
}

// ===========================================================================
// CSR build (replaces gencrs, src/lib/matgen.c:7-74): one lane per row, the
// row start from its closed form, entries in the reference's order with the
// diagonal subtracted in the reference's order, so all three arrays are
// bit-identical.
// ===========================================================================
__global__ __launch_bounds__(kBlock) void k_csr_build(int n, const double *__restrict__ rates,
                                                      uint64_t *__restrict__ rowPtr,
                                                      uint64_t *__restrict__ colInd,
                                                      double *__restrict__ val, uint64_t dim) {
    __shared__ double r[kRateStride];
    stage_rates(r, rates, 0, n);
    const Lattice g = make_lattice(n);
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t s = (uint64_t)blockIdx.x * kBlock + threadIdx.x; s < dim; s += stride) {
        uint64_t p = csr_row_start(n, s);
        rowPtr[s] = p;
        if (s == dim - 1) rowPtr[dim] = csr_row_start(n, dim);
        if (s & g.top) {
            colInd[p] = s - g.top;
            val[p] = r[0];
            p++;
        }
        double d = 0.0;
        if (s & 1) d -= r[n];
        if (!(s & g.top)) d -= r[0];
        uint64_t mo = ~s & (s >> 1) & g.bulk;
        while (mo) {
            int k = __builtin_ctzll(mo);
            mo &= mo - 1;
            d -= r[n - 1 - k];
        }
        colInd[p] = s;
        val[p] = d;
        p++;
        if (!(s & 1)) {
            colInd[p] = s + 1;
            val[p] = r[n];
            p++;
        }
        uint64_t mi = s & ~(s >> 1) & g.bulk;
        while (mi) {
            int k = __builtin_ctzll(mi);
            mi &= mi - 1;
            colInd[p] = s + ((uint64_t)1 << k);
            val[p] = r[n - 1 - k];
            p++;
        }
    }
}

// ===========================================================================
// Moments in one pass (replaces condense/exp_val_eval and the two
// O(n^2 2^n) covariance loops): all E_ik = sum_s b_i b_k x_s (i <= k; E_ii =
// y_i) and M = sum_s x_s.
//
// A block takes rows of 1024 states; lane t holds s = row*1024 + t + 256 j,
// j = 0..3, so bits 0-7 are the lane, bits 8-9 the register j and bits >= 10
// the row (uniform).  Per lane A = (S, P8, P9, P89) = (sum_j v_j, v_1+v_3,
// v_2+v_3, v_3); a superset-sum (zeta) transform over the 8 lane bits -- six
// rounds of wave shuffles and two through LDS -- leaves in lane m the sums
// over all lanes u with u & m == m.  Lane m = 0, 2^i, 2^i+2^k then holds the
// row's mass, site and pair sums for the low bits; pairs that involve bits
// >= 10 take the row mass (or a low-bit sum) when the row has those bits.
// Each lane owns <= 3 output quantities, accumulated with Neumaier sums;
// a second kernel adds the block partials in block order (deterministic).
// ===========================================================================
constexpr int kMomRowBits = 10;
constexpr int kMomMaxQ = kMaxN * (kMaxN + 1) / 2 + 1;  // 596
constexpr int kMomOwn = (kMomMaxQ + kBlock - 1) / kBlock;

__host__ __device__ inline int tri_index(int i, int k) {  // i <= k
    return k * (k + 1) / 2 + i;
}

template <class V>
__global__ __launch_bounds__(kBlock) void k_moments_partial(const V *__restrict__ x, int n,
                                                            uint64_t dim, uint64_t rows,
                                                            double *__restrict__ part) {
    __shared__ double zs[4 * kBlock];
    const int t = threadIdx.x;
    const int nq = n * (n + 1) / 2 + 1;
    int qidx[kMomOwn];
    uint64_t qmask[kMomOwn];
    Kahan acc[kMomOwn];
    for (int o = 0; o < kMomOwn; o++) {
        const int q = t + o * kBlock;
        qidx[o] = -1;
        qmask[o] = 0;
        if (q >= nq) continue;
        if (q == nq - 1) {  // mass
            qidx[o] = 0;
            continue;
        }
        int k = 0;
        while (tri_index(0, k + 1) <= q) k++;
        const int i = q - tri_index(0, k);
        int m = 0, p8 = 0, p9 = 0;
        uint64_t hm = 0;
        const int bits[2] = {i, k};
        for (int b = 0; b < (i == k ? 1 : 2); b++) {
            const int bit = bits[b];
            if (bit < 8) m |= 1 << bit;
            else if (bit == 8) p8 = 1;
            else if (bit == 9) p9 = 1;
            else hm |= (uint64_t)1 << (bit - kMomRowBits);
        }
        qidx[o] = (p8 && p9) ? 3 * kBlock : (p8 ? kBlock + m : (p9 ? 2 * kBlock + m : m));
        qmask[o] = hm;
    }
    for (uint64_t row = blockIdx.x; row < rows; row += gridDim.x) {
        const uint64_t base = row << kMomRowBits;
        double v[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const uint64_t s = base + t + (uint64_t)j * kBlock;
            v[j] = s < dim ? (double)x[s] : 0.0;
        }
        double S = (v[0] + v[1]) + (v[2] + v[3]);
        double P8 = v[1] + v[3], P9 = v[2] + v[3], P89 = v[3];
#pragma unroll
        for (int b = 0; b < 6; b++) {
            const double oS = __shfl_xor(S, 1 << b, kWave);
            const double o8 = __shfl_xor(P8, 1 << b, kWave);
            const double o9 = __shfl_xor(P9, 1 << b, kWave);
            const double o89 = __shfl_xor(P89, 1 << b, kWave);
            if (!(t & (1 << b))) {
                S += oS;
                P8 += o8;
                P9 += o9;
                P89 += o89;
            }
        }
        for (int b = 6; b < 8; b++) {
            __syncthreads();
            zs[t] = S;
            zs[kBlock + t] = P8;
            zs[2 * kBlock + t] = P9;
            zs[3 * kBlock + t] = P89;
            __syncthreads();
            const int u = t | (1 << b);
            if (!(t & (1 << b))) {
                S += zs[u];
                P8 += zs[kBlock + u];
                P9 += zs[2 * kBlock + u];
                P89 += zs[3 * kBlock + u];
            }
        }
        __syncthreads();
        zs[t] = S;
        zs[kBlock + t] = P8;
        zs[2 * kBlock + t] = P9;
        zs[3 * kBlock + t] = P89;
        __syncthreads();
#pragma unroll
        for (int o = 0; o < kMomOwn; o++) {
            if (qidx[o] >= 0 && (row & qmask[o]) == qmask[o]) acc[o].add(zs[qidx[o]]);
        }
    }
    for (int o = 0; o < kMomOwn; o++) {
        const int q = t + o * kBlock;
        if (q < nq) part[(uint64_t)blockIdx.x * nq + q] = acc[o].value();
    }
}

// ===========================================================================
// Window marginals (pattern probabilities behind the reference's high-order
// post-processing, src/masterEquation.c:32-143, :1422-1539, which loops over
// all 2^n states for every value it prints), for windows of WW = 4 or 6
// consecutive sites (B = 2^WW patterns):
//   W[o][p] = sum_s x_s [((s >> o) & (B-1)) == p],  o = 0..n-WW, p < B,
// plus the mass as the last quantity.  Rows of 1024 states go through LDS;
// lane t owns the bins o*B+p = t, t+256, ... with o < 10 and sums its row
// states (windows inside bits 0-9) or, for windows that reach into the row
// bits, the 2^o states of the matching low pattern when the row's bits
// match; lane j < n-WW+1-10 also owns the B bins of offset 10+j, whose
// pattern is fixed per row, and adds the row mass to one of them.  Sums run
// in a fixed order per block and k_moments_final adds the block partials in
// block order: deterministic.
// ===========================================================================
constexpr int kWin = 4, kWin6 = 6;
constexpr int kWinBins = 1 << kWin;
// Quantities of the widest window table: (kMaxN - 5) * 64 + 1.
constexpr int kWinMaxQ = (kMaxN - kWin6 + 1) * (1 << kWin6) + 1;

template <class V, int WW>
__global__ __launch_bounds__(kBlock) void k_windows_partial(const V *__restrict__ x, int n, uint64_t dim,
                                                            uint64_t rows, double *__restrict__ part) {
    constexpr int R = kMomRowBits;
    constexpr int B = 1 << WW;
    constexpr int QL = (R * B + kBlock - 1) / kBlock;    // low bins per lane
    __shared__ double zs[1 << R];
    __shared__ double hi_acc[kMaxN * B];
    __shared__ double red[kWaves];
    const int t = threadIdx.x;
    const int no = n - WW + 1;                           // window offsets
    const int nq = no * B + 1;
    const int nlow = (no < R ? no : R) * B;              // bins with o < R
    const int nhigh = no > R ? no - R : 0;               // offsets >= R
    for (int i = t; i < kMaxN * B; i += kBlock) hi_acc[i] = 0.0;
    Kahan acc[QL], macc;
    for (uint64_t row = blockIdx.x; row < rows; row += gridDim.x) {
        const uint64_t base = row << R;
        double v[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const uint64_t s = base + t + (uint64_t)j * kBlock;
            v[j] = s < dim ? (double)x[s] : 0.0;
            zs[t + j * kBlock] = v[j];
        }
        const double rm = block_sum((v[0] + v[1]) + (v[2] + v[3]), red);  // also orders zs
        macc.add(rm);
#pragma unroll
        for (int k = 0; k < QL; k++) {
            const int q = t + k * kBlock;
            if (q >= nlow) break;
            const int o = q / B, p = q % B;
            double sum = 0.0;
            if (o + WW <= R) {
                for (int i = 0; i < (1 << (R - WW)); i++) {
                    const int lo = i & ((1 << o) - 1), hi = i >> o;
                    sum += zs[(hi << (o + WW)) | (p << o) | lo];
                }
            } else {
                const int lowbits = R - o;                 // window bits inside the row
                const uint64_t hmask = ((uint64_t)1 << (WW - lowbits)) - 1;
                if ((row & hmask) == (uint64_t)(p >> lowbits)) {
                    const int pl = p & ((1 << lowbits) - 1);
                    for (int lo = 0; lo < (1 << o); lo++) sum += zs[(pl << o) | lo];
                }
            }
            acc[k].add(sum);
        }
        if (t < nhigh) hi_acc[t * B + (int)((row >> t) & (B - 1))] += rm;
        __syncthreads();  // zs is rewritten by the next row
    }
    double *out = part + (uint64_t)blockIdx.x * nq;
#pragma unroll
    for (int k = 0; k < QL; k++)
        if (t + k * kBlock < nlow) out[t + k * kBlock] = acc[k].value();
    __syncthreads();
    for (int q = t; q < nhigh * B; q += kBlock) out[R * B + q] = hi_acc[q];
    if (t == 0) out[nq - 1] = macc.value();
}

// One workgroup per quantity: lane t sums partials t, t+256, ... (Neumaier),
// then a fixed-shape tree; deterministic for a given partial count.
__global__ __launch_bounds__(kBlock) void k_moments_final(const double *__restrict__ part,
                                                          int nblocks, int nq,
                                                          double *__restrict__ out) {
    __shared__ double red[kWaves];
    const int q = blockIdx.x;
    Kahan k;
    for (int b = threadIdx.x; b < nblocks; b += kBlock) k.add(part[(uint64_t)b * nq + q]);
    const double v = block_sum(k.value(), red);
    if (threadIdx.x == 0) out[q] = v;
}

// ===========================================================================
// Small helpers.
// ===========================================================================
template <class V>
__global__ void k_fill(V *x, uint64_t dim, double v) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s < dim; s += stride)
        x[s] = (V)v;
}

template <class A, class B>
__global__ void k_convert(const A *__restrict__ src, B *__restrict__ dst, uint64_t dim) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s < dim; s += stride)
        dst[s] = (B)src[s];
}

// Product measure x_s = prod_k (b_k(s) ? p_k : 1 - p_k), k ascending, in
// FP64 (rounded once to V).
struct ProductArgs {
    double p[kMaxN];
};

template <class V>
__global__ void k_product_state(V *__restrict__ x, uint64_t dim, int n, const ProductArgs a) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t s = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; s < dim; s += stride) {
        double v = 1.0;
        for (int k = 0; k < n; k++) v *= ((s >> k) & 1) ? a.p[k] : 1.0 - a.p[k];
        x[s] = (V)v;
    }
}

}  // namespace tasep
