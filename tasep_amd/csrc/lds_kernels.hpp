// lds_kernels.hpp -- whole runs of LDS-resident instances, one workgroup each.
//
// When two (reference loop) or four (RK4) 2^n-entry vectors fit in one
// workgroup's LDS (up to 160 KiB on gfx950: n <= 13, FP64 RK4 n <= 12), a
// workgroup can run every step of an instance without leaving the CU: a
// pass reads its stage input from one LDS buffer and writes the next stage
// input into another, with one barrier per pass.
// The early exit, t += h and the time-variant rates are evaluated inside the
// kernel, so a 10,000-step run of a small lattice is a single launch.  One
// launch runs any number of instances (blockIdx.x = instance), each with its
// own n, rates and run state: the building block of the batched engine.
#pragma once

#include <type_traits>

#include "device_common.hpp"
#include "generator.hpp"

namespace tasep {

struct LdsInstance {
    uint64_t off;      // element offset of the instance's state vector
    int32_t n;
    int32_t tv;        // 1: time-variant rates
    int32_t rate_off;  // constant rates at rates[rate_off .. rate_off+n]
    int32_t tv_off;    // rate functions at tv_*[tv_off .. tv_off+n]
    int32_t id;        // index of the instance's run state in LdsArgs::st
};

// The run state of each instance (t, h, thr*h, patience, step counters) is
// read from st[instance] at entry and written back at exit, exactly as the
// multi-kernel step loop keeps it, so launches can be chained.
struct LdsArgs {
    const LdsInstance *inst;
    void *x;
    const double *rates;
    const int32_t *tv_type;
    const double *tv_amp, *tv_period, *tv_offset, *tv_c;
    RunState *st;      // run states, indexed by LdsInstance::id
    uint64_t steps;    // steps of this launch (fewer if an instance stops)
    // INTEG 2 (reference loop through the stored matrix, one instance)
    const uint64_t *rowPtr, *colInd;
    const double *val;
};

// r(t) = max(0, A f(w t + phi) + c), f = sin / cos / linear
// (the reference's timeVariantProbs, src/lib/matvec.c:44-68)
__device__ __forceinline__ double tv_rate(int type, double A, double w, double phi, double c, double t) {
    double v;
    switch (type) {
    case 0: v = A * sin(t * w + phi) + c; break;
    case 1: v = A * cos(t * w + phi) + c; break;
    default: v = A * t + c; break;
    }
    return v < 0.0 ? 0.0 : v;
}

// ---------------------------------------------------------------------------
// Rows with the rates in registers.  The lanes of a wave hold 64 consecutive
// states, so bits 6 and up of s are the same in every lane: bonds k >= 6
// (bits k, k+1) are tested once per wave with a scalar branch and cost
// nothing when they do not pull into s.  Bonds k < 6 run branch-free, with
// the rate or 0 selected per lane, so their neighbour loads are independent
// and their LDS latencies overlap; the neighbour of an in-bond k (bits
// k+1,k = 0,1) is s + 2^k = s ^ (3 << k), which for every other lane is
// still inside [0, 2^n) and is multiplied by 0.
// ---------------------------------------------------------------------------
constexpr int kLdsMaxN = 13;
constexpr int kLdsBonds = kLdsMaxN - 1;
constexpr int kLaneBits = 6;  // state bits that differ between lanes of a wave

template <class A>
struct LdsRates {
    A alpha, beta;
    A bk[kLdsBonds];  // bk[k] = rate of bond k = r[n-1-k] (0 beyond n-2)
};

template <class A, class R>
__device__ __forceinline__ void load_rates(LdsRates<A> &q, const R *r, int n) {
    q.alpha = (A)r[0];
    q.beta = (A)r[n];
#pragma unroll
    for (int k = 0; k < kLdsBonds; k++) q.bk[k] = k < n - 1 ? (A)r[n - 1 - k] : (A)0;
}

__device__ __forceinline__ uint32_t wave_high_bits(uint32_t s) {
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)s);
}

// Reference diagonal, in the reference's subtraction order
// (src/lib/matgen.c:36-51): 0 - beta[b0] - alpha[!b_top] - out-bonds by k.
__device__ __forceinline__ double ref_diag_lds(int n, uint32_t s, const LdsRates<double> &q) {
    const uint32_t top = 1u << (n - 1), hi = wave_high_bits(s);
    double d = 0.0;
    if (s & 1) d -= q.beta;
    if (!(s & top)) d -= q.alpha;
#pragma unroll
    for (int k = 0; k < kLaneBits; k++)
        if (k < n - 1 && ((s >> k) & 3u) == 2u) d -= q.bk[k];
#pragma unroll
    for (int k = kLaneBits; k < kLdsBonds; k++)
        if (k < n - 1 && ((hi >> k) & 3u) == 2u) d -= q.bk[k];
    return d;
}

// Reference row: T = 0; T += v_j x_j in row order (src/lib/solv.c:8-13), no
// FMA.  A low bond that does not pull into s adds 0 * x = +-0, which leaves
// T unchanged: T starts at +0 and a round-to-nearest sum is -0 only when
// both operands are -0, so T is never -0 and T + (+-0) == T bit for bit.
template <class G>
__device__ __forceinline__ double ref_row_lds(int n, uint32_t s, const G &x, const LdsRates<double> &q,
                                              double d) {
    const uint32_t top = 1u << (n - 1), hi = wave_high_bits(s);
    double T = 0.0;
    T += ((s & top) ? q.alpha : 0.0) * x(s ^ top);
    T += d * x(s);
    T += ((s & 1) ? 0.0 : q.beta) * x(s ^ 1u);
#pragma unroll
    for (int k = 0; k < kLaneBits; k++)
        if (k < n - 1) T += (((s >> k) & 3u) == 1u ? q.bk[k] : 0.0) * x(s ^ (3u << k));
#pragma unroll
    for (int k = kLaneBits; k < kLdsBonds; k++)
        if (k < n - 1 && ((hi >> k) & 3u) == 1u) T += q.bk[k] * x(s + (1u << k));
    return T;
}

// Total outflow rate of s (minus the diagonal), arithmetic type A.
template <class A>
__device__ __forceinline__ A rk_diag_lds(int n, uint32_t s, const LdsRates<A> &q) {
    const uint32_t top = 1u << (n - 1), hi = wave_high_bits(s);
    A d = (s & 1) ? q.beta : (A)0;
    if (!(s & top)) d += q.alpha;
#pragma unroll
    for (int k = 0; k < kLaneBits; k++)
        if (k < n - 1 && ((s >> k) & 3u) == 2u) d += q.bk[k];
#pragma unroll
    for (int k = kLaneBits; k < kLdsBonds; k++)
        if (k < n - 1 && ((hi >> k) & 3u) == 2u) d += q.bk[k];
    return d;
}

// (L u)_s in type A with FMA, given the outflow rate dpos of s.
template <class A, class G>
__device__ __forceinline__ A rk_row_lds(int n, uint32_t s, const G &u, const LdsRates<A> &q, A dpos) {
    const uint32_t top = 1u << (n - 1), hi = wave_high_bits(s);
    A acc = -dpos * u(s);
    acc = fma((s & top) ? q.alpha : (A)0, u(s ^ top), acc);
    acc = fma((s & 1) ? (A)0 : q.beta, u(s ^ 1u), acc);
#pragma unroll
    for (int k = 0; k < kLaneBits; k++)
        if (k < n - 1) acc = fma(((s >> k) & 3u) == 1u ? q.bk[k] : (A)0, u(s ^ (3u << k)), acc);
#pragma unroll
    for (int k = kLaneBits; k < kLdsBonds; k++)
        if (k < n - 1 && ((hi >> k) & 3u) == 1u) acc = fma(q.bk[k], u(s + (1u << k)), acc);
    return acc;
}

// Number of 2^n-entry LDS buffers a run needs: the reference loop ping-pongs
// the state between two; RK4 keeps the state, two stage inputs and either
// the diagonal (Horner form, constant rates) or the stage accumulator
// (stored stages).
__host__ __device__ constexpr int lds_buffers(int integ) { return integ == 0 ? 2 : 4; }

// INTEG: TASEP_REF semantics (0), RK4 (1), TASEP_REF through the stored CSR
// (2: T = sum of val * x over the row in CSR order, src/lib/solv.c:8-13;
// the arrays stay in global memory, x in LDS).  Lane t handles states
// s = t, t + blockDim.x, ... < 2^n; everything an instance needs lives in
// LDS, so registers hold only the row in flight and any 2^n fits the same
// code.  TV = false: constant rates (RK4 in Horner form).  TV = true: rates
// may be time-variant per instance and RK4 keeps stored stages; a separate
// instantiation because the FP64 sin/cos it carries costs ~40 registers.
//
// One barrier per pass.  The step's checksum partials (one per wave) and
// the next step's time-variant rates are written before the step's last
// barrier into slots of the step's parity; every lane then forms the same
// checksum in the same order and takes the same exit decision, so no
// broadcast barrier is needed.
template <class V, int INTEG, bool TV>
__global__ __launch_bounds__(1024) void k_lds_run(const LdsArgs a) {
    extern __shared__ double lds_raw[];
    __shared__ double rr[2][3][kRateStride];  // [step parity][rates at t, t+h/2, t+h]
    __shared__ V rv[2][3][kRateStride];
    __shared__ double red[2][1024 / kWave];   // per-wave checksum partials by parity
    const LdsInstance in = a.inst[blockIdx.x];
    RunState *st = a.st + in.id;
    V *x = reinterpret_cast<V *>(a.x) + in.off;
    const int n = in.n, nt = blockDim.x, t = threadIdx.x;
    const int nw = nt / kWave, w = t / kWave, lane = t & (kWave - 1);
    const uint32_t D = 1u << n;
    V *xs = reinterpret_cast<V *>(lds_raw);  // the state
    V *bA = xs + D, *bB = bA + D, *bC = bB + D;
    const bool tv = TV && in.tv;
    // run state (identical in every lane)
    double tcur = st->t;
    const double h = st->h, thr_h = st->thr_h;
    const int patience = st->patience;
    uint64_t done = st->steps_done, at_stop = st->steps_at_stop;
    int early = st->early_stop, stopped = st->stopped;
    double chk = st->checksum;
    // lanes 0..n write the rates of a step starting at time t0 into slot par
    auto set_rates = [&](int par, double t0) {
        if (t > n) return;
        if (tv) {
            const int o = in.tv_off + t;
            for (int q = 0; q < (INTEG == 0 ? 1 : 3); q++) {
                const double tq = q == 0 ? t0 : (q == 1 ? t0 + 0.5 * h : t0 + h);
                const double v = tv_rate(a.tv_type[o], a.tv_amp[o], a.tv_period[o], a.tv_offset[o],
                                         a.tv_c[o], tq);
                rr[par][q][t] = v;
                rv[par][q][t] = (V)v;
            }
        } else {
            const double v = a.rates[in.rate_off + t];
            for (int q = 0; q < 3; q++) {
                rr[par][q][t] = v;
                rv[par][q][t] = (V)v;
            }
        }
    };
    set_rates(0, tcur);
    if (!tv) set_rates(1, tcur);
    for (uint32_t s = t; s < D; s += nt) xs[s] = x[s];
    __syncthreads();
    LdsRates<double> qd;
    LdsRates<V> qv;
    if (INTEG != 1) load_rates(qd, rr[0][0], n);
    else load_rates(qv, rv[0][0], n);
    if (INTEG == 1 && !TV) {  // Horner form: the diagonal once, in bC
        for (uint32_t s = t; s < D; s += nt) bC[s] = rk_diag_lds<V>(n, s, qv);
    }
    // constant-rate reference loop with at most one state per lane: the
    // diagonal once, in a register
    const bool diag_once = INTEG == 0 && !TV && D <= (uint32_t)nt;
    double d_once = 0.0;
    if (INTEG == 0 && !TV && diag_once && (uint32_t)t < D) d_once = ref_diag_lds(n, (uint32_t)t, qd);
    // reference loop (stored matrix, or constant rates) with at most one
    // state per lane: the row's nonzeros once, in registers, in row order
    // (a row has 1 + b_top + !b_0 + #in-bonds <= 3 + n/2 <= 9 entries for
    // n <= 13).  The matrix-free row leaves out the terms that do not pull
    // into s, which add 0 * x = +-0 and leave T unchanged (T is never -0).
    constexpr bool kRowRegs = INTEG == 2 || (INTEG == 0 && !TV);
    constexpr int kRow = 3 + kLdsMaxN / 2;
    const bool row_once = kRowRegs && D <= (uint32_t)nt;
    uint32_t rcol[kRowRegs ? kRow : 1];
    double rval[kRowRegs ? kRow : 1];
    int rlen = 0;
    if constexpr (INTEG == 2) {
        if (row_once && (uint32_t)t < D) {
            const uint64_t b0 = a.rowPtr[t];
            rlen = (int)(a.rowPtr[t + 1] - b0);
#pragma unroll
            for (int j = 0; j < kRow; j++) {
                rcol[j] = j < rlen ? (uint32_t)a.colInd[b0 + j] : 0u;
                rval[j] = j < rlen ? a.val[b0 + j] : 0.0;
            }
        }
    } else if constexpr (kRowRegs) {
        if (row_once) {
#pragma unroll
            for (int j = 0; j < kRow; j++) {
                rcol[j] = 0u;
                rval[j] = 0.0;
            }
            if ((uint32_t)t < D) {
                // compact the row: term number rlen goes to slot rlen
                const uint32_t s0 = (uint32_t)t, top = 1u << (n - 1);
                auto put = [&](bool on, uint32_t col, double v) {
#pragma unroll
                    for (int i = 0; i < kRow; i++)
                        if (on && i == rlen) {
                            rcol[i] = col;
                            rval[i] = v;
                        }
                    rlen += on ? 1 : 0;
                };
                put((s0 & top) != 0, s0 ^ top, qd.alpha);
                put(true, s0, d_once);
                put(!(s0 & 1u), s0 ^ 1u, qd.beta);
#pragma unroll
                for (int k = 0; k < kLdsBonds; k++)
                    put(k < n - 1 && ((s0 >> k) & 3u) == 1u, s0 + (1u << k), qd.bk[k]);
            }
        }
    }
    int par = 0;
    V *cur = xs, *nxt = bA;  // the reference loop's state alternates
    // end of a step's last pass: checksum partial and next step's rates
    auto close_step = [&](const Kahan &ck) {
        if (tv) set_rates(par ^ 1, tcur + h);
        const double c = wave_sum(ck.value());
        if (lane == 0) red[par][w] = c;
    };
    for (uint64_t k = 0; k < a.steps && !stopped; k++) {
        Kahan ck;
        if constexpr (INTEG == 0 || INTEG == 2) {
            // reference loop: x += (((k+2k)+2k)+k)/6 with k = h (L x)_s
            if (TV) load_rates(qd, rr[par][0], n);
            auto gx = [&](uint32_t j) { return cur[j]; };
            for (uint32_t s = t; s < D; s += nt) {
                double T;
                if (row_once) {
                    T = 0.0;
#pragma unroll
                    for (int j = 0; j < kRow; j++)
                        if (j < rlen) T += rval[j] * (double)cur[rcol[j]];
                } else if constexpr (INTEG == 2) {
                    T = 0.0;
                    const uint64_t e = a.rowPtr[s + 1];
                    for (uint64_t j = a.rowPtr[s]; j < e; j++) T += a.val[j] * (double)cur[a.colInd[j]];
                } else {
                    T = ref_row_lds(n, s, gx, qd, diag_once ? d_once : ref_diag_lds(n, s, qd));
                }
                const double d = ref_delta(h, T);
                nxt[s] = cur[s] + d;
                ck.add(fabs(d));
            }
            close_step(ck);
            __syncthreads();
            V *tmp = cur; cur = nxt; nxt = tmp;
        } else if constexpr (!TV) {
            // classical RK4 in Horner form: u = x + h/4 Lx, v = x + h/3 Lu,
            // u = x + h/2 Lv, x += h Lu  (passes xs->bA, bA->bB, bB->bA, bA->xs)
            const V cs[3] = {(V)(0.25 * h), (V)(h / 3.0), (V)(0.5 * h)};
#pragma unroll
            for (int p = 0; p < 3; p++) {
                const V *src = p == 0 ? xs : (p == 1 ? bA : bB);
                V *dst = p == 1 ? bB : bA;
                auto gp = [&](uint32_t j) { return src[j]; };
                for (uint32_t s = t; s < D; s += nt)
                    dst[s] = fma(cs[p], rk_row_lds<V>(n, s, gp, qv, bC[s]), xs[s]);
                __syncthreads();
            }
            auto gp = [&](uint32_t j) { return bA[j]; };
            const V hv = (V)h;
            for (uint32_t s = t; s < D; s += nt) {
                const V d = hv * rk_row_lds<V>(n, s, gp, qv, bC[s]);
                xs[s] = xs[s] + d;  // only lane-own entries of xs are touched here
                ck.add(fabs((double)d));
            }
            close_step(ck);
            __syncthreads();
        } else {
            // classical RK4 with stored stages: k_i = h L(t_i) y_i,
            // d = (k1 + 2 k2 + 2 k3 + k4) / 6  (accumulator in bC)
            const V cy[3] = {(V)0.5, (V)0.5, (V)1.0};
            const V hv = (V)h;
            for (int p = 0; p < 4; p++) {
                const V *src = p == 0 ? xs : (p == 2 ? bB : bA);
                V *dst = p == 1 ? bB : bA;
                auto gp = [&](uint32_t j) { return src[j]; };
                if (p != 2) load_rates(qv, rv[par][p == 0 ? 0 : (p == 1 ? 1 : 2)], n);
                if (p < 3) {
                    for (uint32_t s = t; s < D; s += nt) {
                        const V kk = hv * rk_row_lds<V>(n, s, gp, qv, rk_diag_lds<V>(n, s, qv));
                        bC[s] = p == 0 ? kk : bC[s] + (V)2 * kk;
                        dst[s] = fma(cy[p], kk, xs[s]);
                    }
                } else {
                    for (uint32_t s = t; s < D; s += nt) {
                        const V kk = hv * rk_row_lds<V>(n, s, gp, qv, rk_diag_lds<V>(n, s, qv));
                        const V d = (bC[s] + kk) / (V)6;
                        xs[s] = xs[s] + d;  // lane-own entries only
                        ck.add(fabs((double)d));
                    }
                    close_step(ck);
                }
                __syncthreads();
            }
        }
        // step bookkeeping, as the reference loop (src/masterEquation.c:644,670-679)
        // (a sequential sum of the wave partials measured faster than a
        // cross-lane butterfly here: 16.7 vs 18.3 ms for 10,000 steps at n=10)
        double tot = 0.0;
        for (int i = 0; i < nw; i++) tot += red[par][i];
        chk = tot;
        tcur = tcur + h;
        if (patience >= 0) {
            if (tot < thr_h) {
                early += 1;
                at_stop = done;
            } else {
                early = 0;
            }
            if (early > patience) stopped = 1;
        }
        done += 1;
        par ^= 1;
    }
    // write back the state and the run summary
    Kahan km;
    for (uint32_t s = t; s < D; s += nt) {
        const V v = cur[s];
        x[s] = v;
        km.add((double)v);
    }
    const double m = wave_sum(km.value());
    __syncthreads();
    if (lane == 0) red[0][w] = m;
    __syncthreads();
    if (t == 0) {
        double tot = 0.0;
        for (int i = 0; i < nw; i++) tot += red[0][i];
        st->t = tcur;
        st->steps_done = done;
        st->steps_at_stop = at_stop;
        st->early_stop = early;
        st->stopped = stopped;
        st->checksum = chk;
        st->mass = tot;
    }
}

}  // namespace tasep
