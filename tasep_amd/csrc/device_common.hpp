// device_common.hpp -- wave/block reductions, run state, rate staging.
//
// Wavefront size is 64 on gfx950; every kernel here is launched with
// blockDim.x == kBlock (4 waves).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace tasep {

constexpr int kBlock = 256;
constexpr int kWave = 64;
constexpr int kWaves = kBlock / kWave;
constexpr int kMaxN = 34;
constexpr int kRateStride = 40;  // doubles per rate set (n+1 <= 35)

// Epilogues of the generator passes (kernels.hpp, tile_kernels.hpp):
// EPI_STORE: out = self + c * (L G).  EPI_HORNER: x += c * (L G) with the
// step's checksum and mass.  EPI_RK4: the staged step's combination from
// stored k1, k2, k3.  EPI_STAGE / EPI_RK4ACC (tiled passes): the staged step
// with materialised stage inputs and a running k-sum.
enum { EPI_STORE = 0, EPI_HORNER = 1, EPI_RK4 = 2, EPI_STAGE = 3, EPI_RK4ACC = 4 };

// Device-resident loop state.  Written by the end-of-step reduction (one
// thread of k_finish_step / k_batch_finish, or of the LDS-resident kernel)
// and by the host between runs; read by every kernel at entry.
struct RunState {
    double t;            // time after the last finished step
    double h;
    double thr_h;        // early-exit threshold thr*h
    double checksum;     // sum |dx| of the last step
    double mass;         // sum x after the last step
    uint64_t steps_done;
    uint64_t steps_at_stop;
    int32_t early_stop;
    int32_t patience;    // < 0: never leave early
    int32_t stopped;
    uint32_t pad0;
    double t_rates;      // start time of the next step whose rates k_tv_rates evaluates
};

// Neumaier-compensated FP64 accumulator.
struct Kahan {
    double s = 0.0, c = 0.0;
    __device__ __forceinline__ void add(double v) {
        double t = s + v;
        if (fabs(s) >= fabs(v)) c += (s - t) + v;
        else c += (v - t) + s;
        s = t;
    }
    __device__ __forceinline__ double value() const { return s + c; }
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

// Fixed-shape block sum (deterministic for a fixed launch shape) over a
// block of NT lanes.  Result is valid in every thread.  `red` must hold
// NT/64 doubles.
template <int NT = kBlock>
__device__ __forceinline__ double block_sum(double v, double *red) {
    v = wave_sum(v);
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    __syncthreads();
    if (lane == 0) red[w] = v;
    __syncthreads();
    double r = 0.0;
#pragma unroll
    for (int i = 0; i < NT / kWave; i++) r += red[i];
    return r;
}

// block_sum of two values at once: one pair of barriers, each result
// bit-identical to block_sum's.  `red` must hold 2*NT/64 doubles.
template <int NT = kBlock>
__device__ __forceinline__ void block_sum2(double &a, double &b, double *red) {
    constexpr int W = NT / kWave;
    a = wave_sum(a);
    b = wave_sum(b);
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    __syncthreads();
    if (lane == 0) {
        red[w] = a;
        red[W + w] = b;
    }
    __syncthreads();
    double ra = 0.0, rb = 0.0;
#pragma unroll
    for (int i = 0; i < W; i++) {
        ra += red[i];
        rb += red[W + i];
    }
    a = ra;
    b = rb;
}

// Advance a run state by one finished step with its checksum sum|dx| and
// mass sum x, exactly as the reference loop does
// (src/masterEquation.c:644,670-679): t += h, then the sub-threshold counter
// and the exit test.  Called by one thread.
__device__ __forceinline__ void advance_state(RunState *st, double chk, double mass) {
    st->checksum = chk;
    st->mass = mass;
    st->t = st->t + st->h;
    uint64_t i = st->steps_done;       // index of this step in the run
    st->steps_done = i + 1;
    if (st->patience >= 0) {
        if (chk < st->thr_h) {
            st->early_stop += 1;
            st->steps_at_stop = i;
        } else {
            st->early_stop = 0;
        }
        if (st->early_stop > st->patience) st->stopped = 1;
    }
}

// Fixed-order FP64 sums of nslots block partials (every thread of the
// block gets both), so the result does not depend on block scheduling.
// The partials come from an earlier launch, so plain loads see them.
template <int NT = kBlock>
__device__ __forceinline__ void reduce_partials(const double *__restrict__ part_chk,
                                                const double *__restrict__ part_mass, uint32_t nslots,
                                                double &chk, double &mass) {
    __shared__ double s_red[2 * NT / kWave];
    Kahan kc, km;
    for (uint32_t b = threadIdx.x; b < nslots; b += NT) {
        kc.add(part_chk[b]);
        km.add(part_mass[b]);
    }
    chk = kc.value();
    mass = km.value();
    block_sum2<NT>(chk, mass, s_red);
}

// Block partials of a step's final kernel: thread 0 stores the block's
// sum|dx| and sum x at slot blockIdx.x; a separate launch (k_finish_step)
// reduces them and advances the run state.  (No last-block-done count: on
// gfx950 a device-scope release in every block writes back its XCD's L2.)
__device__ __forceinline__ void store_partials(double *part_chk, double *part_mass, double blk_chk,
                                               double blk_mass) {
    if (threadIdx.x == 0) {
        part_chk[blockIdx.x] = blk_chk;
        part_mass[blockIdx.x] = blk_mass;
    }
}

// Stage the rate set `set` (n+1 doubles) into LDS.
__device__ __forceinline__ void stage_rates(double *lds, const double *rates,
                                            int set, int n) {
    if (threadIdx.x <= n) lds[threadIdx.x] = rates[set * kRateStride + threadIdx.x];
    __syncthreads();
}

// Tiles (or row groups) of one workgroup's positions P0 + k * step,
// k < count.  With a level-walk table, lane i of every wave holds the tile
// of position k0 + i,
// loaded once per chunk of 63 tiles (64 entries: the chunk's tiles and the
// next one, which a tile prefetches), and a tile is read with readlane.
// Callers walk a chunk in an inner loop with no load of the table inside
// it: a table load per tile would make the wave wait for all its
// outstanding memory operations (the previous tile's output stores, the
// prefetched next tile) at every tile.
struct WalkTiles {
    static constexpr uint64_t kChunk = 63;
    const uint32_t *walk;
    uint64_t P0, step, count, k0;
    uint32_t v;
    __device__ __forceinline__ WalkTiles(const uint32_t *w, uint64_t p0, uint64_t st, uint64_t cnt)
        : walk(w), P0(p0), step(st), count(cnt), k0(0), v(0) {
        if (walk) load(0);
    }
    __device__ __forceinline__ void load(uint64_t k) {
        k0 = k;
        if (!walk) return;
        const uint64_t kl = k + (threadIdx.x & 63);
        const uint32_t raw = kl < count ? walk[P0 + kl * step] : 0u;
        // copied once by a lane-identity DPP move: the wait for the load
        // happens here, and the loop reads a register with no load pending
        v = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)raw, 0xE4, 0xF, 0xF, false);
    }
    // position k of the loaded chunk: k0 <= k <= k0 + 63
    __device__ __forceinline__ uint64_t at(uint64_t k) const {
        if (!walk) return P0 + k * step;
        return (uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)v, (int)(k - k0));
    }
};

}  // namespace tasep
