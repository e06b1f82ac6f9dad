// tile_kernels.hpp -- matrix-free generator passes on 2-D state tiles.
//
// A workgroup of 512 lanes (8 waves) owns a tile of 8 chunks; chunk w (wave
// w) holds C = 4 columns of 64*E consecutive states (E = 16/sizeof(V) per
// lane per column, one 16-byte load), the columns 2^LB states apart, and the
// 8 chunks differ only in the top three state bits.  With e = log2 E,
// LB = e + 6 and TB0 = LB + 2 the bits of a state are
//   bits 0 .. e-1        register index j
//   bits e .. LB-1       lane l (6 bits)
//   bits LB, LB+1        column c (registers)
//   bits TB0 .. n-4      tile index T (uniform over the workgroup)
//   bits n-3 .. n-1      wave w (the chunk)
// so that the entry site and the two bonds at the top of the lattice --
// whose neighbours are 2^(n-1) and 2^(n-2) states away, too far for any
// cache -- are exchanged through LDS.  Bond k (bits k, k+1):
//   k <= e-2          both register bits: static, no loads
//   k == e-1          register bit + lane bit 0: next lane via LDS
//   e <= k <= LB-2    lane bits: per-lane weight, lane l + 2^(k-e) via LDS
//   k == LB-1         lane bit 5 + column bit 0: lane l-32 of column c+1 (LDS)
//   k == LB           column bits: static, register column c+1
//   k == LB+1         column bit 1 + tile bit 0: columns 2, 3 read columns
//                     0, 1 of tile T+1 (global, when the tile bit is 0)
//   TB0 <= k <= n-5   tile bits: uniform, the whole tile reads tile
//                     T + 2^(k-TB0) (coalesced 16-byte loads)
//   k == n-4          tile top bit + wave bit 0: per-wave, chunk w+1 of tile
//                     T - 2^(n-4-TB0), 2^(n-4) states away
//   k == n-3, n-2     wave bits: chunk w+1 / w+2 of the same tile via LDS
//   entry (n-1)       wave bit 2: chunk w-4 via LDS
// The diagonal splits the same way (static, per-lane, per-wave, per-tile).
// Every global neighbour access is a contiguous stream at most 2^(n-4)
// states away; only active tile bonds are read, and the per-tile uniform
// work is shared by 16 states per lane.
//
// Memory-level parallelism: the workgroup is persistent over its tiles.
// Before a tile's LDS phase it issues every global read the tile needs, in
// the order they are consumed -- the seam, bond n-4, the first
// TASEP_TILE_SLOTS tile bonds, the self vector (and the running k-sum of
// the staged passes) and the next tile's own vector -- so their latency
// overlaps the LDS work; further tile bonds are read two at a time
// afterwards.
//
// Tiles are walked in an XCD-aware order (tile_walk): workgroup g runs on
// XCD g % 8, and each XCD sweeps its own eighth of the walk positions in
// order; by default the positions follow the level walk (WalkTiles), which
// places a tile-bond partner one group after its reader.
#pragma once

#include "device_common.hpp"
#include "generator.hpp"

namespace tasep {

constexpr int kTileThreads = 512;
// Waves per SIMD the tile passes are compiled for (register budget 512/W).
#ifndef TASEP_TILE_WAVES
#define TASEP_TILE_WAVES 2
#endif
constexpr int kTileWaves = kTileThreads / kWave;
// Tile-bond streams issued per tile before its LDS phase (16 registers per
// stream and lane); a tile with more active tile bonds reads the rest two at
// a time afterwards.  Per step, one against two: RK4 FP32 n=28 2.825 /
// 2.851 ms, n=29..31 0.8-1.0% faster, n=32, 33 and n=20..26 within +-1%;
// RK4 FP64 n=28 5.893 / 5.947 ms, n=20 1.4% faster, n=22 3% slower.
#ifndef TASEP_TILE_SLOTS
#define TASEP_TILE_SLOTS 1
#endif
// Order within a tile: 1 = the prefetched own vector goes to LDS before the
// tile's streams are issued, 0 = after, for FP32 and FP64 tiles.  Measured
// per step at n=28 (first / after): RK4 FP32 2.89 / 3.07 ms (P2/P3 0.741 /
// 0.783 ms), RK4 FP64 6.10 / 6.02 ms, reference loop 1.634 / 1.586 ms.
#ifndef TASEP_OWN_FIRST_F32
#define TASEP_OWN_FIRST_F32 1
#endif
#ifndef TASEP_OWN_FIRST_F64
#define TASEP_OWN_FIRST_F64 0
#endif
// The same for the reference-loop tiles.
#ifndef TASEP_REF_SLOTS
#define TASEP_REF_SLOTS 1
#endif

template <class V> struct TileGeom;
// E states per lane per column (one 16-byte load), C = 4 columns per lane.
// Tile bits start at TB0 = LB + 2; a tile holds 2^b = 2^(TB0 + 3) states.
template <> struct TileGeom<float> {
    static constexpr int E = 4, e = 2, LB = 8, C = 4, TB0 = LB + 2, b = TB0 + 3;  // 8192 states
};
template <> struct TileGeom<double> {
    static constexpr int E = 2, e = 1, LB = 7, C = 4, TB0 = LB + 2, b = TB0 + 3;  // 4096 states
};

__device__ __forceinline__ void ld16(const float *__restrict__ p, float (&o)[4]) {
    const float4 a = *reinterpret_cast<const float4 *>(p);
    o[0] = a.x; o[1] = a.y; o[2] = a.z; o[3] = a.w;
}
__device__ __forceinline__ void ld16(const double *__restrict__ p, double (&o)[2]) {
    const double2 a = *reinterpret_cast<const double2 *>(p);
    o[0] = a.x; o[1] = a.y;
}
__device__ __forceinline__ void st16(float *__restrict__ p, const float (&o)[4]) {
    *reinterpret_cast<float4 *>(p) = make_float4(o[0], o[1], o[2], o[3]);
}
__device__ __forceinline__ void st16(double *__restrict__ p, const double (&o)[2]) {
    *reinterpret_cast<double2 *>(p) = make_double2(o[0], o[1]);
}

// Streaming (non-temporal) variants for vectors touched once per pass: the
// self-term read of x and the output write.  Keeping them out of the way in
// L2 leaves its capacity to the gathered vector, which neighbouring tiles
// re-read.  The output's policy is a template parameter (NT) chosen per
// launch: a vector that the memory-side cache can keep until the next pass
// reads it is written with ordinary stores.
#ifndef TASEP_NT_SELF
#define TASEP_NT_SELF 1
#endif
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void ld16_nt(const float *p, float (&o)[4]) {
    const f32x4 a = __builtin_nontemporal_load(reinterpret_cast<const f32x4 *>(p));
    o[0] = a.x; o[1] = a.y; o[2] = a.z; o[3] = a.w;
}
__device__ __forceinline__ void ld16_nt(const double *p, double (&o)[2]) {
    const f64x2 a = __builtin_nontemporal_load(reinterpret_cast<const f64x2 *>(p));
    o[0] = a.x; o[1] = a.y;
}
template <class V, int E>
__device__ __forceinline__ void ld_self(const V *p, V (&o)[E]) {
    if constexpr (TASEP_NT_SELF) ld16_nt(p, o);
    else ld16(p, o);
}
// Output stores address a wave's 4 KB chunk through a buffer descriptor of
// the chunk's (wave-uniform) base and a 32-bit lane offset, instead of a
// 64-bit address per lane; NT sets the non-temporal bit, the policy of
// __builtin_nontemporal_store.  Measured at n=28 against 64-bit addresses
// (per step, three runs each): RK4 FP32 2.876 -> 2.858 ms (the
// self-gathering pass 0.638 -> 0.623 ms per launch), reference loop 1.403 ->
// 1.398 ms, RK4 FP64 unchanged; the same bits land at the same addresses.
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ __amdgpu_buffer_rsrc_t chunk_rsrc(void *base) {
    // word 3: the gfx9 raw-buffer format (32-bit data, no swizzle); the
    // range (2 GiB) only has to cover the chunk
    return __builtin_amdgcn_make_buffer_rsrc(base, (short)0, 0x7fffffff, 0x00020000);
}
template <bool NT, class V, int E>
__device__ __forceinline__ void st_out(__amdgpu_buffer_rsrc_t r, uint32_t elem, const V (&o)[E]) {
    static_assert(sizeof(V) * E == 16, "one 16-byte store per lane");
    u32x4 u;
    if constexpr (sizeof(V) == 4) {
        const f32x4 v = {o[0], o[1], o[2], o[3]};
        u = __builtin_bit_cast(u32x4, v);
    } else {
        const f64x2 v = {o[0], o[1]};
        u = __builtin_bit_cast(u32x4, v);
    }
    __builtin_amdgcn_raw_buffer_store_b128(u, r, elem * (uint32_t)sizeof(V), 0, NT ? 2 : 0);
}

template <class V>
struct TileArgs {
    const V *gx;       // gathered vector G
    const V *x;        // state, self term
    V *out;
    V c;               // out = self + c * (L G)
    // EPI_STAGE: k = c (L G); out = self + cy k; acc_out = k (stage 0) or
    // acc_in + 2 k.  EPI_RK4ACC: x_new = self + (acc_in + c (L G)) / 6.
    const V *acc_in;
    V *acc_out;
    V cy;
    int stage;
    int use_self;      // STORE: add X[s] (or G[s] when SELF)
    const double *rates;
    int rset;
    RunState *st;
    double *pchk, *pmass;
    int n;
    uint64_t ntiles;
    int xcd_order;     // tile order (tile_walk)
    uint32_t xcd_res;  // workgroups resident per XCD (order 2)
    const uint32_t *walk;  // tile at each walk position (level walk), or null: the position itself
};


// Stage-input values of one lane: R columns 2^LB states apart.
template <class V, int R, int E>
struct Cols {
    V g[R][E];
};

template <class V, int E, int R, int LB>
__device__ __forceinline__ void issue_cols(const TileArgs<V> &a, uint64_t base, uint32_t lo, Cols<V, R, E> &s) {
#pragma unroll
    for (int c = 0; c < R; c++) ld16(a.gx + base + lo + ((uint32_t)c << LB), s.g[c]);
}

// acc[c0 + c] += r * G[c] for the R loaded columns.
template <class V, int E, int R, int C>
__device__ __forceinline__ void fma_cols(V r, const Cols<V, R, E> &s, V (&acc)[C][E], int c0 = 0) {
#pragma unroll
    for (int c = 0; c < R; c++)
#pragma unroll
        for (int j = 0; j < E; j++) acc[c0 + c][j] = fma(r, s.g[c][j], acc[c0 + c][j]);
}

template <class V, int E, int C, int LB>
__device__ __forceinline__ void load_cols_x(const V *p, uint64_t base, uint32_t lo, V (&g)[C][E]) {
#pragma unroll
    for (int c = 0; c < C; c++) ld_self<V, E>(p + base + lo + ((uint32_t)c << LB), g[c]);
}

// Walk positions of one workgroup: its k-th position is base + k * step;
// the tile at position p is walk[p] with a level-walk table, else p.
//   order 0: grid order, T = g + k * G
//   order 1: XCD eighths -- workgroup g runs on XCD g % 8 and walks that
//            XCD's eighth of the tile range with stride G / 8
//   order 2: XCD eighths, and the R workgroups resident on an XCD at a time
//            (dispatched together) take R consecutive tiles per step, so the
//            XCD sweeps its eighth in order and a tile-bond partner 2^j tiles
//            ahead is read again while it still sits in that XCD's L2.
//            Needs G % 8 == 0, (G / 8) % R == 0 and ntiles % G == 0.
//   order 3: one sweep over the whole range: the R workgroups resident on
//            the device take R consecutive tiles per step, so a partner
//            2^j tiles ahead is read again 2^j tiles of traffic later (a
//            window the memory-side cache holds for larger j than the
//            per-XCD sweeps).  Needs G % R == 0 and ntiles % G == 0.
//   The level walk (a table over the positions of order 2) groups tiles
//   by the particles on their tile bits above the lowest: by count, then by
//   position sum, alternately ascending and descending within a sum.  A
//   hop between two such bits keeps the count and raises the sum by one,
//   so a tile-bond partner sits in the next group, not 2^j tiles ahead.
struct TileWalk {
    uint64_t base, step;
};
__device__ __forceinline__ TileWalk tile_walk(uint64_t ntiles, int order, uint32_t res) {
    const uint64_t g = blockIdx.x, G = gridDim.x;
    if (order == 0) return {g, G};
    if (order == 3) {
        const uint64_t K = ntiles / G;                             // tiles per workgroup
        return {g % res + (uint64_t)res * K * (g / res), res};
    }
    const uint64_t per_xcd = ntiles >> 3, q = g >> 3;
    if (order == 1) return {(g & 7) * per_xcd + q, G >> 3};
    const uint64_t K = per_xcd / (G >> 3);                         // tiles per workgroup
    return {(g & 7) * per_xcd + q % res + (uint64_t)res * K * (q / res), res};
}

// Pop the lowest set bit of a uniform bond mask: returns the bond index
// (LB0 + bit) or -1 when the mask is empty.
__device__ __forceinline__ int pop_bond(uint64_t &m, int LB0) {
    if (!m) return -1;
    const int k = LB0 + __builtin_ctzll(m);
    m &= m - 1;
    return k;
}

// One workgroup's share of a tiled pass: the tiles at its walk positions
// (`tiles`) of the lattice described by `a`, with the pass's rates already in LDS
// (rr) and the tile and reduction buffers supplied by the kernel.  The
// step's final passes leave the workgroup's (sum |dx|, sum x) in partial
// slot `part`.  `cur` holds the first tile's own vector, already issued by
// the caller (the kernel issues it before its prologue waits).
template <class V, int EPI, bool SELF, bool NT>
__device__ __forceinline__ void tile_pass_from(const TileArgs<V> &a, const V *__restrict__ rr, V *tile,
                                               double *red, WalkTiles tiles, uint64_t part,
                                               Cols<V, TileGeom<V>::C, TileGeom<V>::E> cur) {
    using G = TileGeom<V>;
    constexpr int E = G::E, e = G::e, LB = G::LB, C = G::C, TB0 = G::TB0;
    const int n = a.n;
    const int t = threadIdx.x;
    const int l = t & 63;                                          // lane in the chunk
    const int w = __builtin_amdgcn_readfirstlane(t >> 6);          // chunk (uniform per wave)
    const uint32_t lo = (uint32_t)l * E;                           // lane offset (elements)
    const V alpha = rr[0], beta = rr[n];
    const int tb = n - 3 - TB0;                                    // bits of the tile index (>= 1)
    const uint64_t tmask = tb > 1 ? (((uint64_t)1 << (tb - 1)) - 1) : 0;  // bonds TB0..n-5
    const uint64_t wofs = (uint64_t)w << (n - 3);
    // LDS slot of (wave, column, lane): ((w*C + c)*64 + l)*E
    const uint32_t lds0 = ((uint32_t)w * C * 64 + (uint32_t)l) * E;
    constexpr uint32_t LCOL = 64 * E;                              // next column
    // ---- tile-invariant parts of the diagonal: lane bonds, wave bonds, entry
    const uint32_t lout = ~(uint32_t)l & ((uint32_t)l >> 1) & 0x1Fu;
    V d_fix = 0;
#pragma unroll
    for (int q = 0; q < 5; q++)
        if ((lout >> q) & 1) d_fix += rr[n - 1 - (e + q)];
    const bool w0 = w & 1, w1 = (w >> 1) & 1, w2 = (w >> 2) & 1;
    const V r_n2 = rr[1], r_n3 = rr[2], r_n4 = rr[3];           // bonds n-2, n-3, n-4
    if (!w0 && w1) d_fix += r_n3;
    if (!w1 && w2) d_fix += r_n2;
    if (!w2) d_fix += alpha;
    V wl[5];                                                       // lane-bond in-flow weights
    uint32_t ls[5];                                                // and their LDS sources
#pragma unroll
    for (int q = 0; q < 5; q++) {
        const bool on = ((l >> q) & 1) && !((l >> (q + 1)) & 1);
        wl[q] = on ? rr[n - 1 - (e + q)] : (V)0;
        ls[q] = on ? lds0 + ((uint32_t)E << q) : lds0;
    }
    const V r_e1 = rr[n - e];                                      // bond e-1
    const bool lane_hi = (l >> 5) & 1;
    const V r_c0 = rr[n - LB];                                     // bond LB-1: lane bit 5, column bit 0
    const V r_c1 = rr[n - 1 - LB];                                 // bond LB: column bits 0, 1
    const V r_c2 = rr[n - 2 - LB];                                 // bond LB+1: column bit 1, tile bit 0
    const V w_c0 = lane_hi ? r_c0 : (V)0;                          // in-flow for columns 0, 2
    const V d_c0 = lane_hi ? (V)0 : r_c0;                          // out-flow for columns 1, 3
    const bool use_self = (EPI != EPI_STORE) || (!SELF && a.use_self);
    Kahan chk, mass;

    constexpr int NS = TASEP_TILE_SLOTS;
    using Str = Cols<V, C, E>;
    Str gn;                                                        // own vector of the next tile
    const uint64_t count = tiles.count;
    for (uint64_t c0 = 0; c0 < count; c0 += WalkTiles::kChunk) {
      if (c0) tiles.load(c0);
      const uint64_t cend = c0 + WalkTiles::kChunk < count ? c0 + WalkTiles::kChunk : count;
      for (uint64_t kt = c0; kt < cend; kt++) {
        const uint64_t T = tiles.at(kt);
        const uint64_t Tn = kt + 1 < count ? tiles.at(kt + 1) : 0;
        const uint64_t wb = (T << TB0) + wofs;                     // wave base (uniform)
        const bool t_lo = T & 1, t_hi = (T >> (tb - 1)) & 1;
        constexpr bool kOwnFirst = sizeof(V) == 4 ? TASEP_OWN_FIRST_F32 : TASEP_OWN_FIRST_F64;
        const V(&g)[C][E] = cur.g;
        if constexpr (kOwnFirst) {
            // ---- own vector to LDS first.  It was prefetched a tile ago;
            // the reads below are issued after it, so the wait for it (vmcnt
            // counts in issue order) covers only the own vector and the
            // previous tile's stores, not this tile's conditionally issued
            // streams, and the LDS phase overlaps their latency.
            __syncthreads();  // the previous tile's LDS reads are done
#pragma unroll
            for (int c = 0; c < C; c++) st16(&tile[lds0 + c * LCOL], g[c]);
        }
        // ---- issue every global read of this tile in the order it is used:
        // bond LB+1 (columns 2, 3 read columns 0, 1 of tile T+1), bond n-4
        // (per wave), the first NS tile bonds, the self vector, then the next
        // tile's own vector; all are in flight during the LDS phase
        Cols<V, 2, E> seam;
        if (!t_lo) issue_cols<V, E, 2, LB>(a, wb + ((uint64_t)1 << TB0), lo, seam);
        const bool has_n4 = t_hi && !w0;
        Str s4;
        if (has_n4) issue_cols<V, E, C, LB>(a, wb + ((uint64_t)1 << (n - 4)), lo, s4);
        uint64_t mm = T & ~(T >> 1) & tmask;
        int ks[NS > 0 ? NS : 1] = {-1};
        Str S[NS > 1 ? NS : 2];
#pragma unroll
        for (int q = 0; q < NS; q++) {
            ks[q] = pop_bond(mm, TB0);
            if (ks[q] >= 0) issue_cols<V, E, C, LB>(a, wb + ((uint64_t)1 << ks[q]), lo, S[q]);
        }
        V xs[C][E];
        if constexpr (!SELF) load_cols_x<V, E, C, LB>(a.x, wb, lo, xs);
        constexpr bool kAcc = EPI == EPI_STAGE || EPI == EPI_RK4ACC;
        V ac[kAcc ? C : 1][E];                                     // running k-sum
        if constexpr (kAcc) {
            if (EPI == EPI_RK4ACC || a.stage > 0) load_cols_x<V, E, C, LB>(a.acc_in, wb, lo, ac);
        }
        const bool more = kt + 1 < count;
        // issued on every tile (the last re-reads its own vector, an L2
        // hit) so that the compiler can count it: a wait for an earlier
        // stream then leaves it in flight (vmcnt is in issue order, and a
        // conditional load is counted as absent)
        issue_cols<V, E, C, LB>(a, ((more ? Tn : T) << TB0) + wofs, lo, gn);
        if constexpr (!kOwnFirst) {
            // ---- own vector to LDS (waited for only here)
            __syncthreads();  // the previous tile's LDS reads are done
#pragma unroll
            for (int c = 0; c < C; c++) st16(&tile[lds0 + c * LCOL], g[c]);
        }
        // ---- uniform and per-wave parts of the diagonal
        const uint64_t hout = ~T & (T >> 1) & tmask;
        V d_t = d_fix;
        for (uint64_t m = hout; m; m &= m - 1) d_t += rr[n - 1 - (TB0 + __builtin_ctzll(m))];
        if (!t_hi && w0) d_t += r_n4;
        __syncthreads();  // tile in LDS
        V acc[C][E];
#pragma unroll
        for (int c = 0; c < C; c++)
#pragma unroll
            for (int j = 0; j < E; j++) acc[c][j] = 0;
#pragma unroll
        for (int c = 0; c < C; c++) {
            const uint32_t base = lds0 + c * LCOL;
            V lv[E];
            // lane bonds e..LB-2 (weight 0 where inactive)
#pragma unroll
            for (int q = 0; q < 5; q++) {
                ld16(&tile[ls[q] + c * LCOL], lv);
#pragma unroll
                for (int j = 0; j < E; j++) acc[c][j] = fma(wl[q], lv[j], acc[c][j]);
            }
            // bond e-1: register bit e-1, lane bit 0 (next lane)
            ld16(&tile[(l & 1) ? base : base + E], lv);
            const V we = (l & 1) ? (V)0 : r_e1;
#pragma unroll
            for (int j = E / 2; j < E; j++) acc[c][j] = fma(we, lv[j - E / 2], acc[c][j]);
            // bond LB-1: lane bit 5, column bit 0 (columns 0, 2 read lane l-32 of column c+1)
            if ((c & 1) == 0) {
                ld16(&tile[lane_hi ? base + LCOL - 32 * E : base], lv);
#pragma unroll
                for (int j = 0; j < E; j++) acc[c][j] = fma(w_c0, lv[j], acc[c][j]);
            }
        }
        // ---- wave bonds n-3, n-2 and the entry: other chunks, same lane/column
        if (w0 && !w1) {
#pragma unroll
            for (int c = 0; c < C; c++) {
                V lv[E];
                ld16(&tile[lds0 + c * LCOL + C * LCOL], lv);
#pragma unroll
                for (int j = 0; j < E; j++) acc[c][j] = fma(r_n3, lv[j], acc[c][j]);
            }
        }
        if (w1 && !w2) {
#pragma unroll
            for (int c = 0; c < C; c++) {
                V lv[E];
                ld16(&tile[lds0 + c * LCOL + 2 * C * LCOL], lv);
#pragma unroll
                for (int j = 0; j < E; j++) acc[c][j] = fma(r_n2, lv[j], acc[c][j]);
            }
        }
        if (w2) {
#pragma unroll
            for (int c = 0; c < C; c++) {
                V lv[E];
                ld16(&tile[lds0 + c * LCOL - 4 * C * LCOL], lv);
#pragma unroll
                for (int j = 0; j < E; j++) acc[c][j] = fma(alpha, lv[j], acc[c][j]);
            }
        }
        // ---- static register bonds: 0..e-2, column bond LB, exit site
#pragma unroll
        for (int k = 0; k <= e - 2; k++) {
            const V r = rr[n - 1 - k];
#pragma unroll
            for (int c = 0; c < C; c++)
#pragma unroll
                for (int j = 0; j < E; j++)
                    if (((j >> k) & 1) && !((j >> (k + 1)) & 1)) acc[c][j] = fma(r, g[c][j + (1 << k)], acc[c][j]);
        }
#pragma unroll
        for (int j = 0; j < E; j++) acc[1][j] = fma(r_c1, g[2][j], acc[1][j]);
#pragma unroll
        for (int c = 0; c < C; c++)
#pragma unroll
            for (int j = 0; j < E; j += 2) acc[c][j] = fma(beta, g[c][j + 1], acc[c][j]);
        // ---- global streams, in issue order; tile bonds beyond the first NS
        // two at a time
        if (!t_lo) fma_cols<V, E, 2, C>(r_c2, seam, acc, 2);
        if (has_n4) fma_cols<V, E, C, C>(r_n4, s4, acc);
#pragma unroll
        for (int q = 0; q < NS; q++)
            if (ks[q] >= 0) fma_cols<V, E, C, C>(rr[n - 1 - ks[q]], S[q], acc);
        if (NS == 0 || ks[NS - 1] >= 0) {
            int kA = pop_bond(mm, TB0);
            if (kA >= 0) issue_cols<V, E, C, LB>(a, wb + ((uint64_t)1 << kA), lo, S[0]);
            while (kA >= 0) {
                const int kB = pop_bond(mm, TB0);
                if (kB >= 0) issue_cols<V, E, C, LB>(a, wb + ((uint64_t)1 << kB), lo, S[1]);
                fma_cols<V, E, C, C>(rr[n - 1 - kA], S[0], acc);
                if (kB < 0) break;
                kA = pop_bond(mm, TB0);
                if (kA >= 0) issue_cols<V, E, C, LB>(a, wb + ((uint64_t)1 << kA), lo, S[0]);
                fma_cols<V, E, C, C>(rr[n - 1 - kB], S[1], acc);
            }
        }
        // ---- diagonal and epilogue.  The step's final passes sum |dx| and
        // x over the lane's C*E states of the tile in plain FP64 (16 FP32 or
        // 8 FP64 terms: rounding below 2^-49 of their magnitudes' sum) and
        // add that to the compensated totals once per tile: the FP32 P4 at
        // n=28 0.757 -> 0.749 ms per launch against one compensated add per
        // column.
        double tcs = 0, tms = 0;
#pragma unroll
        for (int c = 0; c < C; c++) {
            V dc = d_t;
            if (c & 1) dc += d_c0;                                 // bond LB-1 out-flow
            if (c == 2) dc += r_c1;                                // bond LB out-flow
            if (c < 2 && t_lo) dc += r_c2;                         // bond LB+1 out-flow
            V d[E];
#pragma unroll
            for (int j = 0; j < E; j++) {
                d[j] = dc;
                if (j & 1) d[j] += beta;
                if (j < E / 2 && (l & 1)) d[j] += r_e1;
            }
#pragma unroll
            for (int k = 0; k <= e - 2; k++) {
                const V r = rr[n - 1 - k];
#pragma unroll
                for (int j = 0; j < E; j++)
                    if (!((j >> k) & 1) && ((j >> (k + 1)) & 1)) d[j] += r;
            }
            V o[E];
            if constexpr (EPI == EPI_STORE) {
#pragma unroll
                for (int j = 0; j < E; j++) {
                    const V self = SELF ? g[c][j] : (use_self ? xs[c][j] : (V)0);
                    o[j] = fma(a.c, fma(-d[j], g[c][j], acc[c][j]), self);
                }
            } else if constexpr (EPI == EPI_STAGE) {
                // k rounded as the stored stages round it (fma with +0), the
                // next stage input and the running sum k1 + 2 k2 + 2 k3
                V ao[E];
#pragma unroll
                for (int j = 0; j < E; j++) {
                    const V self = SELF ? g[c][j] : xs[c][j];
                    const V kk = fma(a.c, fma(-d[j], g[c][j], acc[c][j]), (V)0);
                    o[j] = fma(a.cy, kk, self);
                    ao[j] = a.stage > 0 ? (V)(ac[c][j] + (V)2 * kk) : kk;
                }
                st_out<NT, V, E>(chunk_rsrc(a.acc_out + wb), lo + ((uint32_t)c << LB), ao);
            } else {
                V dl[E];
#pragma unroll
                for (int j = 0; j < E; j++) dl[j] = a.c * fma(-d[j], g[c][j], acc[c][j]);
                if constexpr (EPI == EPI_RK4ACC) {
#pragma unroll
                    for (int j = 0; j < E; j++) dl[j] = (ac[c][j] + dl[j]) / (V)6;
                }
#pragma unroll
                for (int j = 0; j < E; j++) {
                    o[j] = xs[c][j] + dl[j];
                    tcs += fabs((double)dl[j]);
                    tms += (double)o[j];
                }
            }
            st_out<NT, V, E>(chunk_rsrc(a.out + wb), lo + ((uint32_t)c << LB), o);
        }
        if constexpr (EPI == EPI_HORNER || EPI == EPI_RK4ACC) {
            chk.add(tcs);
            mass.add(tms);
        }
        if (more) cur = gn;
      }
    }
    if constexpr (EPI == EPI_HORNER || EPI == EPI_RK4ACC) {
        double bc = chk.value(), bm = mass.value();
        block_sum2<kTileThreads>(bc, bm, red);
        if (threadIdx.x == 0) {
            a.pchk[part] = bc;
            a.pmass[part] = bm;
        }
    }
}

template <class V, int EPI, bool SELF, bool NT>
__global__ __launch_bounds__(kTileThreads, TASEP_TILE_WAVES) void k_tile_pass(const TileArgs<V> a) {
    using TG = TileGeom<V>;
    __shared__ V tile[kTileThreads * TG::C * TG::E];
    __shared__ V rr[kRateStride];
    __shared__ double red[2 * kTileWaves];
    // this workgroup's tiles: blockIdx.x, blockIdx.x + gridDim.x, ... < ntiles,
    // placed along the walk
    const TileWalk walk = tile_walk(a.ntiles, a.xcd_order, a.xcd_res);
    const uint64_t g = blockIdx.x, G = gridDim.x;
    const uint64_t count = g < a.ntiles ? (a.ntiles - g + G - 1) / G : 0;
    // the rates, then the first tile's own vector, are in flight before
    // anything waits (the stopped flag is a scalar load)
    const int t = threadIdx.x;
    const double rate = t <= a.n ? a.rates[a.rset * kRateStride + t] : 0.0;
    WalkTiles tiles(a.walk, walk.base, walk.step, count);
    Cols<V, TG::C, TG::E> first;
    if (count > 0) {
        const uint64_t wofs = (uint64_t)__builtin_amdgcn_readfirstlane(t >> 6) << (a.n - 3);
        issue_cols<V, TG::E, TG::C, TG::LB>(a, (tiles.at(0) << TG::TB0) + wofs, (uint32_t)(t & 63) * TG::E, first);
    }
    if (a.st->stopped) return;
    if (t <= a.n) rr[t] = (V)rate;
    __syncthreads();
    tile_pass_from<V, EPI, SELF, NT>(a, rr, tile, red, tiles, g, first);
}

// ===========================================================================
// The reference loop (TASEP_REF) on the same FP64 tiles: one pass per step,
// x' = x + (((k + 2k) + 2k) + k)/6 with k = h (L x)_s, bit-exact.  Each
// state's row is accumulated in the reference's term order
// (src/lib/solv.c:8-13 over the rows of src/lib/matgen.c:29-67): T = 0,
// T += alpha x(entry), T += d x(s), T += beta x(s+1), then bonds k = 0..n-2
// ascending, every product rounded before its add; the diagonal is
// 0 - beta - alpha - out-bonds ascending.  The bond sources of the tile
// layout (see the bit map above; FP64: e = 1, LB = 7, TB0 = 9) come in
// exactly that order: bond 0 (next lane), bonds 1-5 (lane bonds), 6 (lane
// bit 5 / column bit 0), 7 (column), 8 (column / tile seam), tile bonds
// ascending, n-4 (per wave), n-3 and n-2 (other chunks).  A per-lane bond
// that does not pull into a state adds 0 * x = +-0, which leaves T unchanged
// (T starts at +0 and a round-to-nearest sum is -0 only if both operands
// are); a wave- or tile-uniform one is skipped, as in the reference.
// ===========================================================================
struct TileRefArgs {
    const double *x;    // state: gathered and self
    double *out;        // next state
    const double *rates;
    RunState *st;
    double *pchk, *pmass;
    double h;
    int n;
    uint64_t ntiles;
    int xcd_order;
    uint32_t xcd_res;
    const uint32_t *walk;  // tile at each walk position (level walk), or null
};

template <bool NT>
__global__ __launch_bounds__(kTileThreads) void k_tile_ref(const TileRefArgs a) {
    using G = TileGeom<double>;
    constexpr int E = G::E, LB = G::LB, C = G::C, TB0 = G::TB0;
    static_assert(E == 2 && LB == 7 && TB0 == 9, "bond order below assumes the FP64 tile");
    __shared__ double tile[kTileThreads * C * E];
    __shared__ double rr[kRateStride];
    __shared__ double red[2 * kTileWaves];
    const int n = a.n;
    const int t = threadIdx.x;
    const int l = t & 63;
    const int w = __builtin_amdgcn_readfirstlane(t >> 6);
    const uint32_t lo = (uint32_t)l * E;
    const uint64_t wofs = (uint64_t)w << (n - 3);
    const TileWalk walk = tile_walk(a.ntiles, a.xcd_order, a.xcd_res);
    auto issue = [&](uint64_t base, double (&dst)[C][E]) {
#pragma unroll
        for (int c = 0; c < C; c++) ld16(a.x + base + lo + ((uint32_t)c << LB), dst[c]);
    };
    // Prologue: the rates, then the first tile's own vector, are in flight
    // before anything waits (the stopped flag is a scalar load), so that a
    // workgroup with a single tile (mid n) pays one memory latency, not three.
    const double rate = t <= n ? a.rates[t] : 0.0;
    const uint64_t g0 = blockIdx.x;
    uint64_t kt = 0;
    double g[C][E];
    const uint64_t count = g0 < a.ntiles ? (a.ntiles - g0 + gridDim.x - 1) / gridDim.x : 0;
    WalkTiles tiles(a.walk, walk.base, walk.step, count);
    if (count) issue((tiles.at(0) << TB0) + wofs, g);
    if (a.st->stopped) return;
    if (t <= n) rr[t] = rate;
    __syncthreads();
    const double alpha = rr[0], beta = rr[n];
    auto rb = [&](int k) { return rr[n - 1 - k]; };               // rate of bond k
    const int tb = n - 3 - TB0;
    const uint64_t tmask = tb > 1 ? (((uint64_t)1 << (tb - 1)) - 1) : 0;
    const uint32_t lds0 = ((uint32_t)w * C * 64 + (uint32_t)l) * E;
    constexpr uint32_t LCOL = 64 * E;
    const bool w0 = w & 1, w1 = (w >> 1) & 1, w2 = (w >> 2) & 1;
    const bool lane_hi = (l >> 5) & 1;
    // per-lane in-flow weights and LDS sources of bonds 1..5 (lane bits q, q+1)
    double wl[5], dl5[5];
    uint32_t ls[5];
#pragma unroll
    for (int q = 0; q < 5; q++) {
        const bool on = ((l >> q) & 1) && !((l >> (q + 1)) & 1);
        const bool off = !((l >> q) & 1) && ((l >> (q + 1)) & 1);
        wl[q] = on ? rb(1 + q) : 0.0;
        dl5[q] = off ? rb(1 + q) : 0.0;
        ls[q] = on ? lds0 + ((uint32_t)E << q) : lds0;
    }
    const double w_b0 = (l & 1) ? 0.0 : rb(0);                    // bond 0 in-flow (odd j)
    const double d_b0 = (l & 1) ? rb(0) : 0.0;                    // bond 0 out-flow (even j)
    // leading part of the diagonal, the same for every tile: 0 - beta - alpha
    // - bonds 0..5 (subtracting an inactive bond's 0 is exact)
    double d_lane[E];
#pragma unroll
    for (int j = 0; j < E; j++) {
        double d = 0.0;
        if (j & 1) d -= beta;
        if (!w2) d -= alpha;
        if (j == 0) d -= d_b0;
#pragma unroll
        for (int q = 0; q < 5; q++) d -= dl5[q];
        d_lane[j] = d;
    }
    const double w_b6 = lane_hi ? rb(6) : 0.0;                    // bond 6 in-flow, columns 0, 2
    const double d_b6 = lane_hi ? 0.0 : rb(6);                    // bond 6 out-flow, columns 1, 3
    Kahan chk, mass;
    const double h = a.h;

    // One tile bond pre-issued (bond n-4 first): FP64 n=20..26 3-5% faster
    // per step than two, the same at n=28; three are 3% slower.
    constexpr int NS = TASEP_REF_SLOTS;
    for (uint64_t c0 = 0; c0 < count; c0 += WalkTiles::kChunk) {
      if (c0) tiles.load(c0);
      const uint64_t cend = c0 + WalkTiles::kChunk < count ? c0 + WalkTiles::kChunk : count;
      for (kt = c0; kt < cend; kt++) {
        const uint64_t T = tiles.at(kt);
        const uint64_t Tn = kt + 1 < count ? tiles.at(kt + 1) : 0;
        const uint64_t wb = (T << TB0) + wofs;
        const bool t_lo = T & 1, t_hi = (T >> (tb - 1)) & 1;
#if TASEP_OWN_FIRST_F64
        // ---- own vector to LDS first (prefetched a tile ago; waiting for it
        // then does not wait for this tile's streams, issued below)
        __syncthreads();  // the previous tile's LDS reads are done
#pragma unroll
        for (int c = 0; c < C; c++) st16(&tile[lds0 + c * LCOL], g[c]);
#endif
        // ---- issue every global read of this tile before its LDS phase:
        // bond 8 (columns 2, 3 read columns 0, 1 of tile T+1), bond n-4, the
        // first NS tile bonds, then the next tile's own vector
        double seam[2][E];
        if (!t_lo) {
            ld16(a.x + wb + lo + ((uint64_t)1 << TB0), seam[0]);
            ld16(a.x + wb + lo + ((uint64_t)1 << TB0) + (1u << LB), seam[1]);
        }
        const bool has_n4 = t_hi && !w0;
        double s4[C][E];
        if (has_n4) issue(wb + ((uint64_t)1 << (n - 4)), s4);
        uint64_t mm = T & ~(T >> 1) & tmask;
        int ks[NS];
        double S[NS > 1 ? NS : 2][C][E];
#pragma unroll
        for (int q = 0; q < NS; q++) {
            ks[q] = pop_bond(mm, TB0);
            if (ks[q] >= 0) issue(wb + ((uint64_t)1 << ks[q]), S[q]);
        }
        double gn[C][E];
        const bool more = kt + 1 < count;
        // issued on every tile (the last re-reads its own vector) so that
        // the compiler can count it: waits for earlier streams leave it in
        // flight
        issue(((more ? Tn : T) << TB0) + wofs, gn);
#if !TASEP_OWN_FIRST_F64
        __syncthreads();  // the previous tile's LDS reads are done
#pragma unroll
        for (int c = 0; c < C; c++) st16(&tile[lds0 + c * LCOL], g[c]);
#endif
        __syncthreads();
        const uint64_t hout = ~T & (T >> 1) & tmask;
        double Tv[C][E];
        // ---- diagonal (reference subtraction order: the lane part, bonds
        // 6..8, tile bonds ascending, n-4, n-3, n-2), then entry and d x(s)
        {
            double dg[C][E];
#pragma unroll
            for (int c = 0; c < C; c++)
#pragma unroll
                for (int j = 0; j < E; j++) {
                    double d = d_lane[j];
                    if (c & 1) d -= d_b6;
                    if (c == 2) d -= rb(LB);
                    if (c < 2 && t_lo) d -= rb(LB + 1);
                    dg[c][j] = d;
                }
            for (uint64_t m = hout; m; m &= m - 1) {
                const double r = rb(TB0 + __builtin_ctzll(m));
#pragma unroll
                for (int c = 0; c < C; c++)
#pragma unroll
                    for (int j = 0; j < E; j++) dg[c][j] -= r;
            }
            const double r4 = rb(n - 4), r3 = rb(n - 3), r2 = rb(n - 2);
            const bool o4 = !t_hi && w0, o3 = !w0 && w1, o2 = !w1 && w2;
            double lv[E];
            const double wa = w2 ? alpha : 0.0;
            const uint32_t src = w2 ? lds0 - 4 * C * LCOL : lds0;
#pragma unroll
            for (int c = 0; c < C; c++) {
                ld16(&tile[src + c * LCOL], lv);
#pragma unroll
                for (int j = 0; j < E; j++) {
                    double d = dg[c][j];
                    if (o4) d -= r4;
                    if (o3) d -= r3;
                    if (o2) d -= r2;
                    double T0 = 0.0;
                    T0 += wa * lv[j];
                    T0 += d * g[c][j];
                    Tv[c][j] = T0;
                }
            }
        }
        // ---- exit (even j reads j+1), bond 0 (odd j reads lane l+1, j = 0)
#pragma unroll
        for (int c = 0; c < C; c++) {
            Tv[c][0] += beta * g[c][1];
            double lv[E];
            ld16(&tile[(l & 1) ? lds0 + c * LCOL : lds0 + c * LCOL + E], lv);
            Tv[c][1] += w_b0 * lv[0];
        }
        // ---- bonds 1..5: lane l + 2^(k-1)
#pragma unroll
        for (int q = 0; q < 5; q++) {
#pragma unroll
            for (int c = 0; c < C; c++) {
                double lv[E];
                ld16(&tile[ls[q] + c * LCOL], lv);
#pragma unroll
                for (int j = 0; j < E; j++) Tv[c][j] += wl[q] * lv[j];
            }
        }
        // ---- bond 6: columns 0, 2 read lane l-32 of column c+1
#pragma unroll
        for (int c = 0; c < C; c += 2) {
            double lv[E];
            ld16(&tile[lane_hi ? lds0 + c * LCOL + LCOL - 32 * E : lds0 + c * LCOL], lv);
#pragma unroll
            for (int j = 0; j < E; j++) Tv[c][j] += w_b6 * lv[j];
        }
        // ---- bond 7: column 1 reads column 2
#pragma unroll
        for (int j = 0; j < E; j++) Tv[1][j] += rb(LB) * g[2][j];
        // ---- bond 8: columns 2, 3 read columns 0, 1 of tile T+1
        if (!t_lo) {
            const double r8 = rb(LB + 1);
#pragma unroll
            for (int j = 0; j < E; j++) {
                Tv[2][j] += r8 * seam[0][j];
                Tv[3][j] += r8 * seam[1][j];
            }
        }
        // ---- tile bonds ascending: the first NS as issued, the rest two
        // streams in flight at a time
        auto add = [&](int k, const double (&nb)[C][E]) {
            const double r = rb(k);
#pragma unroll
            for (int c = 0; c < C; c++)
#pragma unroll
                for (int j = 0; j < E; j++) Tv[c][j] += r * nb[c][j];
        };
#pragma unroll
        for (int q = 0; q < NS; q++)
            if (ks[q] >= 0) add(ks[q], S[q]);
        if (ks[NS - 1] >= 0) {
            int kA = pop_bond(mm, TB0);
            if (kA >= 0) issue(wb + ((uint64_t)1 << kA), S[0]);
            while (kA >= 0) {
                const int kB = pop_bond(mm, TB0);
                if (kB >= 0) issue(wb + ((uint64_t)1 << kB), S[1]);
                add(kA, S[0]);
                if (kB < 0) break;
                kA = pop_bond(mm, TB0);
                if (kA >= 0) issue(wb + ((uint64_t)1 << kA), S[0]);
                add(kB, S[1]);
            }
        }
        // ---- bond n-4: chunk w+1 of tile T - 2^(tb-1), 2^(n-4) states away
        if (has_n4) add(n - 4, s4);
        // ---- bonds n-3, n-2: chunks w+1, w+2 of this tile
        if (w0 && !w1) {
            const double r = rb(n - 3);
#pragma unroll
            for (int c = 0; c < C; c++) {
                double lv[E];
                ld16(&tile[lds0 + c * LCOL + C * LCOL], lv);
#pragma unroll
                for (int j = 0; j < E; j++) Tv[c][j] += r * lv[j];
            }
        }
        if (w1 && !w2) {
            const double r = rb(n - 2);
#pragma unroll
            for (int c = 0; c < C; c++) {
                double lv[E];
                ld16(&tile[lds0 + c * LCOL + 2 * C * LCOL], lv);
#pragma unroll
                for (int j = 0; j < E; j++) Tv[c][j] += r * lv[j];
            }
        }
        // ---- the reference update, checksum and mass (the lane's 8 states
        // of the tile summed plain, then one compensated add per tile: per
        // step at n=28 1.390 -> 1.379 ms, n=24 75 -> 71 us against one
        // compensated add per column)
        double cs = 0.0, ms = 0.0;
#pragma unroll
        for (int c = 0; c < C; c++) {
            double o[E];
#pragma unroll
            for (int j = 0; j < E; j++) {
                const double d = ref_delta(h, Tv[c][j]);
                o[j] = g[c][j] + d;
                cs += fabs(d);
                ms += o[j];
            }
            st_out<NT, double, E>(chunk_rsrc(a.out + wb), lo + ((uint32_t)c << LB), o);
        }
        chk.add(cs);
        mass.add(ms);
        if (more) {
#pragma unroll
            for (int c = 0; c < C; c++)
#pragma unroll
                for (int j = 0; j < E; j++) g[c][j] = gn[c][j];
        }
      }
    }
    double bc = chk.value(), bm = mass.value();
    block_sum2<kTileThreads>(bc, bm, red);
    store_partials(a.pchk, a.pmass, bc, bm);
}

}  // namespace tasep
