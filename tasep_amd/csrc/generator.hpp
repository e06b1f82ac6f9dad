// generator.hpp -- rows of the TASEP generator L, on the device.
//
// State s in [0, 2^n); bit k <-> site n-k from the entry; r[0] = alpha,
// r[n] = beta, bond k (particle at bit k+1 hops to empty bit k) has rate
// r[n-1-k].  Row s of L (pull form, what s receives):
//   (Lx)_s = [b_{n-1}] alpha x_{s-2^(n-1)} + d_s x_s + [!b_0] beta x_{s+1}
//          + sum_{k: b_k && !b_{k+1}} r[n-1-k] x_{s+2^k}
//   d_s    = -( [b_0] beta + [!b_{n-1}] alpha + sum_{k: !b_k && b_{k+1}} r[n-1-k] )
// (reference src/lib/matgen.c:29-67 and src/lib/solv.c:99-131 emit the
// same terms in the same order: entry, diagonal, exit, bulk k ascending).
//
// Every lane derives its row from the bitmask: the bonds that pull into s
// are the set bits of  s & ~(s >> 1) & (2^(n-1)-1)  and the bonds that drain
// s are the set bits of  ~s & (s >> 1) & (2^(n-1)-1).
#pragma once

#include "device_common.hpp"

namespace tasep {

struct Lattice {
    int n;
    uint64_t top;   // 2^(n-1): entry-site bit
    uint64_t bulk;  // 2^(n-1) - 1: bonds 0..n-2
};

__host__ __device__ inline Lattice make_lattice(int n) {
    Lattice L;
    L.n = n;
    L.top = (uint64_t)1 << (n - 1);
    L.bulk = L.top - 1;
    return L;
}

// ---------------------------------------------------------------------------
// Reference-faithful FP64 row (TASEP_REF): the exact operation sequence of
// the reference's per-row dot product, T = 0; T += v_j * x[c_j] in row order
// (src/lib/solv.c:8-13), with the diagonal built as 0 - beta - alpha - r...
// (src/lib/matgen.c:36-51).  The library is compiled with -ffp-contract=off,
// so no product is fused into the following add.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double ref_row_mf(const Lattice g, uint64_t s,
                                             const double *__restrict__ x,
                                             const double *r) {
    const int n = g.n;
    double T = 0.0;
    if (s & g.top) T += r[0] * x[s - g.top];
    double d = 0.0;
    if (s & 1) d -= r[n];
    if (!(s & g.top)) d -= r[0];
    uint64_t mo = ~s & (s >> 1) & g.bulk;
    while (mo) {
        int k = __builtin_ctzll(mo);
        mo &= mo - 1;
        d -= r[n - 1 - k];
    }
    T += d * x[s];
    if (!(s & 1)) T += r[n] * x[s + 1];
    uint64_t mi = s & ~(s >> 1) & g.bulk;
    while (mi) {
        int k = __builtin_ctzll(mi);
        mi &= mi - 1;
        T += r[n - 1 - k] * x[s + ((uint64_t)1 << k)];
    }
    return T;
}

// a / 6 correctly rounded, without the general division sequence.
// q = RN(a * y) with y = RN(1/6) is within one ulp of a/6, the remainder
// r = a - 6q is exact in one FMA, and RN(q + r * y) is then the correctly
// rounded quotient (Markstein's correction theorem: y within half an ulp of
// 1/b, q faithful, no underflow).  Nonzero |a| < 2^-960, where the products
// could underflow, and |a| > 2^1000 (infinities included) take the division
// itself; a signed zero keeps its sign through copysign (the sequence alone
// would turn -0 into +0).  Checked against a / 6.0 on 5e8 random and
// near-midpoint operands.
__device__ __forceinline__ double div6(double a) {
    constexpr double y = 1.0 / 6.0;
    const double m = fabs(a);
    if ((m < 0x1p-960 || m > 0x1p1000) && a != 0.0) return a / 6.0;
    const double q = a * y;
    const double r = __builtin_fma(-6.0, q, a);
    return copysign(__builtin_fma(r, y, q), a);
}

// The reference update for one state: all four stages equal k = h*T
// (src/lib/solv.c:12-13 reads the unmodified x), combined with the
// reference's rounding (src/masterEquation.c:640).
__device__ __forceinline__ double ref_delta(double h, double T) {
    const double k = h * T;
    return div6(((k + 2.0 * k) + 2.0 * k) + k);
}

// ---------------------------------------------------------------------------
// Performance rows (TASEP_RK4): same terms with explicit FMA, computed in the
// arithmetic type A (the state's storage type), stage input supplied by a
// gather functor G(j) -> A.
// ---------------------------------------------------------------------------
template <class A, class G>
__device__ __forceinline__ A rk_row_mf(const Lattice g, uint64_t s, const G &u, const double *r) {
    const int n = g.n;
    const bool occ_top = (s & g.top) != 0, occ_exit = (s & 1) != 0;
    A d = occ_exit ? (A)r[n] : (A)0;
    if (!occ_top) d += (A)r[0];
    uint64_t mo = ~s & (s >> 1) & g.bulk;
    while (mo) {
        int k = __builtin_ctzll(mo);
        mo &= mo - 1;
        d += (A)r[n - 1 - k];
    }
    A acc = -d * u(s);
    if (occ_top) acc = fma((A)r[0], u(s - g.top), acc);
    if (!occ_exit) acc = fma((A)r[n], u(s + 1), acc);
    uint64_t mi = s & ~(s >> 1) & g.bulk;
    while (mi) {
        int k = __builtin_ctzll(mi);
        mi &= mi - 1;
        acc = fma((A)r[n - 1 - k], u(s + ((uint64_t)1 << k)), acc);
    }
    return acc;
}

template <class G>
__device__ __forceinline__ double rk_row_csr(uint64_t s, const uint64_t *__restrict__ rowPtr,
                                             const uint64_t *__restrict__ colInd,
                                             const double *__restrict__ val, const G &u) {
    double acc = 0.0;
    const uint64_t e = rowPtr[s + 1];
    for (uint64_t j = rowPtr[s]; j < e; j++) acc = __builtin_fma(val[j], u(colInd[j]), acc);
    return acc;
}

// ---------------------------------------------------------------------------
// CSR construction.  rowPtr has a closed form, so every row is independent
// (no scan): with len(j) = 1 + b_{n-1}(j) + !b_0(j) + #{k: b_k(j) && !b_{k+1}(j)}
//   rowPtr[i] = i + max(0, i - 2^(n-1)) + ceil(i/2)
//             + sum_k ( floor(i / 2^(k+2)) 2^k + clamp(i mod 2^(k+2) - 2^k, 0, 2^k) )
// (the pattern b_{k+1} b_k = 01 fills 2^k of every 2^(k+2) consecutive j).
// ---------------------------------------------------------------------------
__host__ __device__ inline uint64_t csr_row_start(int n, uint64_t i) {
    const uint64_t top = (uint64_t)1 << (n - 1);
    uint64_t v = i + (i > top ? i - top : 0) + ((i + 1) >> 1);
    for (int k = 0; k < n - 1; k++) {
        const uint64_t lo = (uint64_t)1 << k;
        const uint64_t rem = i & (((uint64_t)1 << (k + 2)) - 1);
        uint64_t part = rem > lo ? rem - lo : 0;
        if (part > lo) part = lo;
        v += (i >> (k + 2)) * lo + part;
    }
    return v;
}

}  // namespace tasep
