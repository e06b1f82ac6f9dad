// lds_launch.hpp -- host side of the LDS-resident whole-run kernel
// (k_lds_run): LDS budget, workgroup size and the launch by precision,
// integrator and rate kind.
#pragma once

#include <atomic>
#include <mutex>

#include "host_common.hpp"
#include "lds_kernels.hpp"

namespace tasep_host {

using namespace tasep;

// Largest dynamic LDS a workgroup may take on the current device (the
// kernel's static LDS, about 4 KiB, is kept aside).
inline size_t lds_limit() {
    static std::atomic<size_t> cache{0};
    size_t lim = cache.load(std::memory_order_relaxed);
    if (!lim) {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && v > 0)
            lim = (size_t)v;
        else
            lim = 64 * 1024;
        cache.store(lim, std::memory_order_relaxed);
    }
    return lim - 8192;
}

// Workgroup size for a 2^n instance: one state per lane up to 1024 lanes
// (the passes are latency-bound, so more waves per CU win), at least a wave.
inline int lds_auto_threads(int n) {
    const uint64_t q = (uint64_t)1 << n;
    return q < 64 ? 64 : (q > 1024 ? 1024 : (int)q);
}

inline size_t lds_bytes(int n, int esize, int integ) {
    return (size_t)lds_buffers(integ == TASEP_REF ? 0 : 1) * esize << n;
}

inline bool lds_fits(int n, int esize, int integ) { return n <= kLdsMaxN && lds_bytes(n, esize, integ) <= lds_limit(); }

template <class V, int INTEG, bool TV>
inline void launch_lds_k(const LdsArgs &a, int blocks, int nt, size_t shmem, hipStream_t s) {
    static std::once_flag attr;
    std::call_once(attr, [] {
        hipFuncSetAttribute((const void *)k_lds_run<V, INTEG, TV>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)lds_limit());
    });
    hipLaunchKernelGGL((k_lds_run<V, INTEG, TV>), dim3(blocks), dim3(nt), shmem, s, a);
}

// One launch of k_lds_run over `blocks` instances whose largest needs
// `shmem` bytes of LDS.  tv: some instance has time-variant rates, or RK4
// must keep stored stages; csr: the reference loop reads the stored matrix.
inline void launch_lds_any(int prec, int integ, bool tv, const LdsArgs &a, int blocks, int nt, size_t shmem,
                           hipStream_t s, bool csr = false) {
    if (integ == TASEP_REF) {
        if (csr) launch_lds_k<double, 2, false>(a, blocks, nt, shmem, s);
        else if (tv) launch_lds_k<double, 0, true>(a, blocks, nt, shmem, s);
        else launch_lds_k<double, 0, false>(a, blocks, nt, shmem, s);
    } else if (prec == TASEP_F64) {
        if (tv) launch_lds_k<double, 1, true>(a, blocks, nt, shmem, s);
        else launch_lds_k<double, 1, false>(a, blocks, nt, shmem, s);
    } else {
        if (tv) launch_lds_k<float, 1, true>(a, blocks, nt, shmem, s);
        else launch_lds_k<float, 1, false>(a, blocks, nt, shmem, s);
    }
}

}  // namespace tasep_host
