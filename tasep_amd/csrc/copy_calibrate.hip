// copy_calibrate -- plain streaming kernels of known size on 1 GiB arrays.
//
// Used to calibrate hardware traffic counters (rocprofv3 --pmc FETCH_SIZE /
// WRITE_SIZE read back per kernel dispatch) against bytes that are known
// exactly (copy4, triad4, read4: one 16-byte access per lane and
// iteration), and to measure the practical HBM bandwidth of the three stream
// shapes of the generator passes -- read-only, copy (1 read + 1 write) and
// triad (2 reads + 1 write) -- in their fastest form found on MI355X
// (copy_nt, triad_nt, read_nt: 4 or 8 independent 16-byte loads per lane in
// flight, non-temporal loads and stores).
#include <hip/hip_runtime.h>

#include <stdio.h>
#include <stdlib.h>

#define CK(x)                                                                      \
    do {                                                                           \
        hipError_t e_ = (x);                                                       \
        if (e_ != hipSuccess) {                                                    \
            fprintf(stderr, "%s: %s (line %d)\n", #x, hipGetErrorString(e_), __LINE__); \
            exit(1);                                                               \
        }                                                                          \
    } while (0)

__global__ void copy4(const float4 *__restrict__ a, float4 *__restrict__ b, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        b[i] = a[i];
}

__global__ void triad4(const float4 *__restrict__ a, const float4 *__restrict__ c, float4 *__restrict__ b,
                       size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float4 x = a[i], y = c[i];
        b[i] = make_float4(x.x + 0.5f * y.x, x.y + 0.5f * y.y, x.z + 0.5f * y.z, x.w + 0.5f * y.w);
    }
}

__global__ void read4(const float4 *__restrict__ a, float *out, size_t n) {
    float s = 0;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float4 x = a[i];
        s += x.x + x.y + x.z + x.w;
    }
    if (s == 1234.5f) out[0] = s;  // keeps the loads; never true for zero data
}

typedef float f4 __attribute__((ext_vector_type(4)));

// Workgroups walk contiguous chunks of U * blockDim 16-byte elements; n must
// be a multiple of the chunk.
template <int U>
__global__ void copy_nt(const f4 *__restrict__ a, f4 *__restrict__ b, size_t n) {
    const size_t chunk = (size_t)U * blockDim.x;
    for (size_t base = blockIdx.x * chunk; base < n; base += (size_t)gridDim.x * chunk) {
        f4 v[U];
#pragma unroll
        for (int u = 0; u < U; u++) v[u] = __builtin_nontemporal_load(a + base + u * blockDim.x + threadIdx.x);
#pragma unroll
        for (int u = 0; u < U; u++) __builtin_nontemporal_store(v[u], b + base + u * blockDim.x + threadIdx.x);
    }
}

template <int U>
__global__ void triad_nt(const f4 *__restrict__ a, const f4 *__restrict__ c, f4 *__restrict__ b, size_t n) {
    const size_t chunk = (size_t)U * blockDim.x;
    for (size_t base = blockIdx.x * chunk; base < n; base += (size_t)gridDim.x * chunk) {
        f4 v[U], w[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            v[u] = __builtin_nontemporal_load(a + base + u * blockDim.x + threadIdx.x);
            w[u] = __builtin_nontemporal_load(c + base + u * blockDim.x + threadIdx.x);
        }
#pragma unroll
        for (int u = 0; u < U; u++) __builtin_nontemporal_store(v[u] + 0.5f * w[u], b + base + u * blockDim.x + threadIdx.x);
    }
}

template <int U>
__global__ void read_nt(const f4 *__restrict__ a, float *out, size_t n) {
    const size_t chunk = (size_t)U * blockDim.x;
    f4 s = {0, 0, 0, 0};
    for (size_t base = blockIdx.x * chunk; base < n; base += (size_t)gridDim.x * chunk) {
#pragma unroll
        for (int u = 0; u < U; u++) s += __builtin_nontemporal_load(a + base + u * blockDim.x + threadIdx.x);
    }
    if (s.x + s.y + s.z + s.w == 1234.5f) out[0] = 1;  // keeps the loads; never true for zero data
}

int main() {
    const size_t bytes = (size_t)1 << 30, n4 = bytes / 16;
    float4 *a, *b, *c;
    float *o;
    CK(hipMalloc(&a, bytes));
    CK(hipMalloc(&b, bytes));
    CK(hipMalloc(&c, bytes));
    CK(hipMalloc(&o, 64));
    CK(hipMemset(a, 0, bytes));
    CK(hipMemset(b, 0, bytes));
    CK(hipMemset(c, 0, bytes));
    hipEvent_t e0, e1;
    CK(hipEventCreate(&e0));
    CK(hipEventCreate(&e1));
    const int grid = 32768, reps = 10;
    float ms;
    copy4<<<grid, 256>>>(a, b, n4);
    CK(hipEventRecord(e0));
    for (int r = 0; r < reps; r++) copy4<<<grid, 256>>>(a, b, n4);
    CK(hipEventRecord(e1));
    CK(hipEventSynchronize(e1));
    CK(hipEventElapsedTime(&ms, e0, e1));
    printf("copy  1 GiB read + 1 GiB write: %.3f ms  %.0f GB/s\n", ms / reps, 2.0 * bytes / (ms / reps * 1e-3) / 1e9);
    CK(hipEventRecord(e0));
    for (int r = 0; r < reps; r++) triad4<<<grid, 256>>>(a, c, b, n4);
    CK(hipEventRecord(e1));
    CK(hipEventSynchronize(e1));
    CK(hipEventElapsedTime(&ms, e0, e1));
    printf("triad 2 GiB read + 1 GiB write: %.3f ms  %.0f GB/s\n", ms / reps, 3.0 * bytes / (ms / reps * 1e-3) / 1e9);
    CK(hipEventRecord(e0));
    for (int r = 0; r < reps; r++) read4<<<grid, 256>>>(a, o, n4);
    CK(hipEventRecord(e1));
    CK(hipEventSynchronize(e1));
    CK(hipEventElapsedTime(&ms, e0, e1));
    printf("read  1 GiB read: %.3f ms  %.0f GB/s\n", ms / reps, 1.0 * bytes / (ms / reps * 1e-3) / 1e9);
    // fastest forms: 8192 workgroups of 256 lanes, chunks of 1024 / 2048 elements divide n4 = 2^26
    const int g2 = 8192;
    const f4 *fa = (const f4 *)a, *fc = (const f4 *)c;
    f4 *fb = (f4 *)b;
    copy_nt<4><<<g2, 256>>>(fa, fb, n4);
    CK(hipEventRecord(e0));
    for (int r = 0; r < reps; r++) copy_nt<4><<<g2, 256>>>(fa, fb, n4);
    CK(hipEventRecord(e1));
    CK(hipEventSynchronize(e1));
    CK(hipEventElapsedTime(&ms, e0, e1));
    printf("copy_nt  (4 in flight, non-temporal): %.3f ms  %.0f GB/s\n", ms / reps, 2.0 * bytes / (ms / reps * 1e-3) / 1e9);
    CK(hipEventRecord(e0));
    for (int r = 0; r < reps; r++) triad_nt<4><<<g2, 256>>>(fa, fc, fb, n4);
    CK(hipEventRecord(e1));
    CK(hipEventSynchronize(e1));
    CK(hipEventElapsedTime(&ms, e0, e1));
    printf("triad_nt (4 in flight, non-temporal): %.3f ms  %.0f GB/s\n", ms / reps, 3.0 * bytes / (ms / reps * 1e-3) / 1e9);
    CK(hipEventRecord(e0));
    for (int r = 0; r < reps; r++) read_nt<8><<<g2, 256>>>(fa, o, n4);
    CK(hipEventRecord(e1));
    CK(hipEventSynchronize(e1));
    CK(hipEventElapsedTime(&ms, e0, e1));
    printf("read_nt  (8 in flight, non-temporal): %.3f ms  %.0f GB/s\n", ms / reps, 1.0 * bytes / (ms / reps * 1e-3) / 1e9);
    CK(hipFree(a));
    CK(hipFree(b));
    CK(hipFree(c));
    CK(hipFree(o));
    return 0;
}
