// copy_calibrate -- plain streaming kernels of known size on 1 GiB arrays.
//
// Used to calibrate hardware traffic counters (rocprofv3 --pmc FETCH_SIZE /
// WRITE_SIZE read back per kernel dispatch) against bytes that are known
// exactly, and to measure the practical HBM bandwidth of the three stream
// shapes of the generator passes: read-only, copy (1 read + 1 write) and
// triad (2 reads + 1 write).
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
    CK(hipFree(a));
    CK(hipFree(b));
    CK(hipFree(c));
    CK(hipFree(o));
    return 0;
}
