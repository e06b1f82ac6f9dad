// host_common.hpp -- error reporting and device checks shared by the host
// translation units of libtasep_hip.so.
#pragma once

#include <hip/hip_runtime.h>

#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "../../include/tasep_hip.h"
#include "device_common.hpp"

namespace tasep_host {

constexpr int kMaxBlocks = 16384;  // grid cap of the streaming kernels
constexpr int kWindow = 128;      // steps per reduction window with the exit test armed
constexpr int kMomMaxBlocks = 2048;

// Message of the last error on this thread (tasep_strerror).
inline thread_local char g_errmsg[256] = "";

inline int fail(int code, const char *what, hipError_t e = hipSuccess) {
    if (e != hipSuccess)
        snprintf(g_errmsg, sizeof g_errmsg, "%s: %s", what, hipGetErrorString(e));
    else
        snprintf(g_errmsg, sizeof g_errmsg, "%s", what);
    return code;
}

#define HIP_OK(expr)                                                        \
    do {                                                                    \
        hipError_t e_ = (expr);                                             \
        if (e_ != hipSuccess)                                               \
            return fail(e_ == hipErrorOutOfMemory ? TASEP_ERR_NOMEM         \
                                                  : TASEP_ERR_HIP,          \
                        #expr, e_);                                         \
    } while (0)

inline int check_device() {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count < 1)
        return fail(TASEP_ERR_NODEV, "no HIP device visible");
    int dev = 0;
    HIP_OK(hipGetDevice(&dev));
    hipDeviceProp_t p;
    HIP_OK(hipGetDeviceProperties(&p, dev));
    if (strncmp(p.gcnArchName, "gfx950", 6) != 0)
        return fail(TASEP_ERR_NODEV, "device is not gfx950 (MI355X)");
    return TASEP_OK;
}

inline uint64_t grid_for(uint64_t dim, int cap = kMaxBlocks) {
    uint64_t b = (dim + tasep::kBlock - 1) / tasep::kBlock;
    return b < (uint64_t)cap ? (b ? b : 1) : (uint64_t)cap;
}

// Moments of a device state vector (defined with the moment kernels):
// y[n], cov[n*n], mul[n*n], mass as tasep_moments; FP32 vectors are divided
// by their FP64 mass.  *part and *out are work buffers allocated on first use.
int device_moments(hipStream_t stream, const void *x, int n, int prec, double **part, double **out, double *y,
                   double *cov, double *mul, double *mass);

// Window marginals W[o*16 + p] (o = 0..n-4) of a device state vector, as
// tasep_window_marginals.
int device_windows(hipStream_t stream, const void *x, int n, int prec, double **part, double **out, int width,
                   double *W);

}  // namespace tasep_host
