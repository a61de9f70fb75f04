// The six 1-D slope limiters of source/limiters.d:7-37 as device functions, and a
// C entry point that applies one of them elementwise to device-resident pairs
// (Sp, Sm).  ebb-solve never calls them (the reference's solver uses the scalar and LP
// limiters of source/solve.d); app/ms.d:301-305 names them by the config key "limiter".
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdint>

#include "../../include/ebbcfd.h"

namespace ebb {

// std.math.sgn: +1, -1, or the argument itself for +-0 and NaN
__device__ __forceinline__ double sgnD(double x) { return x > 0 ? 1.0 : (x < 0 ? -1.0 : x); }

__device__ __forceinline__ double limiter1d(int kind, double Sp, double Sm)
{
    switch (kind) {
    case 0: return 0.0;                                                              // firstOrderUpwindS
    case 1: return Sp;                                                               // laxWendroffS
    case 2: return fmax(0.0, fmin(fabs(Sm), fabs(Sp))) * sgnD(Sm + Sp);             // minmodS
    case 3: return (Sm * fabs(Sp) + fabs(Sm) * Sp) / (fabs(Sm) + fabs(Sp) + 10e-12); // harmonicS
    case 4: return sqrt((Sm * Sp + fabs(Sm * Sp)) / 2) * sgnD(Sm + Sp);              // geometricS
    default: return fmax(0.0, fmax(fmin(2 * fabs(Sm), fabs(Sp)), fmin(fabs(Sm), 2 * fabs(Sp)))) * sgnD(Sm + Sp);  // superbeeS
    }
}

__global__ __launch_bounds__(256) void k_limiter1d(int kind, uint64_t n, const double* __restrict__ sp,
                                                   const double* __restrict__ sm, double* __restrict__ out)
{
    const uint64_t i = uint64_t(blockIdx.x) * 256 + threadIdx.x;
    if (i < n) out[i] = limiter1d(kind, sp[i], sm[i]);
}

}  // namespace ebb

extern "C" int ebb_limiter1d(int device, int kind, int64_t n, const double* sp, const double* sm, double* out)
{
    if (kind < 0 || kind > 5 || n < 0 || (n > 0 && (!sp || !sm || !out))) return EBB_ERR_ARG;
    if (n == 0) return EBB_OK;
    if (hipSetDevice(device) != hipSuccess) return EBB_ERR_DEVICE;
    double* d = nullptr;
    const size_t bytes = size_t(n) * sizeof(double);
    if (hipMalloc(&d, 3 * bytes) != hipSuccess) return EBB_ERR_ALLOC;
    int rc = EBB_OK;
    if (hipMemcpy(d, sp, bytes, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d + n, sm, bytes, hipMemcpyHostToDevice) != hipSuccess) {
        rc = EBB_ERR_DEVICE;
    } else {
        const uint64_t blocks = (uint64_t(n) + 255) / 256;
        hipLaunchKernelGGL(ebb::k_limiter1d, dim3(uint32_t(blocks)), dim3(256), 0, 0, kind, uint64_t(n), d, d + n, d + 2 * n);
        if (hipGetLastError() != hipSuccess || hipMemcpy(out, d + 2 * n, bytes, hipMemcpyDeviceToHost) != hipSuccess)
            rc = EBB_ERR_DEVICE;
    }
    (void)hipFree(d);
    return rc;
}
