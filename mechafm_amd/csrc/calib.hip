// VALU issue-rate calibration (no reference counterpart): independent chains
// of FMAs in FP32 or FP64, or of FP64 reciprocal square roots each followed
// by an addition, over a fixed grid of 32 768 wave64s.  The instruction count
// of each launch is known exactly, so timing it (and reading the
// SQ_INSTS_VALU / SQ_BUSY_CYCLES counters of it) calibrates those counters
// against the rate at which a SIMD issues VALU instructions.
#include <hip/hip_runtime.h>

#include "device_common.hpp"

namespace mafm {
namespace {

constexpr int CAL_CHAINS = 8;
constexpr int CAL_BLOCKS = 8192;   // 256-thread blocks: 32 768 wave64s
constexpr int CAL_THREADS = 256;

template <typename T>
__global__ void __launch_bounds__(CAL_THREADS) k_calib_fma(T* out, T b, T c, int n) {
    T a[CAL_CHAINS];
#pragma unroll
    for (int u = 0; u < CAL_CHAINS; ++u) a[u] = (T)(threadIdx.x + u);
    for (int i = 0; i < n; ++i) {
#pragma unroll
        for (int u = 0; u < CAL_CHAINS; ++u) a[u] = fma(a[u], b, c);
    }
    T s = 0;
#pragma unroll
    for (int u = 0; u < CAL_CHAINS; ++u) s += a[u];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

__global__ void __launch_bounds__(CAL_THREADS) k_calib_rsq(double* out, double b, int n) {
    double a[CAL_CHAINS];
#pragma unroll
    for (int u = 0; u < CAL_CHAINS; ++u) a[u] = 1.0 + threadIdx.x + u;
    for (int i = 0; i < n; ++i) {
#pragma unroll
        for (int u = 0; u < CAL_CHAINS; ++u) a[u] = __builtin_amdgcn_rsq(a[u]) + b;
    }
    double s = 0;
#pragma unroll
    for (int u = 0; u < CAL_CHAINS; ++u) s += a[u];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

}  // namespace

size_t valu_calibration_doubles() { return (size_t)CAL_BLOCKS * CAL_THREADS; }

// kind 0: FP32 FMA chains, 1: FP64 FMA chains, 2: FP64 v_rsq + add chains;
// n iterations of every chain.  out: valu_calibration_doubles() doubles.
// *wave_insts: the launch's VALU wave instructions in the chains.
hipError_t launch_valu_calibration(int kind, int n, double* out, int64_t* wave_insts, hipStream_t st) {
    const int64_t waves = (int64_t)CAL_BLOCKS * CAL_THREADS / 64;
    if (kind == 0) {
        hipLaunchKernelGGL(k_calib_fma<float>, dim3(CAL_BLOCKS), dim3(CAL_THREADS), 0, st, reinterpret_cast<float*>(out),
                           0.999f, 0.001f, n);
        *wave_insts = waves * CAL_CHAINS * n;
    } else if (kind == 1) {
        hipLaunchKernelGGL(k_calib_fma<double>, dim3(CAL_BLOCKS), dim3(CAL_THREADS), 0, st, out, 0.999, 0.001, n);
        *wave_insts = waves * CAL_CHAINS * n;
    } else if (kind == 2) {
        hipLaunchKernelGGL(k_calib_rsq, dim3(CAL_BLOCKS), dim3(CAL_THREADS), 0, st, out, 0.5, n);
        *wave_insts = waves * CAL_CHAINS * n * 2;
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace mafm
