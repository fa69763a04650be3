// Small elementwise kernels.
#include "device_common.hpp"

namespace mafm {
namespace {

// FP64 {fx,fy,fz,e} grid samples -> FP32 copy used by the FP32 scan.
__global__ void __launch_bounds__(256) k_to_f32(const double4* in, float4* out, long n) {
    for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (long)gridDim.x * blockDim.x) {
        double4 v = in[p];
        out[p] = make_float4((float)v.x, (float)v.y, (float)v.z, (float)v.w);
    }
}

}  // namespace

hipError_t launch_to_f32(const double4* in, float4* out, long n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    long blocks = std::min<long>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_to_f32, dim3((unsigned)blocks), dim3(256), 0, st, in, out, n);
    return hipGetLastError();
}

}  // namespace mafm
