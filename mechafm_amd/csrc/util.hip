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

__device__ inline unsigned long long wave_sum(unsigned long long v) {
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Totals of a scan from its records: out[0] minimiser iterations, [1]
// unconverged positions, [2] grid-clamp warnings, [3] out-of-grid lookups
// (added to what out holds).
__global__ void __launch_bounds__(256) k_record_totals(const DevRecord* rec, long n, unsigned long long* out) {
    unsigned long long s = 0, u = 0, w = 0, o = 0;
    for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (long)gridDim.x * blockDim.x) {
        const int2 v = *reinterpret_cast<const int2*>(&rec[p].n_steps);
        s += (unsigned)v.x;
        u += (v.y & REC_UNCONVERGED) != 0;
        w += (v.y & REC_GRID_CLAMP) != 0;
        o += (v.y & REC_OUT_OF_GRID) != 0;
    }
    s = wave_sum(s);
    u = wave_sum(u);
    w = wave_sum(w);
    o = wave_sum(o);
    if ((threadIdx.x & 63) == 0) {
        if (s) atomicAdd(&out[0], s);
        if (u) atomicAdd(&out[1], u);
        if (w) atomicAdd(&out[2], w);
        if (o) atomicAdd(&out[3], o);
    }
}

}  // namespace

hipError_t launch_record_totals(const DevRecord* rec, long n, int64_t* out, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    long blocks = std::min<long>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_record_totals, dim3((unsigned)blocks), dim3(256), 0, st, rec, n,
                       reinterpret_cast<unsigned long long*>(out));
    return hipGetLastError();
}

hipError_t launch_to_f32(const double4* in, float4* out, long n, hipStream_t st) {
    if (n <= 0) return hipSuccess;
    long blocks = std::min<long>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_to_f32, dim3((unsigned)blocks), dim3(256), 0, st, in, out, n);
    return hipGetLastError();
}

}  // namespace mafm
