// 3D complex FFT (FP64) and the electrostatic-potential pipeline of
// ElectrostaticPotentialInteraction (interactions.cpp:53-234).
//
// FFT: mixed-radix Stockham autosort, radices 4, 2, 3, 5 and a direct DFT
// for any other prime factor.  A 3D transform is three batched 1D passes
// (z, then y, then x); each workgroup stages B lines in LDS ([index][line],
// the line index fastest so the strided axes load B contiguous elements per
// row), runs all radix passes between two LDS buffers and writes the lines
// back in place.  The twiddle table w_n^k = exp(-2 pi i k / n) is built on
// the host in double.  Forward is unnormalised with the e^{-2 pi i jk/n}
// sign of kiss_fft (kiss_fft.c:359-362); the inverse of the scan pipeline is
// scaled by 1/N afterwards as ffti_data_grid does (fft.cpp:74-84).
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "device_common.hpp"

namespace mafm {
namespace {

constexpr int FFT_THREADS = 256;
constexpr int MAX_FACTORS = 24;

struct AxisPlan {
    int n;
    int nf;
    int radix[MAX_FACTORS];
};

__device__ inline double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ inline double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
// multiply by -i (forward) or +i (inverse)
__device__ inline double2 mul_mi(double2 a, bool inv) {
    return inv ? make_double2(-a.y, a.x) : make_double2(a.y, -a.x);
}

// In-place DFT of v[0..R-1]: X_k = sum_r v_r w^{rk}, w = exp(-+2 pi i / R).
template <int R>
__device__ inline void dft_small(double2* v, bool inv, const double2* tw, int n) {
    if (R == 2) {
        double2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if (R == 4) {
        double2 a = v[0], b = v[1], c = v[2], d = v[3];
        double2 s0 = cadd(a, c), s1 = csub(a, c), s2 = cadd(b, d), s3 = mul_mi(csub(b, d), inv);
        v[0] = cadd(s0, s2);
        v[2] = csub(s0, s2);
        v[1] = cadd(s1, s3);
        v[3] = csub(s1, s3);
    } else if (R == 3) {
        const double c1 = -0.5, s1 = inv ? 0.86602540378443864676 : -0.86602540378443864676;
        double2 a = v[0], b = v[1], c = v[2];
        double2 t = cadd(b, c);
        double2 u = csub(b, c);
        double2 m = make_double2(a.x + c1 * t.x, a.y + c1 * t.y);
        double2 r = make_double2(-s1 * u.y, s1 * u.x);     // i * s1 * u
        v[0] = cadd(a, t);
        v[1] = cadd(m, r);
        v[2] = csub(m, r);
    } else if (R == 5) {
        const double c1 = 0.30901699437494742410, c2 = -0.80901699437494742410;
        const double sa = inv ? 0.95105651629515357212 : -0.95105651629515357212;
        const double sb = inv ? 0.58778525229247312917 : -0.58778525229247312917;
        double2 a = v[0];
        double2 t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
        double2 u1 = csub(v[1], v[4]), u2 = csub(v[2], v[3]);
        double2 m1 = make_double2(a.x + c1 * t1.x + c2 * t2.x, a.y + c1 * t1.y + c2 * t2.y);
        double2 m2 = make_double2(a.x + c2 * t1.x + c1 * t2.x, a.y + c2 * t1.y + c1 * t2.y);
        // i * (sa u1 + sb u2), i * (sb u1 - sa u2)
        double2 r1 = make_double2(-(sa * u1.y + sb * u2.y), sa * u1.x + sb * u2.x);
        double2 r2 = make_double2(-(sb * u1.y - sa * u2.y), sb * u1.x - sa * u2.x);
        v[0] = cadd(a, cadd(t1, t2));
        v[1] = cadd(m1, r1);
        v[4] = csub(m1, r1);
        v[2] = cadd(m2, r2);
        v[3] = csub(m2, r2);
    }
}

// Direct DFT for a prime radix R (<= MAX_GENERIC) with twiddles from the length-n table.
constexpr int MAX_GENERIC = 13;
__device__ inline void dft_generic(double2* v, int R, bool inv, const double2* tw, int n) {
    double2 o[MAX_GENERIC];
    const int step = n / R;
    for (int k = 0; k < R; ++k) {
        double2 acc = make_double2(0, 0);
        for (int r = 0; r < R; ++r) {
            double2 w = tw[((r * k) % R) * step];
            if (inv) w.y = -w.y;
            acc = cadd(acc, cmul(v[r], w));
        }
        o[k] = acc;
    }
    for (int k = 0; k < R; ++k) v[k] = o[k];
}

template <int R>
__device__ void stockham_pass(const double2* src, double2* dst, int n, int Ns, int B, int nb, bool inv,
                              const double2* tw) {
    const int m = n / R;
    for (int idx = threadIdx.x; idx < B * m; idx += blockDim.x) {
        const int b = idx % B;
        const int j = idx / B;
        if (b >= nb) continue;
        double2 v[R];
        const int jm = j % Ns;
        const int tstep = (n / (Ns * R)) * jm;   // angle -2 pi jm r / (Ns R) = index r*tstep of w_n
#pragma unroll
        for (int r = 0; r < R; ++r) {
            double2 x = src[(j + r * m) * B + b];
            if (r > 0 && jm > 0) {
                double2 w = tw[(r * tstep) % n];
                if (inv) w.y = -w.y;
                x = cmul(x, w);
            }
            v[r] = x;
        }
        dft_small<R>(v, inv, tw, n);
        const int idxD = (j / Ns) * Ns * R + jm;
#pragma unroll
        for (int r = 0; r < R; ++r) dst[(idxD + r * Ns) * B + b] = v[r];
    }
}

__device__ void stockham_pass_generic(const double2* src, double2* dst, int n, int R, int Ns, int B, int nb,
                                      bool inv, const double2* tw) {
    const int m = n / R;
    for (int idx = threadIdx.x; idx < B * m; idx += blockDim.x) {
        const int b = idx % B;
        const int j = idx / B;
        if (b >= nb) continue;
        double2 v[MAX_GENERIC];
        const int jm = j % Ns;
        const int tstep = (n / (Ns * R)) * jm;
        for (int r = 0; r < R; ++r) {
            double2 x = src[(j + r * m) * B + b];
            if (r > 0 && jm > 0) {
                double2 w = tw[(r * tstep) % n];
                if (inv) w.y = -w.y;
                x = cmul(x, w);
            }
            v[r] = x;
        }
        dft_generic(v, R, inv, tw, n);
        const int idxD = (j / Ns) * Ns * R + jm;
        for (int r = 0; r < R; ++r) dst[(idxD + r * Ns) * B + b] = v[r];
    }
}

// One batched 1D pass along an axis of length n:
//   element (i, b) of block blk is at base + i*s_i + b*s_b with
//   base = (blk / n_chunks) * outer_stride + (blk % n_chunks) * B * s_b.
__global__ void __launch_bounds__(FFT_THREADS)
k_fft_axis(double2* data, AxisPlan plan, const double2* tw, long s_i, long s_b, long outer_stride, int n_chunks,
           int nb_total, int B, int inv) {
    extern __shared__ double2 lds[];
    const int n = plan.n;
    double2* buf0 = lds;
    double2* buf1 = lds + (size_t)n * B;
    const int outer = blockIdx.x / n_chunks;
    const int chunk = blockIdx.x % n_chunks;
    const long base = outer * outer_stride + (long)chunk * B * s_b;
    const int nb = min(B, nb_total - chunk * B);
    const bool contiguous_lines = (s_i == 1);
    for (int idx = threadIdx.x; idx < B * n; idx += blockDim.x) {
        int i, b;
        if (contiguous_lines) { b = idx / n; i = idx % n; }
        else { i = idx / B; b = idx % B; }
        if (b < nb) buf0[i * B + b] = data[base + i * s_i + b * s_b];
    }
    __syncthreads();
    double2* src = buf0;
    double2* dst = buf1;
    int Ns = 1;
    for (int f = 0; f < plan.nf; ++f) {
        const int R = plan.radix[f];
        switch (R) {
            case 2: stockham_pass<2>(src, dst, n, Ns, B, nb, inv, tw); break;
            case 3: stockham_pass<3>(src, dst, n, Ns, B, nb, inv, tw); break;
            case 4: stockham_pass<4>(src, dst, n, Ns, B, nb, inv, tw); break;
            case 5: stockham_pass<5>(src, dst, n, Ns, B, nb, inv, tw); break;
            default: stockham_pass_generic(src, dst, n, R, Ns, B, nb, inv, tw); break;
        }
        __syncthreads();
        Ns *= R;
        double2* t = src;
        src = dst;
        dst = t;
    }
    for (int idx = threadIdx.x; idx < B * n; idx += blockDim.x) {
        int i, b;
        if (contiguous_lines) { b = idx / n; i = idx % n; }
        else { i = idx / B; b = idx % B; }
        if (b < nb) data[base + i * s_i + b * s_b] = src[i * B + b];
    }
}

__global__ void k_real_to_complex(const double* in, double2* out, long n) {
    for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (long)gridDim.x * blockDim.x)
        out[p] = make_double2(in[p], 0.0);
}

__global__ void k_zero2(double2* out, long n) {
    for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (long)gridDim.x * blockDim.x)
        out[p] = make_double2(0.0, 0.0);
}

// Gaussian tip charge on the grid (interactions.cpp:63-105): points
// ix in [-cx, cx], iy in [-cy, cy], iz in [-cz, cz) at positionAt(ix,iy,iz)
// (origin 0), value q*norm*exp(-0.5 r^2 / w^2), added at the periodic image
// (DataGrid::atPBC, data_grid.cpp:105-127).
__global__ void k_gauss_tip(double2* rho, int nx, int ny, int nz, int cx, int cy, int cz, int ortho,
                            double b0, double b1, double b2, double b3, double b4, double b5, double b6,
                            double b7, double b8, double qn, double w2) {
    const long bx = 2L * cx + 1, by = 2L * cy + 1, bz = 2L * cz;
    const long total = bx * by * bz;
    for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < total; p += (long)gridDim.x * blockDim.x) {
        const int iz = (int)(p % bz) - cz;
        const int iy = (int)((p / bz) % by) - cy;
        const int ix = (int)(p / (bz * by)) - cx;
        double x, y, z;
        if (ortho) {
            x = 0.0 + ix * b0;
            y = 0.0 + iy * b4;
            z = 0.0 + iz * b8;
        } else {
            x = 0.0 + (b0 * ix + b1 * iy + b2 * iz);
            y = 0.0 + (b3 * ix + b4 * iy + b5 * iz);
            z = 0.0 + (b6 * ix + b7 * iy + b8 * iz);
        }
        const double r2 = (x * x + y * y) + z * z;
        const double val = qn * exp(-0.5 * r2 / w2);
        int wx = ix % nx; if (wx < 0) wx += nx;
        int wy = iy % ny; if (wy < 0) wy += ny;
        int wz = iz % nz; if (wz < 0) wz += nz;
        atomicAdd(&rho[((long)wx * ny + wy) * nz + wz].x, val);
    }
}

// pot_k *= vol * rho_k (interactions.cpp:139-146)
__global__ void k_mul_vol(double2* pot, const double2* rho, double vol, long n) {
    for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (long)gridDim.x * blockDim.x) {
        double2 r = rho[p];
        double2 s = make_double2(vol * r.x, vol * r.y);
        pot[p] = cmul(pot[p], s);
    }
}

// temp = sum_a -2 pi i k_a[c] E_k with k_a = idx * kb_col_a, aliased to
// idx - n for idx >= n/2 (interactions.cpp:165-225).
__global__ void k_spectral_grad(const double2* ek, double2* out, int nx, int ny, int nz, int c, double k0x,
                                double k0y, double k0z, double k1x, double k1y, double k1z, double k2x,
                                double k2y, double k2z) {
    const long total = (long)nx * ny * nz;
    const double PI2 = 2.0 * 3.14159265358979323846;
    for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < total; p += (long)gridDim.x * blockDim.x) {
        int iz = (int)(p % nz);
        int iy = (int)((p / nz) % ny);
        int ix = (int)(p / ((long)nz * ny));
        double ka[3] = {ix * k0x, ix * k0y, ix * k0z};
        double kb[3] = {iy * k1x, iy * k1y, iy * k1z};
        double kc[3] = {iz * k2x, iz * k2y, iz * k2z};
        if (ix >= nx / 2) { ka[0] -= nx * k0x; ka[1] -= nx * k0y; ka[2] -= nx * k0z; }
        if (iy >= ny / 2) { kb[0] -= ny * k1x; kb[1] -= ny * k1y; kb[2] -= ny * k1z; }
        if (iz >= nz / 2) { kc[0] -= nz * k2x; kc[1] -= nz * k2y; kc[2] -= nz * k2z; }
        const double2 e = ek[p];
        double re = 0, im = 0;
        for (int a = 0; a < 3; ++a) {
            const double kk = a == 0 ? ka[c] : (a == 1 ? kb[c] : kc[c]);
            const double b = -PI2 * kk;        // (-2 pi i k) E = (-b E.im) + i (b E.re)
            re += -(b * e.y);
            im += b * e.x;
        }
        out[p] = make_double2(re, im);
    }
}

// Real part / N into component `slot` of the {fx,fy,fz,e} grid; flags an
// imaginary part / N above 1e-8 (the one-sided test of fft.cpp:79-83).
__global__ void k_extract(const double2* in, double* fe, int slot, double inv_n, long n, int* imag_flag) {
    for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (long)gridDim.x * blockDim.x) {
        const double2 v = in[p];
        fe[4 * p + slot] = v.x * inv_n;
        if (v.y * inv_n > 1.0e-8) *imag_flag = 1;
    }
}

__global__ void k_scale(double2* x, double s, long n) {
    for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (long)gridDim.x * blockDim.x) {
        x[p].x *= s;
        x[p].y *= s;
    }
}

unsigned grid_for(long n) { return (unsigned)std::max<long>(1, std::min<long>((n + 255) / 256, 8192)); }

AxisPlan make_plan(int n) {
    AxisPlan p;
    p.n = n;
    p.nf = 0;
    int m = n;
    while (m % 4 == 0 && p.nf < MAX_FACTORS) { p.radix[p.nf++] = 4; m /= 4; }
    while (m % 2 == 0 && p.nf < MAX_FACTORS) { p.radix[p.nf++] = 2; m /= 2; }
    for (int r = 3; m > 1 && p.nf < MAX_FACTORS; r += 2)
        while (m % r == 0 && p.nf < MAX_FACTORS) { p.radix[p.nf++] = r; m /= r; }
    return p;
}

}  // namespace

// Lengths 1..2048 whose prime factors are <= MAX_GENERIC.
bool fft_length_supported(int n) {
    if (n < 1 || n > 2048) return false;
    int m = n;
    for (int r = 2; m > 1; ++r) {
        while (m % r == 0) {
            if (r > MAX_GENERIC) return false;
            m /= r;
        }
    }
    return true;
}

// Twiddle table exp(-2 pi i k / n), k < n.
void fft_twiddles(int n, std::vector<double>& out) {
    out.resize(2 * (size_t)n);
    for (int k = 0; k < n; ++k) {
        long double a = -2.0L * 3.141592653589793238462643383279502884L * k / n;
        out[2 * k] = (double)cosl(a);
        out[2 * k + 1] = (double)sinl(a);
    }
}

// In-place 3D FFT of data[n0][n1][n2] (z fastest); tw[a] is the twiddle
// table of axis a on the device.
hipError_t fft3d_device(double2* data, const int* n, const double2* const* tw, bool inv, hipStream_t st) {
    const long n0 = n[0], n1 = n[1], n2 = n[2];
    for (int axis = 2; axis >= 0; --axis) {
        const int len = n[axis];
        if (len == 1) continue;
        AxisPlan plan = make_plan(len);
        int B = std::max(1, std::min(16, (48 * 1024) / (2 * len * (int)sizeof(double2))));
        long s_i, s_b, outer_stride;
        int n_outer, nb_total;
        if (axis == 2) { s_i = 1; s_b = n2; outer_stride = 0; n_outer = 1; nb_total = (int)(n0 * n1); }
        else if (axis == 1) { s_i = n2; s_b = 1; outer_stride = n1 * n2; n_outer = (int)n0; nb_total = (int)n2; }
        else { s_i = n1 * n2; s_b = 1; outer_stride = n2; n_outer = (int)n1; nb_total = (int)n2; }
        const int n_chunks = (nb_total + B - 1) / B;
        const size_t lds = 2 * (size_t)len * B * sizeof(double2);
        hipLaunchKernelGGL(k_fft_axis, dim3(n_outer * n_chunks), dim3(FFT_THREADS), lds, st, data, plan, tw[axis],
                           s_i, s_b, outer_stride, n_chunks, nb_total, B, inv ? 1 : 0);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return e;
    }
    return hipSuccess;
}

hipError_t launch_real_to_complex(const double* in, double2* out, long n, hipStream_t st) {
    hipLaunchKernelGGL(k_real_to_complex, dim3(grid_for(n)), dim3(256), 0, st, in, out, n);
    return hipGetLastError();
}

hipError_t launch_zero2(double2* out, long n, hipStream_t st) {
    hipLaunchKernelGGL(k_zero2, dim3(grid_for(n)), dim3(256), 0, st, out, n);
    return hipGetLastError();
}

hipError_t launch_gauss_tip(double2* rho, const int* n, const int* cut, bool ortho, const double* b, double qn,
                            double w2, hipStream_t st) {
    const long total = (2L * cut[0] + 1) * (2L * cut[1] + 1) * (2L * cut[2]);
    if (total <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_gauss_tip, dim3(grid_for(total)), dim3(256), 0, st, rho, n[0], n[1], n[2], cut[0], cut[1],
                       cut[2], ortho ? 1 : 0, b[0], b[1], b[2], b[3], b[4], b[5], b[6], b[7], b[8], qn, w2);
    return hipGetLastError();
}

hipError_t launch_mul_vol(double2* pot, const double2* rho, double vol, long n, hipStream_t st) {
    hipLaunchKernelGGL(k_mul_vol, dim3(grid_for(n)), dim3(256), 0, st, pot, rho, vol, n);
    return hipGetLastError();
}

hipError_t launch_spectral_grad(const double2* ek, double2* out, const int* n, int c, const double* kb,
                                hipStream_t st) {
    const long total = (long)n[0] * n[1] * n[2];
    // kb row-major; column a = (kb[a], kb[3+a], kb[6+a])
    hipLaunchKernelGGL(k_spectral_grad, dim3(grid_for(total)), dim3(256), 0, st, ek, out, n[0], n[1], n[2], c,
                       kb[0], kb[3], kb[6], kb[1], kb[4], kb[7], kb[2], kb[5], kb[8]);
    return hipGetLastError();
}

hipError_t launch_extract(const double2* in, double* fe, int slot, double inv_n, long n, int* flag, hipStream_t st) {
    hipLaunchKernelGGL(k_extract, dim3(grid_for(n)), dim3(256), 0, st, in, fe, slot, inv_n, n, flag);
    return hipGetLastError();
}

hipError_t launch_scale2(double2* x, double s, long n, hipStream_t st) {
    hipLaunchKernelGGL(k_scale, dim3(grid_for(n)), dim3(256), 0, st, x, s, n);
    return hipGetLastError();
}

}  // namespace mafm
