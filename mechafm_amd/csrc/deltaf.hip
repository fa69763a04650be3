// Frequency-shift (Delta f) images from the scan records, the computation of
// simple-vafm-image.py:76-204 (Giessibl's weighting, Beilstein J.
// Nanotechnol. 3:238 (2012)) done on the device over the records the scan
// left in HBM instead of over re-parsed text files.
//
// The script's arithmetic is followed operation by operation, including the
// Python / numpy semantics its results depend on:
//   * the height step is the mean of the differences of the heights read
//     back from the file names ("scan-%06.3f.dat", mechafm.cpp:59-75),
//     summed the way numpy sums (pairwise, 8 partial sums per block of 128);
//   * the half-step amplitude adjustment tests (1 + A/dz) % (1 + A//dz)
//     with Python's float modulo and floor division;
//   * the window is numpy.arange(0, A + dz, dz) mirrored without its first
//     two points (zosc[-1:1:-1]), so the negative side lacks -dz, and both
//     end points are dropped (W[1:-1]);
//   * an image is -prefactor * sum(Fz * W) * dz with numpy's pairwise sum.
// Everything here is compiled without FMA contraction, so the products and
// sums round exactly as numpy's.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../include/mafm.h"
#include "device_common.hpp"

namespace mafm {

// Python float modulo and floor division (the sign of the result follows
// the divisor; floor division is consistent with the modulo).
double py_mod(double a, double b) {
    double m = std::fmod(a, b);
    if (m != 0.0) {
        if ((b < 0) != (m < 0)) m += b;
    } else {
        m = std::copysign(0.0, b);
    }
    return m;
}

double py_floordiv(double a, double b) {
    const double m = std::fmod(a, b);
    double d = (a - m) / b;
    if (m != 0.0 && ((b < 0) != (m < 0))) d -= 1.0;
    if (d == 0.0) return std::copysign(0.0, a / b);
    double f = std::floor(d);
    if (d - f > 0.5) f += 1.0;
    return f;
}

// numpy's pairwise summation of n doubles (blocks of at most 128 summed
// with 8 running partial sums, larger ranges split in halves).
double np_pairwise_sum(const double* a, int n) {
    if (n < 8) {
        double r = 0.0;
        for (int i = 0; i < n; ++i) r += a[i];
        return r;
    }
    if (n <= 128) {
        double r[8];
        for (int j = 0; j < 8; ++j) r[j] = a[j];
        int i = 8;
        for (; i < n - (n % 8); i += 8)
            for (int j = 0; j < 8; ++j) r[j] += a[i + j];
        double s = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
        for (; i < n; ++i) s += a[i];
        return s;
    }
    int n2 = n / 2;
    n2 -= n2 % 8;
    return np_pairwise_sum(a, n2) + np_pairwise_sum(a + n2, n - n2);
}

constexpr double NM_TO_KCAL_A2 = 1.438978;   // simple-vafm-image.py:80

// Weights of the window and the constants of an image; returns false with a
// message when the scan or the window does not fit.
// z_low (optional) receives the lowest height of every image's window, the
// height the script names the image by.
bool deltaf_weights(const mafm_deltaf_desc& d, int nz, double zhigh, double dz_input, mafm_deltaf_weights_out& o,
                    std::vector<double>* z_low, char* err, size_t err_len) {
    const double A0 = d.amplitude > 0 ? d.amplitude : 0.50;
    const double kc = (d.k_cantilever > 0 ? d.k_cantilever : 1800.0) * NM_TO_KCAL_A2;
    const double f0 = d.f0 > 0 ? d.f0 : 25.000e3;
    if (nz < 2) {
        snprintf(err, err_len, "a frequency shift needs at least two heights");
        return false;
    }
    // heights as the file names give them, in ascending order
    std::vector<double> zpos(nz);
    for (int i = 0; i < nz; ++i) {
        char name[64];
        snprintf(name, sizeof name, "%06.3f", zhigh - i * dz_input);
        zpos[nz - 1 - i] = std::strtod(name, nullptr);
    }
    std::vector<double> diff(nz - 1);
    for (int i = 0; i + 1 < nz; ++i) diff[i] = zpos[i + 1] - zpos[i];
    const double dz = np_pairwise_sum(diff.data(), nz - 1) / (double)(nz - 1);
    const double prefactor = (f0 / (2 * kc)) * (2 / (M_PI * A0 * A0));
    double A = A0;
    const double ampcut = py_mod(1 + (A / dz), 1 + py_floordiv(A, dz));
    if (ampcut < 0.01) A += dz / 2;
    const double len = std::ceil((A + dz - 0.0) / dz);
    if (!(len >= 3 && len <= 66)) {
        snprintf(err, err_len, "oscillation window of %.0f heights is outside 3..66 (amplitude %g, dz %g)", len, A, dz);
        return false;
    }
    const int n = (int)len;
    std::vector<double> zosc;
    for (int i = n - 1; i >= 2; --i) zosc.push_back(-(i * dz));
    for (int i = 0; i < n; ++i) zosc.push_back(i * dz);
    const int n_w = (int)zosc.size() - 2;
    o.n_weights = n_w;
    o.n_images = nz - n_w > 0 ? nz - n_w : 0;
    o.dz = dz;
    o.amplitude = A;
    o.prefactor = prefactor;
    const double a2 = A * A;
    for (int t = 0; t < n_w; ++t) {
        const double z = zosc[t + 1];
        o.weights[t] = z / std::sqrt(std::fabs(a2 - z * z));
    }
    for (int t = n_w; t < 128; ++t) o.weights[t] = 0.0;
    if (z_low) z_low->assign(zpos.begin(), zpos.begin() + o.n_images);
    return true;
}

namespace {

struct DeltaFArgs {
    double w[128];
    double neg_prefactor;
    double dz;
    int n_w, n_img, nx, ny, nz;
    int text;
};

// v as printf("%8.4f") prints it and float() reads it back: the exact binary
// value times 10^4 (p + err, err exact by FMA) rounded to an integer q, ties
// to even, then q / 10^4 rounded once (float() of a decimal string is the
// nearest double).  |v| 10^4 < 2^52 for any force here; larger values are
// returned as they are.
__device__ inline double text4(double v) {
    const double a = fabs(v);
    const double p = a * 1e4;
    if (!(p < 4503599627370496.0)) return v;
    const double err = fma(a, 1e4, -p);
    double fl = floor(p);
    const double t = ((p - fl) - 0.5) + err;
    if (t > 0 || (t == 0 && fmod(fl, 2.0) == 1.0)) fl += 1.0;
    const double q = fl / 1e4;
    return signbit(v) ? -q : q;
}

// One lane per (image, pixel): image p covers ascending heights p .. p+n_w-1,
// i.e. records k = nz-1-p .. nz-n_w-p of the pixel's column.
__global__ void __launch_bounds__(256) k_delta_f(const DevRecord* rec, DeltaFArgs a, double* out) {
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long total = (long)a.n_img * a.nx * a.ny;
    if (q >= total) return;
    const long pix = q % ((long)a.nx * a.ny);
    const int p = (int)(q / ((long)a.nx * a.ny));
    const DevRecord* col = rec + pix * a.nz;
    // np_pairwise_sum over the products, streamed (n_w <= 128: one block)
    auto term = [&](int t) {
        const double fz = col[a.nz - 1 - p - t].fz;
        return (a.text ? text4(fz) : fz) * a.w[t];
    };
    const int n = a.n_w;
    double s;
    if (n < 8) {
        s = 0.0;
        for (int t = 0; t < n; ++t) s += term(t);
    } else {
        double r[8];
        for (int j = 0; j < 8; ++j) r[j] = term(j);
        int i = 8;
        for (; i < n - (n % 8); i += 8)
            for (int j = 0; j < 8; ++j) r[j] += term(i + j);
        s = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
        for (; i < n; ++i) s += term(i);
    }
    out[q] = (a.neg_prefactor * s) * a.dz;
}

}  // namespace

hipError_t launch_delta_f(const DevRecord* rec, const mafm_deltaf_weights_out& w, int nx, int ny, int nz, bool text,
                          double* out, hipStream_t st) {
    DeltaFArgs a;
    a.text = text ? 1 : 0;
    for (int t = 0; t < 128; ++t) a.w[t] = w.weights[t];
    a.neg_prefactor = -w.prefactor;
    a.dz = w.dz;
    a.n_w = w.n_weights;
    a.n_img = w.n_images;
    a.nx = nx;
    a.ny = ny;
    a.nz = nz;
    const long total = (long)a.n_img * nx * ny;
    if (total == 0) return hipSuccess;
    hipLaunchKernelGGL(k_delta_f, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, rec, a, out);
    return hipGetLastError();
}

}  // namespace mafm
