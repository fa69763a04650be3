// Fast scan: persistent, work-stealing, one force evaluation per loop trip.
//
// Work decomposition
//   * A workgroup of W waves (W = 1 .. 8) owns one molecule at a time and
//     stages its sites once in LDS; the molecule index is wave-uniform, so its
//     DevMol record is read with scalar loads into SGPRs.  A molecule too
//     large for LDS is streamed through scalar loads instead (G = 1).
//   * G lanes form a pixel group; the G lanes split the molecule's sites and
//     combine their partial sums with DPP lane moves.  Groups of a wave read
//     the same sites (LDS broadcast).
//   * Each group steals pixel columns from its molecule's atomic counter
//     (one atomic per wave per refill, ballot + prefix count).  Molecules
//     are visited longest columns first; when a molecule runs dry the
//     workgroup moves to the next one with work left.
//     A wave whose groups have mostly finished regroups the remaining
//     positions over all 64 lanes.
//   * The minimiser / integrator of simulation.cpp:102-122 runs as a per-group
//     state machine in registers.  Every trip of the main loop performs exactly
//     one evaluation of the interaction list per group: an integrator
//     sub-step (midpoint 2, RK4 4, Euler 1, SD 1 per iteration) or the final
//     tip-surface evaluation of System::getOutput.  Groups never wait for
//     each other at convergence, only at the end of the launch.
//
// Arithmetic
//   * Minimiser state, convergence tests (including the etol energy
//     difference), tip-dummy terms and the integrator updates are FP64 and
//     follow minimiser.cpp:67-129 / integrators.cpp:8-111.
//   * Pair terms: F32 = FP32 with one hardware reciprocal or reciprocal
//     square root per site, summed in FP32 blocks of 4-6 sites folded into FP64
//     (ACC_BLOCK), Kahan-compensated (ACC_KAHAN) or plain (ACC_NONE);
//     F64 = FP64 with Newton-refined hardware reciprocals.
#pragma once

#include "device_common.hpp"

namespace mafm {
namespace {

constexpr int PH_OUT = 7;

__device__ inline double rcp64(double x) {
    double r = __builtin_amdgcn_rcp(x);
    r = fma(r, fma(-x, r, 1.0), r);
    r = fma(r, fma(-x, r, 1.0), r);
    return r;
}
__device__ inline double rsq64(double x) {
    double y = __builtin_amdgcn_rsq(x);
    double hx = 0.5 * x;
    y = y * fma(-hx * y, y, 1.5);
    y = y * fma(-hx * y, y, 1.5);
    return y;
}

// Pair-term reciprocals: the hardware seeds are good to ~5e-8 relative
// (v_rsq_f64, v_rcp_f64 on gfx950), one Newton step takes them to ~4e-15,
// two orders below what the FP64 trajectories need to keep the reference's
// FIRE decisions.
#ifndef MAFM_PAIR64_NEWTON
#define MAFM_PAIR64_NEWTON 1
#endif
__device__ inline double rsq64p(double x) {
    if (MAFM_PAIR64_NEWTON == 2) return rsq64(x);
    const double y = __builtin_amdgcn_rsq(x);
    const double e = fma(-x * y, y, 1.0);          // 1 - x y^2
    return fma(0.5 * y, e, y);
}
__device__ inline double rcp64p(double x) {
    if (MAFM_PAIR64_NEWTON == 2) return rcp64(x);
    const double r = __builtin_amdgcn_rcp(x);
    return fma(r, fma(-x, r, 1.0), r);
}

// Site source of an evaluation: the copy of the molecule's table staged in
// the workgroup's LDS (the G lanes of a pixel read G different sites, groups
// read the same ones: broadcast), or, SC = true, for a molecule too large for
// LDS with one lane per pixel, the table itself through scalar loads (every
// lane needs the same site).
template <typename T>
__device__ inline T load_const(const T* p) {
    static_assert(sizeof(T) % 4 == 0, "dword-sized records only");
    T out;
    cptr<uint32_t> src = (cptr<uint32_t>)(p);
    uint32_t* dst = reinterpret_cast<uint32_t*>(&out);
#pragma unroll
    for (int i = 0; i < (int)(sizeof(T) / 4); ++i) dst[i] = src[i];
    return out;
}

template <bool SC, typename S>
struct Sites {
    const S* g;        // global table of the molecule (SC)
    const S* l;        // LDS copy
    __device__ inline S operator[](int t) const {
        if (SC) return load_const(g + t);
        return l[t];
    }
    __device__ inline void advance(int n) {
        g += n;
        l += n;
    }
};

// FP32 sites as two arrays so a pair reads only what it needs: A = {x, y, z,
// 6 es6 | 12 De} (16 B) and B = {12 es12 | a log2(e), qq, - | a log2(e) re,
// - | 2 De a};
// an LJ pair reads 20 B, LJ + Coulomb 24 B, Morse 32 B.
template <bool SC>
struct Sites32 {
    const float4* ga;
    const float4* gb;
    const float4* la;
    const float4* lb;
    __device__ inline float4 A(int t) const { return SC ? load_const(ga + t) : la[t]; }
    __device__ inline float B1(int t) const {
        return SC ? load_const(reinterpret_cast<const float*>(gb + t)) : reinterpret_cast<const float*>(lb + t)[0];
    }
    __device__ inline float2 B2(int t) const {
        return SC ? load_const(reinterpret_cast<const float2*>(gb + t)) : reinterpret_cast<const float2*>(lb + t)[0];
    }
    __device__ inline float4 B4(int t) const { return SC ? load_const(gb + t) : lb[t]; }
    __device__ inline void advance(int n) {
        ga += n; gb += n; la += n; lb += n;
    }
};

// Accumulators of the FP32 pair sums.
//   ACC_BLOCK  FP32 partial sums over register blocks of sites, folded into FP64
//   ACC_KAHAN  Kahan-compensated FP32 sum, term by term
//   ACC_NONE   plain FP32 sum
enum { ACC_NONE = 0, ACC_KAHAN = 1, ACC_BLOCK = 2 };

template <int MODE>
struct Acc32 {
    float s[4], c[4];
    double d[4];
    __device__ void init() {
        for (int q = 0; q < 4; ++q) { s[q] = 0.f; c[q] = 0.f; d[q] = 0.0; }
    }
    __device__ inline void begin() {
        if (MODE == ACC_BLOCK)
            for (int q = 0; q < 4; ++q) s[q] = 0.f;
    }
    // E = false: the energy component was not accumulated
    template <bool E = true>
    __device__ inline void end() {
        if (MODE == ACC_BLOCK)
            for (int q = 0; q < (E ? 4 : 3); ++q) d[q] += (double)s[q];
    }
    // s[q] += u * v, fused where the mode allows
    __device__ inline void madd(int q, float u, float v) {
        if (MODE == ACC_KAHAN) add(q, u * v);
        else s[q] = fmaf(u, v, s[q]);
    }
    __device__ inline void add(int q, float x) {
        if (MODE == ACC_KAHAN) {
            float y = x - c[q];
            float t = s[q] + y;
            c[q] = (t - s[q]) - y;
            s[q] = t;
        } else {
            s[q] += x;
        }
    }
    __device__ double get(int q) const {
        if (MODE == ACC_BLOCK) return d[q];
        if (MODE == ACC_KAHAN) return (double)s[q] - (double)c[q];
        return (double)s[q];
    }
};

// One site of segment SEG in registers: a = {x, y, z, 6 es6 | 12 De}, b = {12 es12 |
// a log2(e), qq, - | a log2(e) re, - | 2 De a}; only the B words the segment uses
// are read.
template <int SEG, bool SC>
__device__ inline void load_site(const Sites32<SC>& s, int t, float4& a, float4& b) {
    a = s.A(t);
    if (SEG == 1) {
        b.x = s.B1(t);
    } else if (SEG == 0) {
        const float2 v = s.B2(t);
        b.x = v.x;
        b.y = v.y;
    } else {
        b = s.B4(t);
    }
}

// One site term in FP32 (LJ: interactions.cpp:15-27, Morse :29-40, Coulomb
// :42-51), accumulated into acc: the force on the tip into components 0-2
// and, when E, 12x the energy into component 3.  Site constants are
// pre-scaled so that each quantity is one FMA away from the powers of 1/r:
//   LJ     a.w = 6 es6, b.x = 12 es12
//          F/r^2 factor fl = 12 es12 / r^6 - 6 es6 (then / r^8),
//          12 E = (fl - 6 es6) / r^6
//   Morse  a.w = 12 De, b.x = a log2(e), b.z = a log2(e) re, b.w = 2 De a
//          x = exp(-a (r - re)) = exp2(b.z - b.x r),
//          F/r = 2 De a x (x - 1) / r, 12 E = 12 De (x - 1)^2
//   Coulomb b.y = qq: E = qq / r, F/r = E / r^2
// SEG 0 LJ + Coulomb, 1 LJ, 2 Morse + Coulomb, 3 Morse.
// pair32_d takes the separation d = tip - site and r2 = |d|^2 (computed as
// fma(dx, dx, fma(dy, dy, dz * dz)), so callers that hoist dz * dz out of an
// image loop get the same bits).
template <int SEG, bool E, int MODE>
__device__ inline void pair32_d(const float4 a, const float4 b, float dx, float dy, float dz, float r2,
                                Acc32<MODE>& acc) {
    float fs;
    if (SEG <= 1) {
        float ir = 0.f, ir2;
        if (SEG == 0) {
            ir = __builtin_amdgcn_rsqf(r2);
            ir2 = ir * ir;
        } else {
            ir2 = __builtin_amdgcn_rcpf(r2);
        }
        const float ir6 = ir2 * ir2 * ir2;
        const float fl = fmaf(b.x, ir6, -a.w);
        if (E) acc.madd(3, ir6, fl - a.w);
        if (SEG == 0) {
            const float ec = b.y * ir;
            fs = fmaf(fl, ir6, ec) * ir2;
            if (E) acc.madd(3, 12.f, ec);
        } else {
            fs = fl * ir6 * ir2;
        }
    } else {
        const float ir = __builtin_amdgcn_rsqf(r2);
        const float r = r2 * ir;
        const float x = __builtin_amdgcn_exp2f(fmaf(-b.x, r, b.z));   // hardware exp2; below 2^-126 flushes to 0
        fs = b.w * fmaf(x, x, -x) * ir;
        if (E) {
            const float xm = x - 1.f;
            acc.madd(3, a.w * xm, xm);
        }
        if (SEG == 2) {
            const float ec = b.y * ir;
            fs = fmaf(ec, ir * ir, fs);
            if (E) acc.madd(3, 12.f, ec);
        }
    }
    acc.madd(0, fs, dx);
    acc.madd(1, fs, dy);
    acc.madd(2, fs, dz);
}

template <int SEG, bool E, int MODE>
__device__ inline void pair32(const float4 a, const float4 b, float tx, float ty, float tz, Acc32<MODE>& acc) {
    const float dx = tx - a.x, dy = ty - a.y, dz = tz - a.z;
    pair32_d<SEG, E>(a, b, dx, dy, dz, fmaf(dx, dx, fmaf(dy, dy, dz * dz)), acc);
}

// Sites t = gl, gl + S, ... of one segment against every periodic image of
// the tip: full blocks of 6 (LJ) or 4 (Morse) sites per lane (loaded once,
// reused for all images, no bound test), then the remainder one site at a
// time.  The image
// shifts are wave-uniform (scalar loads); the tip of image im is
// t - shift[im] in FP32 relative to the molecule's origin.  IMG = false is
// the single zero-shift image of a non-periodic system; E = false skips the
// energy (the compiler drops its arithmetic).
template <int SEG, bool SC, int MODE, bool IMG, bool E>
__device__ inline void seg_sum32(cptr<DevMol> M, const Sites32<SC>& s, int n, int gl, int S, float tx, float ty,
                                 float tz, Acc32<MODE>& acc) {
    const int n_img = IMG ? M->n_images : 1;
    constexpr int B = SEG <= 1 ? 6 : 4;    // register blocks: 6 LJ sites (6 floats each), 4 Morse (8 floats)
    const int full = (n / (B * S)) * (B * S);
    // one FP32 partial per register block (and all its images), folded into FP64
    for (int t0 = 0; t0 < full; t0 += B * S) {
        float4 sa[B], sb[B];
#pragma unroll
        for (int u = 0; u < B; ++u) load_site<SEG, SC>(s, t0 + u * S + gl, sa[u], sb[u]);
        acc.begin();
        if (IMG && M->planar) {
            // in-plane images: dz and dz^2 of a site are the same for all of them
            float dz[B], dz2[B];
#pragma unroll
            for (int u = 0; u < B; ++u) {
                dz[u] = tz - sa[u].z;
                dz2[u] = dz[u] * dz[u];
            }
            for (int im = 0; im < n_img; ++im) {
                const float ix = tx - M->shift32[im][0], iy = ty - M->shift32[im][1];
#pragma unroll
                for (int u = 0; u < B; ++u) {
                    const float dx = ix - sa[u].x, dy = iy - sa[u].y;
                    pair32_d<SEG, E>(sa[u], sb[u], dx, dy, dz[u], fmaf(dx, dx, fmaf(dy, dy, dz2[u])), acc);
                }
            }
        } else {
            for (int im = 0; im < n_img; ++im) {
                float ix = tx, iy = ty, iz = tz;
                if (IMG) {
                    ix -= M->shift32[im][0];
                    iy -= M->shift32[im][1];
                    iz -= M->shift32[im][2];
                }
#pragma unroll
                for (int u = 0; u < B; ++u) {
                    pair32<SEG, E>(sa[u], sb[u], ix, iy, iz, acc);
                }
            }
        }
        acc.template end<E>();
    }
    if (!IMG) acc.begin();
    for (int t = full + gl; t < n; t += S) {
        float4 sa, sb;
        load_site<SEG, SC>(s, t, sa, sb);
        if (IMG) acc.begin();
        for (int im = 0; im < n_img; ++im) {
            float ix = tx, iy = ty, iz = tz;
            if (IMG) {
                ix -= M->shift32[im][0];
                iy -= M->shift32[im][1];
                iz -= M->shift32[im][2];
            }
            pair32<SEG, E>(sa, sb, ix, iy, iz, acc);
        }
        if (IMG) acc.template end<E>();
    }
    if (!IMG) acc.template end<E>();
}

template <bool SC, int MODE, bool IMG, bool E>
__device__ inline void seg_sums32(cptr<DevMol> M, const Sites32<SC> s0, float tx, float ty, float tz, int gl, int S,
                                  Acc32<MODE>& acc) {
    Sites32<SC> s = s0;
    int n = M->n_seg[0];
    seg_sum32<0, SC, MODE, IMG, E>(M, s, n, gl, S, tx, ty, tz, acc);
    s.advance(n);
    n = M->n_seg[1];
    seg_sum32<1, SC, MODE, IMG, E>(M, s, n, gl, S, tx, ty, tz, acc);
    s.advance(n);
    n = M->n_seg[2];
    if (n) seg_sum32<2, SC, MODE, IMG, E>(M, s, n, gl, S, tx, ty, tz, acc);
    s.advance(n);
    n = M->n_seg[3];
    if (n) seg_sum32<3, SC, MODE, IMG, E>(M, s, n, gl, S, tx, ty, tz, acc);
}

// Sum of all tip-surface site terms of molecule M at tip position (px,py,pz):
// this lane's share, sites gl, gl+S, ... of each segment (every image), S
// lanes per position.  Coordinates are relative to the molecule's origin so
// FP32 keeps ~1e-6 A resolution.
template <bool SC, int MODE, bool E>
__device__ void site_sums32(cptr<DevMol> M, const Sites32<SC> s0, double px, double py, double pz, int gl, int S,
                            double out[4]) {
    Acc32<MODE> acc;
    acc.init();
    const float tx = (float)(px - M->origin[0]);
    const float ty = (float)(py - M->origin[1]);
    const float tz = (float)(pz - M->origin[2]);
    if (M->n_images > 1)
        seg_sums32<SC, MODE, true, E>(M, s0, tx, ty, tz, gl, S, acc);
    else
        seg_sums32<SC, MODE, false, E>(M, s0, tx, ty, tz, gl, S, acc);
    out[0] = acc.get(0);
    out[1] = acc.get(1);
    out[2] = acc.get(2);
    out[3] = E ? acc.get(3) * (1.0 / 12.0) : 0.0;
}

// The FP64 site term, the same one-FMA forms as pair32: SiteF64 constants
// p0 = 6 es6 | 12 De, p1 = 12 es12 | a, p2 = - | re, p3 = - | 2 De a;
// returns the force and 12x the energy.
template <int SEG>
__device__ inline void pair64(const SiteF64& st, double tx, double ty, double tz, double& fx, double& fy,
                              double& fz, double& e12) {
    const double dx = tx - st.x, dy = ty - st.y, dz = tz - st.z;
    const double r2 = fma(dx, dx, fma(dy, dy, dz * dz));
    double fs;
    if (SEG <= 1) {
        double ir = 0, ir2;
        if (SEG == 0) {
            ir = rsq64p(r2);
            ir2 = ir * ir;
        } else {
            ir2 = rcp64p(r2);
        }
        const double ir6 = ir2 * ir2 * ir2;
        const double fl = fma(st.p1, ir6, -st.p0);   // 12 es12 / r^6 - 6 es6
        e12 = ir6 * (fl - st.p0);
        if (SEG == 0) {
            const double ec = st.qq * ir;
            fs = fma(fl, ir6, ec) * ir2;
            e12 = fma(12.0, ec, e12);
        } else {
            fs = fl * ir6 * ir2;
        }
    } else {
        const double ir = rsq64p(r2);
        const double r = r2 * ir;
        const double x = exp(st.p1 * (st.p2 - r));
        const double xm = x - 1.0;
        fs = st.p3 * x * xm * ir;
        e12 = st.p0 * xm * xm;
        if (SEG == 2) {
            const double ec = st.qq * ir;
            fs = fma(ec, ir * ir, fs);
            e12 = fma(12.0, ec, e12);
        }
    }
    fx = fs * dx;
    fy = fs * dy;
    fz = fs * dz;
}

// B: sites per register block (loaded once, evaluated against every image,
// B independent dependency chains).  The FIRE + midpoint kernels, built for
// three waves per SIMD, use 2: with 4 the block's FP64 temporaries push ~45
// values of the minimiser state into scratch, rewritten every loop trip.
// IMG = false (one image): the tip minus the image shift is formed once for
// the whole segment instead of once per block; the same values.
template <int SEG, bool SC, bool E, int B, bool IMG>
__device__ inline void seg_sum64(cptr<DevMol> M, const Sites<SC, SiteF64>& s, int n, int gl, int S, double px,
                                 double py, double pz, double acc[4]) {
    const int n_img = IMG ? M->n_images : 1;
    const double x0 = px - M->shift[0][0], y0 = py - M->shift[0][1], z0 = pz - M->shift[0][2];
    const int full = (n / (B * S)) * (B * S);
    for (int t0 = 0; t0 < full; t0 += B * S) {
        SiteF64 st[B];
#pragma unroll
        for (int u = 0; u < B; ++u) st[u] = s[t0 + u * S + gl];
        for (int im = 0; im < n_img; ++im) {
            const double tx = IMG ? px - M->shift[im][0] : x0;
            const double ty = IMG ? py - M->shift[im][1] : y0;
            const double tz = IMG ? pz - M->shift[im][2] : z0;
#pragma unroll
            for (int u = 0; u < B; ++u) {
                double fx, fy, fz, e;
                pair64<SEG>(st[u], tx, ty, tz, fx, fy, fz, e);
                acc[0] += fx; acc[1] += fy; acc[2] += fz;
                if (E) acc[3] += e;
            }
        }
    }
    for (int t = full + gl; t < n; t += S) {
        const SiteF64 st = s[t];
        for (int im = 0; im < n_img; ++im) {
            const double tx = IMG ? px - M->shift[im][0] : x0;
            const double ty = IMG ? py - M->shift[im][1] : y0;
            const double tz = IMG ? pz - M->shift[im][2] : z0;
            double fx, fy, fz, e;
            pair64<SEG>(st, tx, ty, tz, fx, fy, fz, e);
            acc[0] += fx; acc[1] += fy; acc[2] += fz;
            if (E) acc[3] += e;
        }
    }
}

template <bool SC, bool E, int B, bool IMG>
__device__ inline void seg_sums64(cptr<DevMol> M, const Sites<SC, SiteF64> s0, double px, double py, double pz,
                                  int gl, int S, double acc[4]) {
    Sites<SC, SiteF64> s = s0;
    int n = M->n_seg[0];
    seg_sum64<0, SC, E, B, IMG>(M, s, n, gl, S, px, py, pz, acc);
    s.advance(n);
    n = M->n_seg[1];
    seg_sum64<1, SC, E, B, IMG>(M, s, n, gl, S, px, py, pz, acc);
    s.advance(n);
    n = M->n_seg[2];
    if (n) seg_sum64<2, SC, E, B, IMG>(M, s, n, gl, S, px, py, pz, acc);
    s.advance(n);
    n = M->n_seg[3];
    if (n) seg_sum64<3, SC, E, B, IMG>(M, s, n, gl, S, px, py, pz, acc);
}

template <bool SC, bool E, int B>
__device__ void site_sums64(cptr<DevMol> M, const Sites<SC, SiteF64> s0, double px, double py, double pz, int gl,
                            int S, double out[4]) {
    double acc[4] = {0, 0, 0, 0};
    if (M->n_images > 1)
        seg_sums64<SC, E, B, true>(M, s0, px, py, pz, gl, S, acc);
    else
        seg_sums64<SC, E, B, false>(M, s0, px, py, pz, gl, S, acc);
    out[0] = acc[0];
    out[1] = acc[1];
    out[2] = acc[2];
    out[3] = acc[3] * (1.0 / 12.0);
}

// Trilinear lookup of ForceGrid::interpolate (force_grid.cpp:88-155) with the
// clamp / wrap rules of getGridPoint / getGridPointIndex / pbcGridPoint
// (force_grid.cpp:158-275).  Fractional coordinates in FP64; on a diagonal
// basis they are (p - offset) times the reciprocal spacing where the
// reference divides (a last-place difference that can only move a point
// across a cell face, where the interpolant is continuous).  FP32 samples
// are weighted and summed in FP32 (weights from the FP64 fractions), FP64
// samples in FP64.
template <bool F32>
__device__ void grid_lookup(const DevGrid& g, double px, double py, double pz, double out[4], int& flags) {
    double q[3];
    if (g.orthogonal) {
        q[0] = (px - g.offset[0]) * g.inv[0];
        q[1] = (py - g.offset[1]) * g.inv[4];
        q[2] = (pz - g.offset[2]) * g.inv[8];
    } else {
        mat_vec(g.inv, px - g.offset[0], py - g.offset[1], pz - g.offset[2], q[0], q[1], q[2]);
    }
    int gp[3];
    for (int d = 0; d < 3; ++d) {
        gp[d] = (int)floor(q[d]);
        if (!g.periodic) {
            if (gp[d] < 0) { gp[d] = 0; flags |= REC_GRID_CLAMP; }
            else if (gp[d] >= g.n[d]) { gp[d] = g.n[d] - 1; flags |= REC_GRID_CLAMP; }
        }
    }
    int i0[3], i1[3];
    for (int d = 0; d < 3; ++d) {
        int a0 = gp[d], a1 = gp[d] + 1;
        if (g.periodic) {   // pbcGridPoint; the division only for points outside the cell
            if (a0 < 0 || a0 >= g.n[d]) a0 -= (int)floor((double)a0 / g.n[d]) * g.n[d];
            if (a1 < 0 || a1 >= g.n[d]) a1 -= (int)floor((double)a1 / g.n[d]) * g.n[d];
        } else if (a1 >= g.n[d]) {
            flags |= REC_OUT_OF_GRID;
            a1 = g.n[d] - 1;
        }
        // final clamp: keeps every gather inside the grid even for non-finite positions
        i0[d] = a0 < 0 ? 0 : (a0 >= g.n[d] ? g.n[d] - 1 : a0);
        i1[d] = a1 < 0 ? 0 : (a1 >= g.n[d] ? g.n[d] - 1 : a1);
    }
    const double w[3] = {q[0] - gp[0], q[1] - gp[1], q[2] - gp[2]};
    // corners are summed in the order c = 0..7 (x from bit 2, y from bit 1,
    // z from bit 0)
    if (F32) {
        // one index for corner (i0, i0, i0) and one stride per axis (0 where
        // the upper corner is clamped onto the lower, negative where it
        // wraps): the eight addresses cost seven additions, and the eight
        // gathers are issued together
        const long sy = g.n[2], sx = (long)g.n[1] * g.n[2];
        const long base = i0[0] * sx + i0[1] * sy + i0[2];
        const long dx = (i1[0] - i0[0]) * sx, dy = (i1[1] - i0[1]) * sy, dz = i1[2] - i0[2];
        const long off[8] = {0, dz, dy, dy + dz, dx, dx + dz, dx + dy, dx + dy + dz};
        const float wf[3] = {(float)w[0], (float)w[1], (float)w[2]};
        const float4* pt = g.fe32 + base;
        float4 v[8];
#pragma unroll
        for (int c = 0; c < 8; ++c) v[c] = pt[off[c]];
        float acc[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int c = 0; c < 8; ++c) {
            const float wt = ((c & 4) ? wf[0] : 1.f - wf[0]) * ((c & 2) ? wf[1] : 1.f - wf[1]) *
                             ((c & 1) ? wf[2] : 1.f - wf[2]);
            acc[0] = fmaf(wt, v[c].x, acc[0]);
            acc[1] = fmaf(wt, v[c].y, acc[1]);
            acc[2] = fmaf(wt, v[c].z, acc[2]);
            acc[3] = fmaf(wt, v[c].w, acc[3]);
        }
        for (int k = 0; k < 4; ++k) out[k] += (double)acc[k];
    } else {
        double acc[4] = {0, 0, 0, 0};
        for (int c = 0; c < 8; ++c) {
            const int ix = (c & 4) ? i1[0] : i0[0];
            const int iy = (c & 2) ? i1[1] : i0[1];
            const int iz = (c & 1) ? i1[2] : i0[2];
            const double wt = ((c & 4) ? w[0] : 1 - w[0]) * ((c & 2) ? w[1] : 1 - w[1]) * ((c & 1) ? w[2] : 1 - w[2]);
            const double4 v = g.fe64[((long)ix * g.n[1] + iy) * g.n[2] + iz];
            acc[0] = fma(wt, v.x, acc[0]);
            acc[1] = fma(wt, v.y, acc[1]);
            acc[2] = fma(wt, v.z, acc[2]);
            acc[3] = fma(wt, v.w, acc[3]);
        }
        for (int k = 0; k < 4; ++k) out[k] += acc[k];
    }
}

// Tip-dummy VDW, Coulomb and lateral spring on the tip (simulation.cpp:503-515,
// interactions.cpp:15-51, 252-266), FP64 with Newton-refined hardware
// reciprocals in place of divisions and square roots.
__device__ void tip_dummy(cptr<DevMol> M, double dx_, double dy_, double dz_, double tx, double ty, double tz,
                          double out[4]) {
    const double rx = dx_ - tx, ry = dy_ - ty, rz = dz_ - tz;
    const double r2 = rx * rx + ry * ry + rz * rz;
    const double ir2 = rcp64(r2);
    double s, e;
    if (M->td_vdw == 0) {
        const double ir6 = ir2 * ir2 * ir2;
        const double ta = M->td_p1 * ir6 * ir6;
        const double tb = M->td_p0 * ir6;
        e = ta - tb;
        s = (12 * ta - 6 * tb) * ir2;
    } else {
        const double ir = rsq64(r2);
        const double x = exp(-M->td_p1 * (r2 * ir - M->td_p2));
        e = M->td_p0 * (x * x - 2 * x + 1);
        s = 2 * M->td_p0 * M->td_p1 * (x * x - x) * ir;
    }
    if (M->td_coul) {
        const double ir = rsq64(r2);
        e += M->td_qq * ir;
        s += M->td_qq * ir * ir2;
    }
    double fx = -s * rx, fy = -s * ry, fz = -s * rz;
    const double rl2 = rx * rx + ry * ry;
    if (rl2 > 1e-20) {
        const double irl = rsq64(rl2);
        const double dr = rl2 * irl - M->td_r0;
        e += M->td_k * dr * dr;
        const double h = -2 * M->td_k * dr * irl;
        fx -= h * rx;
        fy -= h * ry;
    } else {
        const double dr = sqrt(rl2) - M->td_r0;
        e += M->td_k * dr * dr;
    }
    out[0] += fx; out[1] += fy; out[2] += fz; out[3] += e;
}

// Position of the k-th (0-based) set bit of m; k < popcount(m).
__device__ inline int nth_set_bit(uint64_t m, int k) {
    int pos = 0;
#pragma unroll
    for (int w = 32; w >= 1; w >>= 1) {
        const int c = __popcll((m >> pos) & ((1ull << w) - 1));
        if (k >= c) {
            k -= c;
            pos += w;
        }
    }
    return pos;
}

// v of the lane paired with this one by DPP pattern CTRL (a register move
// across lanes, no LDS round trip); every lane must be active.
template <int CTRL>
__device__ inline double dpp_f64(double v) {
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

// Sum of v over aligned groups of n lanes (n a power of two, wave-uniform),
// the same value on every lane of a group: butterflies within a quad (lane
// ^ 1, lane ^ 2), then mirrors within 8 and 16 lanes (after the quad sums
// every lane of a quad holds the same value, so a mirror pairs each lane
// with the other half), then cross-row shuffles.  Called with all lanes
// active.
__device__ inline double lanes_sum(double v, int n) {
    if (n >= 2) v += dpp_f64<0xB1>(v);     // quad_perm [1, 0, 3, 2]
    if (n >= 4) v += dpp_f64<0x4E>(v);     // quad_perm [2, 3, 0, 1]
    if (n >= 8) v += dpp_f64<0x141>(v);    // row_half_mirror: i <-> 7 - i
    if (n >= 16) v += dpp_f64<0x140>(v);   // row_mirror: i <-> 15 - i
    for (int m = 16; m < n; m <<= 1) v += __shfl_xor(v, m, 64);
    return v;
}

template <int G>
__device__ inline double group_sum(double v) {
    return lanes_sum(v, G);
}

// checkConvergence (minimiser.cpp:15-37); |F| < ftol tested as |F|^2 < ftol^2
__device__ inline bool converged(const ScanArgs& a, double fx, double fy, double fz, double ediff) {
    const double f2 = fx * fx + fy * fy + fz * fz;
    const double t2 = a.ftol * a.ftol;
    switch (a.minterm) {
        case 0: return fabs(ediff) < a.etol;
        case 1: return f2 < t2;
        default: return fabs(ediff) < a.etol && f2 < t2;
    }
}

// How near an iteration's decisions came to flipping, folded into the
// height's running minima (force units for f and p, energy units for e):
//   f  |f2 - ftol^2| / (2 ftol), the force's distance from ftol to first
//      order, counted when the energy part of the test passes;
//   e  ||ediff| - etol|, counted when the force part passes
//      (checkConvergence, minimiser.cpp:15-37);
// FIRE's P < 0 test (minimiser.cpp:107-126) is folded in by fire_margin.
__device__ inline void test_margins(const ScanArgs& a, double f2, double ediff, float& mgf, float& mge) {
    const double ae = fabs(ediff);
    if (a.minterm == 1 || (a.minterm == 2 && ae < a.etol))
        mgf = fminf(mgf, (float)(fabs(f2 - a.ftol * a.ftol) * a.inv_2ftol));
    if (a.minterm == 0 || (a.minterm == 2 && f2 < a.ftol * a.ftol)) mge = fminf(mge, (float)fabs(ae - a.etol));
}

// P = v.f over |v| (iv = 1 / |v|): the component of the force along the
// velocity, whose sign decides FIRE's velocity reset.  v = 0 gives P = 0
// exactly on any trajectory that also has v = 0, so no margin is counted.
__device__ inline void fire_margin(double p, double v2, double iv, float& mgp) {
    if (v2 > 0) mgp = fminf(mgp, (float)(fabs(p) * iv));
}

// Molecules are visited by rank (decreasing work per column: the longest
// columns start first and the tail of the launch is made of short ones).
// Rank of workgroup `block`'s first molecule from the initial shares.
__device__ int pick_rank(const ScanArgs& a, int block, int n_blocks) {
    const float target = (block + 0.5f) / n_blocks;
    int lo = 0, hi = a.n_mols - 1;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (a.mol_cum[mid + 1] > target) hi = mid; else lo = mid + 1;
    }
    return lo;
}

// Next rank after rk whose molecule has columns left, or -1.
__device__ int next_rank(const ScanArgs& a, int rk) {
    for (int q = 1; q <= a.n_mols; ++q) {
        int cand = rk + q;
        if (cand >= a.n_mols) cand -= a.n_mols;
        if (__atomic_load_n(&a.next_col[a.mol_order[cand]], __ATOMIC_RELAXED) < a.nx * a.ny) return cand;
    }
    return -1;
}

// blockDim.x = 64 W (W = 1 .. 8 waves).  The W waves of a workgroup share
// one LDS copy of their molecule's sites (so a large molecule costs its LDS
// once per W waves, not once per wave); each wave still steals columns on its
// own and the workgroup moves to the next molecule together once every wave
// has finished its last column.
// MID = true: FIRE with the midpoint integrator (the reference's default),
// fixed at compile time so the RK4 / Euler / SD state is not allocated.
template <int G, bool F32, int COMP, bool SC, bool MID>
__device__ __attribute__((always_inline)) inline void scan_fast_body(const ScanArgs& a) {
    static_assert(!SC || G == 1, "scalar site streaming serves one lane per pixel");
    constexpr int B64 = MID ? 2 : 4;   // FP64 sites per register block
    const int lane = threadIdx.x & 63;
    const bool shared_lds = !SC && blockDim.x > 64;
    const int gl = lane % G;
    const int gbase = lane - gl;
    const bool leader = gl == 0;
    const int n_cols = a.nx * a.ny;
    const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const bool track = a.margins != nullptr;

    int rk = __builtin_amdgcn_readfirstlane(pick_rank(a, blockIdx.x, gridDim.x));
    int m = a.mol_order[rk];
    extern __shared__ float4 lds_raw[];
    __shared__ int next_shared;
    int lds_n = 0;   // site count of the molecule staged in LDS
    // copy molecule m's site segments into the workgroup's LDS (FP32: A array
    // then B array; FP64: SiteF64 records)
    auto stage = [&](int mol) {
        if (!SC) {
            cptr<DevMol> Ms = as_const(a.mols) + mol;
            const int n = Ms->n_seg[0] + Ms->n_seg[1] + Ms->n_seg[2] + Ms->n_seg[3];
            __syncthreads();
            if (F32) {
                for (int t = threadIdx.x; t < n; t += blockDim.x) {
                    lds_raw[t] = a.seg32a[Ms->seg_off + t];
                    lds_raw[n + t] = a.seg32b[Ms->seg_off + t];
                }
            } else {
                SiteF64* dst = reinterpret_cast<SiteF64*>(lds_raw);
                for (int t = threadIdx.x; t < n; t += blockDim.x) dst[t] = a.segf64[Ms->seg_off + t];
            }
            __syncthreads();
            lds_n = n;
        }
    };
    stage(m);
    int64_t trips = 0, wide_trips = 0;
    int32_t active = 0;

    // group state (identical on the G lanes of a group)
    int col = -1, k = 0, ph = 0, n_tot = 1, nnn = 0, flags = 0, i = 0, j = 0;
    double dmx = 0, dmy = 0, dmz = 0, tx = 0, ty = 0, tz = 0, vx = 0, vy = 0, vz = 0;
    double fx = 0, fy = 0, fz = 0, e = 0, prev_e = 0, dt = a.dt, alpha = 0.1;
    double ax = 0, ay = 0, az = 0, bx = 0, by = 0, bz = 0, cx = 0, cy = 0, cz = 0, ce = 0;
    double Px = 0, Py = 0, Pz = 1.0e3;
    bool ts = false;
    float mgf = INFINITY, mge = INFINITY, mgp = INFINITY;   // the height's decision margins
    const double dt_max = 10 * a.dt;
    const bool align = MID || (a.minimiser != 0 && a.integrator == 1);

    for (;;) {
        // ---- refill: groups without a column steal one from molecule m
        const bool need = leader && col == -1;
        const uint64_t need_mask = __ballot(need);
        if (need_mask) {
            const int first = __ffsll((unsigned long long)need_mask) - 1;
            int base = 0;
            if (lane == first) base = atomicAdd(&a.next_col[m], __popcll(need_mask));
            base = __shfl(base, first, 64);
            if (need) {
                const int c = base + __popcll(need_mask & lt_mask);
                col = c < n_cols ? c : -2;
            }
            const int gcol = (G > 1) ? __shfl(col, gbase, 64) : col;
            const bool fresh = (col == -1) && gcol >= 0;
            col = gcol;
            if (fresh || (need && col >= 0)) {
                // new pixel column: System copy, setDummyXY, setDummyZ (simulation.cpp:93-101)
                cptr<DevMol> M = as_const(a.mols) + m;
                i = col / a.ny;
                j = col - i * a.ny;
                const double x = i * a.dx, y = j * a.dy;
                double px = x, py = y;
                if (M->tip_pbc) wrap_xy(M->cell, M->cell_inv, M->offset, x, y, px, py);
                dmx = px; dmy = py; dmz = a.zhigh;
                tx = px; ty = py; tz = a.zhigh - M->tip_dummy_d;
                vx = vy = vz = 0;
                fx = fy = fz = 0;
                e = 0;
                k = 0;
                // begin z-step
                n_tot = 1; dt = a.dt; alpha = 0.1; nnn = 0; flags = 0;
                mgf = mge = mgp = INFINITY;
                if (n_tot >= a.maxsteps) { ph = PH_OUT; ts = true; Px = tx; Py = ty; Pz = tz; }
                else { prev_e = e; ph = 0; ts = false; Px = tx; Py = ty; Pz = tz; }
            }
        }
        if (!__any(col >= 0)) {
            // molecule m exhausted for this wave: move to one with columns left
            int nr;
            if (shared_lds) {
                // m's counter is spent, so every wave of the workgroup arrives
                // here after its last column; one lane picks for all of them
                __syncthreads();
                if (threadIdx.x == 0) next_shared = next_rank(a, rk);
                __syncthreads();
                nr = next_shared;
            } else {
                nr = next_rank(a, rk);
            }
            nr = __builtin_amdgcn_readfirstlane(nr);
            if (nr < 0) break;
            rk = nr;
            m = a.mol_order[rk];
            col = -1;
            stage(m);
            continue;
        }

        // ---- one evaluation of the interaction list at P for every group
        // FIRE + midpoint: every group of the wave runs half-steps (ph 0) on
        // even trips, full steps (ph 1) and final evaluations on odd trips; a
        // group out of step waits one trip (once per height).  The wave then
        // runs one branch of the state machine per trip, and even trips skip
        // the energy sums, which a half-step does not use.
        const bool odd = trips & 1;
        ++trips;
        const bool stall = align && col >= 0 && (odd ? ph == 0 : ph == PH_OUT);
        const bool live = col >= 0 && !stall;
        const bool need_e = !align || (odd && (a.minterm != 1 || __any(live && ph == PH_OUT)));
        active += live;
        cptr<DevMol> M = as_const(a.mols) + m;
        if constexpr (MID) {
            // the evaluation point follows from the state (half-step: t + dt/2 v,
            // the same expression midpointStep uses), so it is not carried in
            // registers from trip to trip
            if (ph == 1) {
                Px = tx + dt / 2 * vx; Py = ty + dt / 2 * vy; Pz = tz + dt / 2 * vz;
            } else {
                Px = tx; Py = ty; Pz = tz;
            }
        }
        double r[4];
        if (M->force_grid >= 0) {
            r[0] = r[1] = r[2] = r[3] = 0;
            grid_lookup<F32>(a.grids[M->force_grid], Px, Py, Pz, r, flags);
        } else {
            // Few groups left with work (a molecule or the whole scan running
            // dry): spread the n_act active positions over all 64 lanes,
            // Ge = 64 / pow2ceil(n_act) lanes each, when that cuts the sites
            // per lane by at least wide_tail (a.wide_tail = 0: never).
            const uint64_t lead_act = __ballot(leader && live);
            const int n_act = __popcll(lead_act);
            const int p2 = n_act <= 1 ? 1 : 1 << (32 - __builtin_clz((unsigned)(n_act - 1)));   // pow2ceil
            const int Ge = 64 / p2;
            const bool wide = G < 64 && a.wide_tail > 0 && Ge >= a.wide_tail * G;
            if (wide) {
                ++wide_trips;
                const int slot = lane / Ge, sub = lane & (Ge - 1);
                const bool slot_live = slot < n_act;
                const int src = slot_live ? nth_set_bit(lead_act, slot) : 0;
                const double qx = __shfl(Px, src, 64), qy = __shfl(Py, src, 64), qz = __shfl(Pz, src, 64);
                double part[4] = {0, 0, 0, 0};
                if (slot_live) {
                    if constexpr (F32) {
                        // SC keeps the sites in global memory: wide loads read them there
                        const float4* wa = SC ? a.seg32a + M->seg_off : lds_raw;
                        const float4* wb = SC ? a.seg32b + M->seg_off : lds_raw + lds_n;
                        Sites32<false> src64{wa, wb, wa, wb};
                        if (need_e) site_sums32<false, COMP, true>(M, src64, qx, qy, qz, sub, Ge, part);
                        else site_sums32<false, COMP, false>(M, src64, qx, qy, qz, sub, Ge, part);
                    } else {
                        const SiteF64* w = SC ? a.segf64 + M->seg_off : reinterpret_cast<const SiteF64*>(lds_raw);
                        Sites<false, SiteF64> src64{w, w};
                        if (need_e) site_sums64<false, true, B64>(M, src64, qx, qy, qz, sub, Ge, part);
                        else site_sums64<false, false, B64>(M, src64, qx, qy, qz, sub, Ge, part);
                    }
                }
                // the energy (q = 3) only on trips that use it
#pragma unroll
                for (int q = 0; q < 3; ++q) part[q] = lanes_sum(part[q], Ge);
                if (need_e) part[3] = lanes_sum(part[3], Ge);
                // group of rank k (k-th active leader) reads slot k
                const int rank = __popcll(lead_act & ((1ull << gbase) - 1));
                const int from = (rank * Ge) & 63;
#pragma unroll
                for (int q = 0; q < 3; ++q) r[q] = __shfl(part[q], from, 64);
                r[3] = need_e ? __shfl(part[3], from, 64) : 0.0;
            } else {
                if constexpr (F32) {
                    Sites32<SC> src{a.seg32a + M->seg_off, a.seg32b + M->seg_off, lds_raw, lds_raw + lds_n};
                    if (need_e) site_sums32<SC, COMP, true>(M, src, Px, Py, Pz, gl, G, r);
                    else site_sums32<SC, COMP, false>(M, src, Px, Py, Pz, gl, G, r);
                } else {
                    Sites<SC, SiteF64> src{a.segf64 + M->seg_off, reinterpret_cast<const SiteF64*>(lds_raw)};
                    if (need_e) site_sums64<SC, true, B64>(M, src, Px, Py, Pz, gl, G, r);
                    else site_sums64<SC, false, B64>(M, src, Px, Py, Pz, gl, G, r);
                }
                if (G > 1) {
#pragma unroll
                    for (int q = 0; q < 3; ++q) r[q] = group_sum<G>(r[q]);
                    if (need_e) r[3] = group_sum<G>(r[3]);
                }
            }
            if (M->potential_grid >= 0) grid_lookup<F32>(a.grids[M->potential_grid], Px, Py, Pz, r, flags);
        }
        if (!ts) tip_dummy(M, dmx, dmy, dmz, Px, Py, Pz, r);

        if (!live) continue;

        // ---- state machine (FP64)
        bool step_done = false;
        if (ph == PH_OUT) {
            if (leader) {
                DevRecord rec;
                rec.x = i * a.dx; rec.y = j * a.dy; rec.z = dmz;
                rec.fx = r[0]; rec.fy = r[1]; rec.fz = r[2];
                rec.rx = tx - dmx; rec.ry = ty - dmy; rec.rz = tz - dmz;
                rec.e = r[3];
                rec.n_steps = n_tot;
                rec.flags = flags | (n_tot >= a.maxsteps ? REC_UNCONVERGED : 0);
                const long ri = M->rec_off + ((long)col * a.nz + k);
                a.rec[ri] = rec;
                if (a.margins) a.margins[ri] = make_float4(mgf, mge, mgp, 0.f);
                if (a.col_done) column_progress(a, (long)m * n_cols + col, k + 1);
            }
            dmz -= a.dz;
            tz -= a.dz;
            ++k;
            if (k >= a.nz) {
                col = -1;
                continue;
            }
            n_tot = 1; dt = a.dt; alpha = 0.1; nnn = 0; flags = 0;
            mgf = mge = mgp = INFINITY;
            if (n_tot >= a.maxsteps) { ph = PH_OUT; ts = true; }
            else { prev_e = e; ph = 0; ts = false; }
            Px = tx; Py = ty; Pz = tz;
            continue;
        }

        if (!MID && a.minimiser == 0) {
            // SDMinimisation (minimiser.cpp:39-65): evaluate, test, move
            fx = r[0]; fy = r[1]; fz = r[2]; e = r[3];
            if (track) test_margins(a, fx * fx + fy * fy + fz * fz, e - prev_e, mgf, mge);
            if (converged(a, fx, fy, fz, e - prev_e)) {
                ph = PH_OUT;
            } else {
                tx += a.dt * fx; ty += a.dt * fy; tz += a.dt * fz;
                ++n_tot;
                ph = n_tot >= a.maxsteps ? PH_OUT : 0;
                prev_e = e;
            }
            ts = ph == PH_OUT;
            Px = tx; Py = ty; Pz = tz;
            continue;
        }

        if (MID || a.integrator == 1) {                // midpointStep
            if (ph == 0) {
                Px = tx + dt / 2 * vx; Py = ty + dt / 2 * vy; Pz = tz + dt / 2 * vz;
                ax = vx + dt / 2 * r[0]; ay = vy + dt / 2 * r[1]; az = vz + dt / 2 * r[2];
                ph = 1;
            } else {
                tx += dt * ax; ty += dt * ay; tz += dt * az;
                vx += dt * r[0]; vy += dt * r[1]; vz += dt * r[2];
                fx = r[0]; fy = r[1]; fz = r[2]; e = r[3];
                step_done = true;
            }
        } else if (a.integrator == 0) {                // eulerStep
            tx += dt * vx; ty += dt * vy; tz += dt * vz;
            vx += dt * r[0]; vy += dt * r[1]; vz += dt * r[2];
            fx = r[0]; fy = r[1]; fz = r[2]; e = r[3];
            step_done = true;
        } else {                                       // rk4Step
            if (ph == 0) {
                cx = r[0]; cy = r[1]; cz = r[2]; ce = r[3];
                bx = vx; by = vy; bz = vz;
                Px = tx + dt / 2 * vx; Py = ty + dt / 2 * vy; Pz = tz + dt / 2 * vz;
                ax = vx + dt / 2 * r[0]; ay = vy + dt / 2 * r[1]; az = vz + dt / 2 * r[2];
                ph = 1;
            } else if (ph == 1) {
                cx += 2 * r[0]; cy += 2 * r[1]; cz += 2 * r[2]; ce += 2 * r[3];
                bx += 2 * ax; by += 2 * ay; bz += 2 * az;
                Px = tx + dt / 2 * ax; Py = ty + dt / 2 * ay; Pz = tz + dt / 2 * az;
                ax = vx + dt / 2 * r[0]; ay = vy + dt / 2 * r[1]; az = vz + dt / 2 * r[2];
                ph = 2;
            } else if (ph == 2) {
                cx += 2 * r[0]; cy += 2 * r[1]; cz += 2 * r[2]; ce += 2 * r[3];
                bx += 2 * ax; by += 2 * ay; bz += 2 * az;
                Px = tx + dt * ax; Py = ty + dt * ay; Pz = tz + dt * az;
                ax = vx + dt * r[0]; ay = vy + dt * r[1]; az = vz + dt * r[2];
                ph = 3;
            } else {
                cx += r[0]; cy += r[1]; cz += r[2]; ce += r[3];
                bx += ax; by += ay; bz += az;
                tx += dt / 6 * bx; ty += dt / 6 * by; tz += dt / 6 * bz;
                vx += dt / 6 * cx; vy += dt / 6 * cy; vz += dt / 6 * cz;
                fx = cx / 6; fy = cy / 6; fz = cz / 6; e = ce / 6;
                step_done = true;
            }
        }
        if (!step_done) continue;

        // FIREMinimisation iteration tail (minimiser.cpp:102-126)
        if (track) test_margins(a, fx * fx + fy * fy + fz * fz, e - prev_e, mgf, mge);
        if (converged(a, fx, fy, fz, e - prev_e)) {
            ph = PH_OUT;
        } else {
            const double p = vx * fx + vy * fy + vz * fz;
            const double f2 = fx * fx + fy * fy + fz * fz;
            double hx = 0, hy = 0, hz = 0;
            if (f2 > 1e-20) {                          // |f| > 1e-10 (vectors.hpp:107-114)
                const double ifl = rsq64(f2);
                hx = fx * ifl; hy = fy * ifl; hz = fz * ifl;
            }
            const double v2 = vx * vx + vy * vy + vz * vz;
            const double iv = v2 > 0 ? rsq64(v2) : 0.0;
            const double vl = v2 * iv;
            if (track) fire_margin(p, v2, iv, mgp);
            vx = (1 - alpha) * vx + alpha * hx * vl;
            vy = (1 - alpha) * vy + alpha * hy * vl;
            vz = (1 - alpha) * vz + alpha * hz * vl;
            if (p < 0) {
                vx = vy = vz = 0;
                nnn = 0;
                dt *= 0.5;
                alpha = 0.1;
            } else {
                if (nnn > 5) {
                    dt = fmin(dt * 1.1, dt_max);
                    alpha *= 0.99;
                }
                ++nnn;
            }
            ++n_tot;
            ph = n_tot >= a.maxsteps ? PH_OUT : 0;
            prev_e = e;
        }
        ts = ph == PH_OUT;
        Px = tx; Py = ty; Pz = tz;
    }
    // step, convergence and grid-warning totals are summed from the records
    // after the launch
    if (lane == 0) {
        atomicAdd((unsigned long long*)&a.counters[4], (unsigned long long)(trips * (64 / G)));
        if (wide_trips) atomicAdd((unsigned long long*)&a.counters[6], (unsigned long long)wide_trips);
    }
    if (leader && active) atomicAdd((unsigned long long*)&a.counters[5], (unsigned long long)active);
}

// Register budgets: the FIRE + midpoint FP32 kernel is held to three waves
// per SIMD (168 VGPRs; a few cold FIRE state values spill, which measured
// faster than two waves without spills), the others to two waves (left free,
// the compiler can take all 256 VGPRs and drop to one wave per SIMD).
template <int G, int COMP, bool SC, bool MID>
__global__ void __launch_bounds__(512, MID ? 3 : 2) k_scan_fast32(ScanArgs a) {
    scan_fast_body<G, true, COMP, SC, MID>(a);
}

#ifndef MAFM_F64_MID_WAVES
#define MAFM_F64_MID_WAVES 3
#endif
template <int G, bool SC, bool MID>
__global__ void __launch_bounds__(512, MID ? MAFM_F64_MID_WAVES : 2) k_scan_fast64(ScanArgs a) {
    scan_fast_body<G, false, 0, SC, MID>(a);
}

// More than 64 KiB of dynamic LDS per workgroup (a molecule's sites staged
// whole; gfx950 allows up to 160 KiB) has to be granted per kernel first.
template <int G, bool F32, int COMP, bool MID, bool SC = false>
hipError_t launch_one(const ScanArgs& a, int n_blocks, int W, size_t lds, hipStream_t st) {
    const size_t bytes = SC ? 0 : lds;
    if constexpr (F32) {
        if (bytes > 64 * 1024) {
            const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_scan_fast32<G, COMP, SC, MID>),
                                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
            if (e != hipSuccess) return e;
        }
        hipLaunchKernelGGL((k_scan_fast32<G, COMP, SC, MID>), dim3(n_blocks), dim3(64 * W), bytes, st, a);
    } else {
        if (bytes > 64 * 1024) {
            const hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_scan_fast64<G, SC, MID>),
                                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
            if (e != hipSuccess) return e;
        }
        hipLaunchKernelGGL((k_scan_fast64<G, SC, MID>), dim3(n_blocks), dim3(64 * W), bytes, st, a);
    }
    return hipGetLastError();
}

template <bool F32, int COMP, bool MID>
hipError_t launch_g(const ScanArgs& a, int G, bool scalar_sites, int n_blocks, int W, size_t lds, hipStream_t st) {
    if (scalar_sites) {
        if (G != 1) return hipErrorInvalidValue;
        return launch_one<1, F32, COMP, MID, true>(a, n_blocks, W, lds, st);
    }
    switch (G) {
        case 1: return launch_one<1, F32, COMP, MID>(a, n_blocks, W, lds, st);
        case 2: return launch_one<2, F32, COMP, MID>(a, n_blocks, W, lds, st);
        case 4: return launch_one<4, F32, COMP, MID>(a, n_blocks, W, lds, st);
        case 8: return launch_one<8, F32, COMP, MID>(a, n_blocks, W, lds, st);
        case 16: return launch_one<16, F32, COMP, MID>(a, n_blocks, W, lds, st);
        case 32: return launch_one<32, F32, COMP, MID>(a, n_blocks, W, lds, st);
        case 64: return launch_one<64, F32, COMP, MID>(a, n_blocks, W, lds, st);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace

// Launch of the fast kernels of one minimiser family (MID: FIRE + midpoint);
// each family is compiled in a translation unit of its own.
template <bool MID>
hipError_t launch_scan_fast_family(const ScanArgs& a, bool f32, int comp, int G, int n_blocks, int W, int lds_sites,
                                   bool scalar_sites, hipStream_t st) {
    if (W < 1 || W > 8) return hipErrorInvalidValue;
    if (!f32)
        return launch_g<false, 0, MID>(a, G, scalar_sites, n_blocks, W, (size_t)lds_sites * sizeof(SiteF64), st);
    const size_t lds = (size_t)lds_sites * 2 * sizeof(float4);
    if (comp == ACC_KAHAN) return launch_g<true, ACC_KAHAN, MID>(a, G, scalar_sites, n_blocks, W, lds, st);
    if (comp == ACC_NONE) return launch_g<true, ACC_NONE, MID>(a, G, scalar_sites, n_blocks, W, lds, st);
    return launch_g<true, ACC_BLOCK, MID>(a, G, scalar_sites, n_blocks, W, lds, st);
}

}  // namespace mafm
