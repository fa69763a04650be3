// Adjoint of the scan by the implicit function theorem.
//
// At a relaxed tip p* the total tip force vanishes, F(p*, theta) = 0, so
// dp*/dtheta = -J^-1 dF/dtheta with J = dF/dp (3x3, the negative Hessian of
// the tip energy: tip-surface pairs + tip-dummy VDW / Coulomb + lateral
// spring, interactions.cpp:15-51, 252-266).  The reported image value is the
// tip-surface force Fz_ts(p*, theta) (system.cpp:41-51), hence for a loss
// with g = dL/dFz:
//     dL/dtheta = sum_pos v . dF_ts/dtheta,   v = g (e_z - w),   J w = grad_p Fz_ts.
//
// k_adjoint_vec   one lane per position: J and grad_p Fz_ts from all sites,
//                 3x3 solve, v.
// k_adjoint_site  one workgroup per site: sum over positions of
//                 v . dF_i/d(es6, es12, qq) and -J_i v (the gradient with
//                 respect to the atom position), reduced in the workgroup
//                 (deterministic order, no atomics).
// Pair arithmetic is FP32 (Real = float) or FP64; sums are FP64.
#include "device_common.hpp"

namespace mafm {
namespace {

struct Sym3 {
    double xx, xy, xz, yy, yz, zz;
};

// Radial pair force F = f(r) d and its Jacobian f I + fp d d^T for an LJ site
// with optional Coulomb (es6 = B, es12 = A):
//   f  = 12 A r^-14 - 6 B r^-8 + qq r^-3
//   fp = f'(r)/r = -168 A r^-16 + 48 B r^-10 - 3 qq r^-5
template <typename Real>
__device__ inline void lj_coul_radial(Real r2, Real A, Real B, Real qq, Real& f, Real& fp, Real& ir2, Real& ir3) {
    const Real i2 = Real(1) / r2;
    const Real i6 = i2 * i2 * i2;
    const Real i8 = i6 * i2;
    const Real i14 = i8 * i6;
    const Real ir = sqrt(i2);
    ir2 = i2;
    ir3 = ir * i2;
    f = Real(12) * A * i14 - Real(6) * B * i8 + qq * ir3;
    fp = Real(-168) * A * i14 * i2 + Real(48) * B * i8 * i2 - Real(3) * qq * ir3 * i2;
}

__device__ inline void add_radial(Sym3& J, double f, double fp, double dx, double dy, double dz) {
    J.xx += f + fp * dx * dx;
    J.yy += f + fp * dy * dy;
    J.zz += f + fp * dz * dz;
    J.xy += fp * dx * dy;
    J.xz += fp * dx * dz;
    J.yz += fp * dy * dz;
}

// Tip-dummy Jacobian: VDW (LJ or Morse), Coulomb and the lateral spring.
__device__ void tip_dummy_jacobian(cptr_mol M, double tx, double ty, double tz, double dmx, double dmy, double dmz,
                                   Sym3& J) {
    const double dx = tx - dmx, dy = ty - dmy, dz = tz - dmz;
    const double r2 = dx * dx + dy * dy + dz * dz;
    const double r = sqrt(r2);
    double f, fp;
    if (M->td_vdw == 0) {
        const double i2 = 1.0 / r2, i6 = i2 * i2 * i2, i8 = i6 * i2, i14 = i8 * i6;
        f = 12 * M->td_p1 * i14 - 6 * M->td_p0 * i8;
        fp = -168 * M->td_p1 * i14 * i2 + 48 * M->td_p0 * i8 * i2;
    } else {
        // F = c (x^2 - x)/r d, x = exp(-a (r - re)), c = 2 De a
        const double a = M->td_p1, c = 2 * M->td_p0 * M->td_p1;
        const double x = exp(-a * (r - M->td_p2));
        f = c * (x * x - x) / r;
        const double dfdr = c * ((-2 * a * x * x + a * x) / r - (x * x - x) / r2);
        fp = dfdr / r;
    }
    if (M->td_coul) {
        const double ir3 = 1.0 / (r2 * r);
        f += M->td_qq * ir3;
        fp += -3 * M->td_qq * ir3 / r2;
    }
    add_radial(J, f, fp, dx, dy, dz);
    // spring: F_xy = -2k (1 - r0/rl) d_xy
    const double rl = sqrt(dx * dx + dy * dy);
    if (rl > 1e-10) {
        const double k = M->td_k, r0 = M->td_r0;
        const double a = -2 * k * (1 - r0 / rl);
        const double b = -2 * k * r0 / (rl * rl * rl);
        J.xx += a + b * dx * dx;
        J.yy += a + b * dy * dy;
        J.xy += b * dx * dy;
    }
}

template <typename Real>
__global__ void __launch_bounds__(256) k_adjoint_vec(const ScanArgs a, int mol, long n, const double* tip,
                                                      const double* dummy, const double* g, double* v_out,
                                                      int* singular) {
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    cptr_mol M = (cptr_mol)(a.mols + mol);
    const double tx = tip[3 * p], ty = tip[3 * p + 1], tz = tip[3 * p + 2];
    Sym3 Jts = {0, 0, 0, 0, 0, 0};
    const Site64* s = a.sites_ref + M->ref_off;
    for (int im = 0; im < M->n_images; ++im)
    for (int i = 0; i < M->n_sites; ++i) {
        const Site64 st = s[i];
        const double ax = st.x + M->shift[im][0], ay = st.y + M->shift[im][1], az = st.z + M->shift[im][2];
        const Real dx = (Real)(tx - ax), dy = (Real)(ty - ay), dz = (Real)(tz - az);
        const Real r2 = dx * dx + dy * dy + dz * dz;
        Real f, fp, ir2, ir3;
        lj_coul_radial<Real>(r2, (Real)st.p1, (Real)st.p0, (Real)(st.coul ? st.qq : 0.0), f, fp, ir2, ir3);
        add_radial(Jts, (double)f, (double)fp, (double)dx, (double)dy, (double)dz);
    }
    Sym3 J = Jts;
    tip_dummy_jacobian(M, tx, ty, tz, dummy[3 * p], dummy[3 * p + 1], dummy[3 * p + 2], J);
    // solve J w = u, u = grad_p Fz_ts = row z of Jts (J symmetric)
    const double ux = Jts.xz, uy = Jts.yz, uz = Jts.zz;
    const double c00 = J.yy * J.zz - J.yz * J.yz;
    const double c01 = J.xz * J.yz - J.xy * J.zz;
    const double c02 = J.xy * J.yz - J.xz * J.yy;
    const double c11 = J.xx * J.zz - J.xz * J.xz;
    const double c12 = J.xy * J.xz - J.xx * J.yz;
    const double c22 = J.xx * J.yy - J.xy * J.xy;
    const double det = J.xx * c00 + J.xy * c01 + J.xz * c02;
    double wx = 0, wy = 0, wz = 0;
    const double scale = fabs(J.xx) + fabs(J.yy) + fabs(J.zz);
    if (fabs(det) > 1e-300 && fabs(det) > 1e-14 * scale * scale * scale) {
        const double id = 1.0 / det;
        wx = (c00 * ux + c01 * uy + c02 * uz) * id;
        wy = (c01 * ux + c11 * uy + c12 * uz) * id;
        wz = (c02 * ux + c12 * uy + c22 * uz) * id;
    } else {
        *singular = 1;
    }
    const double gp = g[p];
    v_out[3 * p] = -gp * wx;
    v_out[3 * p + 1] = -gp * wy;
    v_out[3 * p + 2] = gp * (1.0 - wz);
}

template <typename Real>
__global__ void __launch_bounds__(256) k_adjoint_site(const ScanArgs a, int mol, long n, const double* tip,
                                                       const double* v, double* out6) {
    const int site = blockIdx.x;
    cptr_mol M = (cptr_mol)(a.mols + mol);
    const Site64 st = a.sites_ref[M->ref_off + site];
    const Real A = (Real)st.p1, B = (Real)st.p0, qq = (Real)(st.coul ? st.qq : 0.0);
    double acc[6] = {0, 0, 0, 0, 0, 0};   // d es6, d es12, d qq, d ax, d ay, d az
    const int n_img = M->n_images;
    for (long pi = threadIdx.x; pi < n * n_img; pi += blockDim.x) {
        const long p = pi / n_img;
        const int im = (int)(pi - p * n_img);
        const double ax = st.x + M->shift[im][0], ay = st.y + M->shift[im][1], az = st.z + M->shift[im][2];
        const Real dx = (Real)(tip[3 * p] - ax), dy = (Real)(tip[3 * p + 1] - ay), dz = (Real)(tip[3 * p + 2] - az);
        const Real r2 = dx * dx + dy * dy + dz * dz;
        Real f, fp, ir2, ir3;
        lj_coul_radial<Real>(r2, A, B, qq, f, fp, ir2, ir3);
        const double vx = v[3 * p], vy = v[3 * p + 1], vz = v[3 * p + 2];
        const double vd = vx * (double)dx + vy * (double)dy + vz * (double)dz;
        const double i2 = (double)ir2, i8 = i2 * i2 * i2 * i2, i14 = i8 * i2 * i2 * i2;
        acc[0] += -6 * i8 * vd;                   // dF/d es6 = -6 r^-8 d
        acc[1] += 12 * i14 * vd;                  // dF/d es12 = 12 r^-14 d
        acc[2] += (double)ir3 * vd;               // dF/d qq = r^-3 d
        const double fd = (double)f, fpd = (double)fp;
        acc[3] -= fd * vx + fpd * vd * (double)dx;  // dF/da = -(f I + fp d d^T)
        acc[4] -= fd * vy + fpd * vd * (double)dy;
        acc[5] -= fd * vz + fpd * vd * (double)dz;
    }
    __shared__ double red[6][256];
    for (int q = 0; q < 6; ++q) red[q][threadIdx.x] = acc[q];
    __syncthreads();
    for (int w = blockDim.x / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int q = 0; q < 6; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x < 6) out6[6 * site + threadIdx.x] = red[threadIdx.x][0];
}

}  // namespace

hipError_t launch_adjoint(const ScanArgs& a, int mol, int n_sites, long n, const double* tip, const double* dummy,
                          const double* g, double* v, double* out6, int* singular, bool f32, hipStream_t st) {
    const int blocks = (int)((n + 255) / 256);
    if (f32)
        hipLaunchKernelGGL(k_adjoint_vec<float>, dim3(blocks), dim3(256), 0, st, a, mol, n, tip, dummy, g, v, singular);
    else
        hipLaunchKernelGGL(k_adjoint_vec<double>, dim3(blocks), dim3(256), 0, st, a, mol, n, tip, dummy, g, v, singular);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || n_sites == 0) return e;
    if (f32)
        hipLaunchKernelGGL(k_adjoint_site<float>, dim3(n_sites), dim3(256), 0, st, a, mol, n, tip, v, out6);
    else
        hipLaunchKernelGGL(k_adjoint_site<double>, dim3(n_sites), dim3(256), 0, st, a, mol, n, tip, v, out6);
    return hipGetLastError();
}

}  // namespace mafm
