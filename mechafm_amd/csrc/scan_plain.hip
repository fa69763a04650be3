// Plain FP64 scan: one lane per pixel column, the reference's arithmetic.
//
// This translation unit is compiled with -ffp-contract=off so every operation
// rounds exactly as the reference's -O3 x86-64 build does: the operator
// formulas, integrator updates and FIRE bookkeeping below keep the reference's
// evaluation order (interactions.cpp:15-51, 244-266; integrators.cpp:8-111;
// minimiser.cpp:15-129; force_grid.cpp:88-275) and use IEEE division and
// square root.  Its results agree with the reference to rounding, its FIRE
// step counts exactly.  It is also the "plain FP64 GPU scan" baseline: atoms
// are read straight from global memory and the nested z / FIRE loops run per
// lane, so lanes of a wave wait for the slowest pixel.
#include "device_common.hpp"
#include "ref_terms.hpp"

namespace mafm {
namespace {

struct Ctx {
    const ScanArgs* a;
    const DevMol* m;
    int flags;
};

// ForceGrid::interpolate with getGridPoint / getGridPointIndex / pbcGridPoint
// (force_grid.cpp:88-275).
__device__ void grid_interp(Ctx& c, const DevGrid& g, V3 pos, V3& f, double& e) {
    int gp[3];
    double pg[3];
    if (g.orthogonal) {
        gp[0] = (int)floor((pos.x - g.offset[0]) / g.basis[0]);
        gp[1] = (int)floor((pos.y - g.offset[1]) / g.basis[4]);
        gp[2] = (int)floor((pos.z - g.offset[2]) / g.basis[8]);
    } else {
        double qx, qy, qz;
        mat_vec(g.inv, pos.x - g.offset[0], pos.y - g.offset[1], pos.z - g.offset[2], qx, qy, qz);
        pg[0] = qx; pg[1] = qy; pg[2] = qz;
        gp[0] = (int)floor(qx); gp[1] = (int)floor(qy); gp[2] = (int)floor(qz);
    }
    if (!g.periodic) {
        for (int d = 0; d < 3; ++d) {
            if (gp[d] < 0) { gp[d] = 0; c.flags |= REC_GRID_CLAMP; }
            else if (gp[d] >= g.n[d]) { gp[d] = g.n[d] - 1; c.flags |= REC_GRID_CLAMP; }
        }
    }
    double fs[8][3], es[8];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int k = 0; k < 2; ++k) {
                int q[3] = {gp[0] + i, gp[1] + j, gp[2] + k};
                bool out = false;
                for (int d = 0; d < 3; ++d) out |= (q[d] < 0 || q[d] >= g.n[d]);
                if (out) {
                    for (int d = 0; d < 3; ++d) {
                        if (g.periodic) {
                            if (q[d] < 0 || q[d] >= g.n[d]) q[d] -= (int)floor((double)q[d] / g.n[d]) * g.n[d];
                        } else {
                            c.flags |= REC_OUT_OF_GRID;
                            q[d] = q[d] < 0 ? 0 : (q[d] >= g.n[d] ? g.n[d] - 1 : q[d]);
                        }
                    }
                }
                // final clamp: keeps every gather inside the grid even for
                // non-finite positions
                for (int d = 0; d < 3; ++d) q[d] = q[d] < 0 ? 0 : (q[d] >= g.n[d] ? g.n[d] - 1 : q[d]);
                long idx = ((long)q[0] * g.n[1] + q[1]) * g.n[2] + q[2];
                double4 v = g.fe64[idx];
                int s = i * 4 + j * 2 + k;
                fs[s][0] = v.x; fs[s][1] = v.y; fs[s][2] = v.z; es[s] = v.w;
            }
    double d[3];
    if (g.orthogonal) {
        d[0] = (pos.x - g.offset[0]) / g.basis[0] - gp[0];
        d[1] = (pos.y - g.offset[1]) / g.basis[4] - gp[1];
        d[2] = (pos.z - g.offset[2]) / g.basis[8] - gp[2];
    } else {
        d[0] = pg[0] - gp[0]; d[1] = pg[1] - gp[1]; d[2] = pg[2] - gp[2];
    }
    double o[3];
    for (int q = 0; q < 3; ++q) {
        double f00 = fs[0][q] * (1 - d[0]) + fs[4][q] * d[0];
        double f01 = fs[1][q] * (1 - d[0]) + fs[5][q] * d[0];
        double f10 = fs[2][q] * (1 - d[0]) + fs[6][q] * d[0];
        double f11 = fs[3][q] * (1 - d[0]) + fs[7][q] * d[0];
        double f0 = f00 * (1 - d[1]) + f10 * d[1];
        double f1 = f01 * (1 - d[1]) + f11 * d[1];
        o[q] = f0 * (1 - d[2]) + f1 * (d[2]);
    }
    f.x = o[0]; f.y = o[1]; f.z = o[2];
    double e00 = es[0] * (1 - d[0]) + es[4] * d[0];
    double e01 = es[1] * (1 - d[0]) + es[5] * d[0];
    double e10 = es[2] * (1 - d[0]) + es[6] * d[0];
    double e11 = es[3] * (1 - d[0]) + es[7] * d[0];
    double e0 = e00 * (1 - d[1]) + e10 * d[1];
    double e1 = e01 * (1 - d[1]) + e11 * d[1];
    e = e0 * (1 - d[2]) + e1 * (d[2]);
}

// The interaction list evaluated in order for the tip (index 1): tip-surface
// sites / grids, then tip-dummy VDW, Coulomb and spring unless only the
// tip-surface terms are requested (System::evalTipSurfaceForces).
__device__ void eval_tip(Ctx& c, V3 dummy, V3 tip, bool ts_only, V3& F, double& E) {
    const DevMol& m = *c.m;
    V3 f = {0, 0, 0};
    double e = 0;
    if (m.force_grid >= 0) {
        V3 gf; double ge;
        grid_interp(c, c.a->grids[m.force_grid], tip, gf, ge);
        f.x += gf.x; f.y += gf.y; f.z += gf.z;
        e += ge;
    } else {
        const Site64* s = c.a->sites_ref + m.ref_off;
        for (int im = 0; im < m.n_images; ++im)
        for (int n = 0; n < m.n_sites; ++n) {
            Site64 st = s[n];
            // positions[i2] + pbc_shift, subtracted from the tip (interactions.cpp:16)
            V3 r = {tip.x - (st.x + m.shift[im][0]), tip.y - (st.y + m.shift[im][1]),
                    tip.z - (st.z + m.shift[im][2])};
            V3 fv; double ev;
            vdw(st.vdw, st.p0, st.p1, st.p2, r, fv, ev);
            f.x += fv.x; f.y += fv.y; f.z += fv.z;
            e += ev;
            if (st.coul) {
                coulomb(st.qq, r, fv, ev);
                f.x += fv.x; f.y += fv.y; f.z += fv.z;
                e += ev;
            }
        }
        if (m.potential_grid >= 0) {
            V3 gf; double ge;
            grid_interp(c, c.a->grids[m.potential_grid], tip, gf, ge);
            f.x += gf.x; f.y += gf.y; f.z += gf.z;
            e += ge;
        }
    }
    if (!ts_only) {
        // VDW(0,1), Coulomb(0,1): r = dummy - tip, tip receives -f.
        V3 r = {dummy.x - tip.x, dummy.y - tip.y, dummy.z - tip.z};
        V3 fv; double ev;
        vdw(m.td_vdw, m.td_p0, m.td_p1, m.td_p2, r, fv, ev);
        f.x -= fv.x; f.y -= fv.y; f.z -= fv.z;
        e += ev;
        if (m.td_coul) {
            coulomb(m.td_qq, r, fv, ev);
            f.x -= fv.x; f.y -= fv.y; f.z -= fv.z;
            e += ev;
        }
        // TipHarmonicInteraction (interactions.cpp:252-266)
        double rl = sqrt(sq(r.x) + sq(r.y));
        double dr = rl - m.td_r0;
        ev = m.td_k * dr * dr;
        if (rl > 1e-10) {
            double s = -2 * m.td_k * dr / rl;
            f.x -= s * r.x; f.y -= s * r.y;
            f.z -= 0.0;
        }
        e += ev;
    }
    F = f;
    E = e;
}

struct State {
    V3 dummy, tip, v, f;
    double e;
};

__device__ inline bool converged(const ScanArgs& a, V3 f, double ediff) {
    switch (a.minterm) {
        case 0: return fabs(ediff) < a.etol;
        case 1: return len3(f) < a.ftol;
        default: return fabs(ediff) < a.etol && len3(f) < a.ftol;
    }
}

// eulerStep / midpointStep / rk4Step (integrators.cpp:8-111), masses = 1.
__device__ void step(Ctx& c, State& s, double dt) {
    const ScanArgs& a = *c.a;
    const double m = 1.0;
    if (a.integrator == 0) {
        V3 f; double e;
        eval_tip(c, s.dummy, s.tip, false, f, e);
        s.tip.x += dt * s.v.x; s.tip.y += dt * s.v.y; s.tip.z += dt * s.v.z;
        s.v.x += dt * f.x / m; s.v.y += dt * f.y / m; s.v.z += dt * f.z / m;
        s.f = f; s.e = e;
    } else if (a.integrator == 1) {
        V3 f1, f2; double e1, e2;
        eval_tip(c, s.dummy, s.tip, false, f1, e1);
        V3 p2 = s.tip, v2 = s.v;
        p2.x += dt / 2 * s.v.x; p2.y += dt / 2 * s.v.y; p2.z += dt / 2 * s.v.z;
        v2.x += dt / 2 * f1.x / m; v2.y += dt / 2 * f1.y / m; v2.z += dt / 2 * f1.z / m;
        eval_tip(c, s.dummy, p2, false, f2, e2);
        s.tip.x += dt * v2.x; s.tip.y += dt * v2.y; s.tip.z += dt * v2.z;
        s.v.x += dt * f2.x / m; s.v.y += dt * f2.y / m; s.v.z += dt * f2.z / m;
        s.f = f2; s.e = e2;
    } else {
        V3 f1, f2, f3, f4; double e1, e2, e3, e4;
        eval_tip(c, s.dummy, s.tip, false, f1, e1);
        V3 p2 = s.tip, v2 = s.v, p3 = s.tip, v3 = s.v, p4 = s.tip, v4 = s.v;
        p2.x += dt / 2 * s.v.x; p2.y += dt / 2 * s.v.y; p2.z += dt / 2 * s.v.z;
        v2.x += dt / 2 * f1.x / m; v2.y += dt / 2 * f1.y / m; v2.z += dt / 2 * f1.z / m;
        eval_tip(c, s.dummy, p2, false, f2, e2);
        p3.x += dt / 2 * v2.x; p3.y += dt / 2 * v2.y; p3.z += dt / 2 * v2.z;
        v3.x += dt / 2 * f2.x / m; v3.y += dt / 2 * f2.y / m; v3.z += dt / 2 * f2.z / m;
        eval_tip(c, s.dummy, p3, false, f3, e3);
        p4.x += dt * v3.x; p4.y += dt * v3.y; p4.z += dt * v3.z;
        v4.x += dt * f3.x / m; v4.y += dt * f3.y / m; v4.z += dt * f3.z / m;
        eval_tip(c, s.dummy, p4, false, f4, e4);
        s.tip.x += dt / 6 * (s.v.x + 2 * v2.x + 2 * v3.x + v4.x);
        s.tip.y += dt / 6 * (s.v.y + 2 * v2.y + 2 * v3.y + v4.y);
        s.tip.z += dt / 6 * (s.v.z + 2 * v2.z + 2 * v3.z + v4.z);
        s.v.x += dt / 6 * (f1.x + 2 * f2.x + 2 * f3.x + f4.x) / m;
        s.v.y += dt / 6 * (f1.y + 2 * f2.y + 2 * f3.y + f4.y) / m;
        s.v.z += dt / 6 * (f1.z + 2 * f2.z + 2 * f3.z + f4.z) / m;
        s.f.x = (f1.x + 2 * f2.x + 2 * f3.x + f4.x) / 6;
        s.f.y = (f1.y + 2 * f2.y + 2 * f3.y + f4.y) / 6;
        s.f.z = (f1.z + 2 * f2.z + 2 * f3.z + f4.z) / 6;
        s.e = (e1 + 2 * e2 + 2 * e3 + e4) / 6;
    }
}

// FIREMinimisation (minimiser.cpp:67-129).
__device__ int fire(Ctx& c, State& s) {
    const ScanArgs& a = *c.a;
    const int n_min = 5;
    const double f_inc = 1.1, f_dec = 0.5, f_a = 0.99, a_start = 0.1;
    const double dt_max = 10 * a.dt;
    double alpha = a_start, dt = a.dt;
    int n_non_neg = 0;
    int n_tot = 1;
    for (; n_tot < a.maxsteps; ++n_tot) {
        double prev_tip_e = s.e;
        step(c, s, dt);
        if (converged(a, s.f, s.e - prev_tip_e)) break;
        double p = 0;
        p += s.v.x * s.f.x + s.v.y * s.f.y + s.v.z * s.f.z;
        double l = len3(s.f);
        V3 fh = {0, 0, 0};
        if (l > 1e-10) { fh.x = s.f.x / l; fh.y = s.f.y / l; fh.z = s.f.z / l; }
        double vl = len3(s.v);
        s.v.x = (1 - alpha) * s.v.x + alpha * fh.x * vl;
        s.v.y = (1 - alpha) * s.v.y + alpha * fh.y * vl;
        s.v.z = (1 - alpha) * s.v.z + alpha * fh.z * vl;
        if (p < 0) {
            s.v.x = 0; s.v.y = 0; s.v.z = 0;
            n_non_neg = 0;
            dt *= f_dec;
            alpha = a_start;
        } else {
            if (n_non_neg > n_min) {
                dt = fmin(dt * f_inc, dt_max);
                alpha *= f_a;
            }
            n_non_neg++;
        }
    }
    return n_tot;
}

// SDMinimisation (minimiser.cpp:39-65).
__device__ int sd(Ctx& c, State& s) {
    const ScanArgs& a = *c.a;
    int n = 1;
    for (; n < a.maxsteps; ++n) {
        double prev_tip_e = s.e;
        eval_tip(c, s.dummy, s.tip, false, s.f, s.e);
        if (converged(a, s.f, s.e - prev_tip_e)) break;
        s.tip.x += a.dt * s.f.x; s.tip.y += a.dt * s.f.y; s.tip.z += a.dt * s.f.z;
    }
    return n;
}

__global__ void __launch_bounds__(256) k_scan_plain(ScanArgs a) {
    const long n_cols = (long)a.nx * a.ny;
    const long col = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= n_cols * a.n_mols) return;
    const int mi = (int)(col / n_cols);
    const long local = col - (long)mi * n_cols;
    const int i = (int)(local / a.ny);
    const int j = (int)(local - (long)i * a.ny);
    const DevMol& m = a.mols[mi];
    Ctx c = {&a, &m, 0};
    const double x = i * a.dx;
    const double y = j * a.dy;
    State s;
    s.v = {0, 0, 0};
    s.f = {0, 0, 0};
    s.e = 0;
    // System::setDummyXY (system.cpp:139-162) and setDummyZ (system.hpp:47-50)
    double px = x, py = y;
    if (m.tip_pbc) wrap_xy(m.cell, m.cell_inv, m.offset, x, y, px, py);
    s.dummy = {px, py, a.zhigh};
    s.tip = {px, py, a.zhigh - m.tip_dummy_d};
    int64_t steps = 0, unconv = 0;
    int warn = 0, oog = 0;
    for (int k = 0; k < a.nz; ++k) {
        c.flags = 0;
        int n = a.minimiser == 0 ? sd(c, s) : fire(c, s);
        V3 tf; double te;
        eval_tip(c, s.dummy, s.tip, true, tf, te);
        DevRecord r;
        r.x = x; r.y = y; r.z = s.dummy.z;
        r.fx = tf.x; r.fy = tf.y; r.fz = tf.z;
        r.rx = s.tip.x - s.dummy.x; r.ry = s.tip.y - s.dummy.y; r.rz = s.tip.z - s.dummy.z;
        r.e = te;
        r.n_steps = n;
        r.flags = c.flags | (n >= a.maxsteps ? REC_UNCONVERGED : 0);
        a.rec[m.rec_off + (local * a.nz + k)] = r;
        if (a.col_done) column_progress(a, col, k + 1);
        steps += n;
        unconv += n >= a.maxsteps;
        warn += (c.flags & REC_GRID_CLAMP) != 0;
        oog += (c.flags & REC_OUT_OF_GRID) != 0;
        s.dummy.z -= a.dz;   // System::lowerTip (system.hpp:51-54)
        s.tip.z -= a.dz;
    }
    atomicAdd((unsigned long long*)&a.counters[0], (unsigned long long)steps);
    if (unconv) atomicAdd((unsigned long long*)&a.counters[1], (unsigned long long)unconv);
    if (warn) atomicAdd((unsigned long long*)&a.counters[2], (unsigned long long)warn);
    if (oog) atomicAdd((unsigned long long*)&a.counters[3], (unsigned long long)oog);
}

// System::evalTipSurfaceForces at arbitrary tip positions.
__global__ void __launch_bounds__(256) k_eval_plain(ScanArgs a, int mol, long n, const double* pos,
                                                     double* out) {
    long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    Ctx c = {&a, &a.mols[mol], 0};
    V3 t = {pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]};
    V3 f; double e;
    eval_tip(c, t, t, true, f, e);
    out[4 * p] = f.x; out[4 * p + 1] = f.y; out[4 * p + 2] = f.z; out[4 * p + 3] = e;
    if (c.flags & REC_OUT_OF_GRID) atomicAdd((unsigned long long*)&a.counters[3], 1ull);
}

// ForceGrid::interpolate (force_grid.cpp:88-155) of grid `grid` at
// arbitrary positions; out-of-grid lookups counted in counters[3].
__global__ void __launch_bounds__(256) k_grid_interp(ScanArgs a, int grid, long n, const double* pos, double* out) {
    const long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    Ctx c = {&a, nullptr, 0};
    const V3 t = {pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]};
    V3 f;
    double e;
    grid_interp(c, a.grids[grid], t, f, e);
    out[4 * p] = f.x; out[4 * p + 1] = f.y; out[4 * p + 2] = f.z; out[4 * p + 3] = e;
    if (c.flags & REC_OUT_OF_GRID) atomicAdd((unsigned long long*)&a.counters[3], 1ull);
}

// Grid sampling of buildTipGridInteractions (simulation.cpp:450-479): the tip
// sits on the dummy, every tip-surface term of molecule `mol` is evaluated at
// (i*s + o) (xy wrapped by setDummyXY when the system is periodic).  One lane
// per grid point, z fastest, so consecutive lanes write consecutive points.
__global__ void __launch_bounds__(256) k_grid_build(ScanArgs a, int mol, int nx, int ny, int nz,
                                                     double sx, double sy, double sz, double ox,
                                                     double oy, double oz, double4* out) {
    long p = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long total = (long)nx * ny * nz;
    if (p >= total) return;
    int k = (int)(p % nz);
    long ij = p / nz;
    int j = (int)(ij % ny);
    int i = (int)(ij / ny);
    const DevMol& m = a.mols[mol];
    Ctx c = {&a, &m, 0};
    double x = i * sx + ox;
    double y = j * sy + oy;
    double px = x, py = y;
    if (m.tip_pbc) wrap_xy(m.cell, m.cell_inv, m.offset, x, y, px, py);
    V3 t = {px, py, k * sz + oz};
    V3 f; double e;
    eval_tip(c, t, t, true, f, e);
    out[p] = make_double4(f.x, f.y, f.z, e);
}

}  // namespace

hipError_t launch_scan_plain(const ScanArgs& a, hipStream_t st) {
    long cols = (long)a.nx * a.ny * a.n_mols;
    int blocks = (int)((cols + 255) / 256);
    hipLaunchKernelGGL(k_scan_plain, dim3(blocks), dim3(256), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_eval_plain(const ScanArgs& a, int mol, long n, const double* pos, double* out,
                             hipStream_t st) {
    int blocks = (int)((n + 255) / 256);
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(k_eval_plain, dim3(blocks), dim3(256), 0, st, a, mol, n, pos, out);
    return hipGetLastError();
}

hipError_t launch_grid_interp(const ScanArgs& a, int grid, long n, const double* pos, double* out, hipStream_t st) {
    const int blocks = (int)((n + 255) / 256);
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(k_grid_interp, dim3(blocks), dim3(256), 0, st, a, grid, n, pos, out);
    return hipGetLastError();
}

hipError_t launch_grid_build(const ScanArgs& a, int mol, const int* n, const double* s, const double* o,
                             double4* out, hipStream_t st) {
    long total = (long)n[0] * n[1] * n[2];
    int blocks = (int)((total + 255) / 256);
    if (blocks == 0) return hipSuccess;
    hipLaunchKernelGGL(k_grid_build, dim3(blocks), dim3(256), 0, st, a, mol, n[0], n[1], n[2], s[0], s[1],
                       s[2], o[0], o[1], o[2], out);
    return hipGetLastError();
}

}  // namespace mafm
