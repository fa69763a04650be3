// Flexible scan (flexible on): every free surface atom relaxes with the tip.
//
// One workgroup of 1-4 wavefronts relaxes one pixel column at a time
// (persistent grid, columns handed out by an atomic counter).  The column's
// whole state lives with the workgroup -- in LDS when it fits (gfx950: up to
// 160 KiB), else in a per-block slice of HBM: positions, the integrator's
// evaluation point and stage velocities, velocities, forces, the term
// results and the index tables.
//
// One integrator stage evaluates the interaction list in two phases:
//   1. term-parallel: each thread evaluates the terms the host's schedule
//      gives it (64 terms of one kind per wave-round, so a wave runs one
//      code path at a time, the rounds spread over the waves by cost) and
//      writes one result slot {fx, fy, fz} per atom the term acts on, the
//      force that atom receives (interactions.cpp:15-51, 252-362;
//      `forces[i] -= f` is stored as -f, which adds with the same rounding),
//      and the term's energy once (only the tip's energy is ever read);
//   2. atom-parallel: thread l sums the slots of its atoms in the
//      interaction list's order, from zero -- the order in which the
//      reference's Interaction::eval calls accumulate forces_[i] and
//      energies_[i] (integrators.cpp:8-111) -- so each force is the
//      reference's to the last bit given the same term arithmetic.  The tip,
//      whose list holds every tip-surface term, is summed by the block's last
//      thread, whose wave has no other atom when the block has several waves.
// FIRE's power P = sum_i v_i . f_i (minimiser.cpp:107-109) is summed in atom
// order by one thread.  Fixed atoms (fixed = 1) never move and have zero
// velocity, so their forces (which the reference computes and multiplies by
// that zero velocity in P) are not summed.  Built with -ffp-contract=off;
// IEEE divisions and square roots.  The angle, dihedral and Morse terms call
// the device acos / sin / atan / exp, which may differ from glibc's in the
// last place.
#include <algorithm>

#include "flex.hpp"
#include "ref_terms.hpp"

namespace mafm {
namespace {

constexpr double TOL = 1e-10;   // TOLERANCE, globals.hpp:11

__device__ inline V3 ld3(const double* a, int i) { return {a[3 * i], a[3 * i + 1], a[3 * i + 2]}; }
__device__ inline void st3(double* a, int i, V3 v) {
    a[3 * i] = v.x;
    a[3 * i + 1] = v.y;
    a[3 * i + 2] = v.z;
}
__device__ inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ inline V3 neg(V3 a) { return {-a.x, -a.y, -a.z}; }
__device__ inline V3 scl(double s, V3 v) { return {s * v.x, s * v.y, s * v.z}; }   // T * Vector3
__device__ inline V3 scr(V3 v, double s) { return {v.x * s, v.y * s, v.z * s}; }   // Vector3 * T
__device__ inline V3 dvs(V3 v, double s) { return {v.x / s, v.y / s, v.z / s}; }   // Vector3 / T
__device__ inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }   // vectors.hpp:116-118
__device__ inline V3 cross(V3 a, V3 b) {                                                   // vectors.hpp:120-122
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

__device__ inline void put(double* R, int slot, V3 f) {
    R[3 * slot] = f.x;
    R[3 * slot + 1] = f.y;
    R[3 * slot + 2] = f.z;
}

// One Interaction::eval of the list at positions X: the force on each atom
// into the term's slots R (slot r = the term's atom[r]), the energy into e.
__device__ void eval_term(const mafm_flex_term& t, const double* X, double* R, double& e_out) {
    switch (t.kind) {
        case MAFM_FX_LJ:
        case MAFM_FX_MORSE: {   // interactions.cpp:15-40
            const V3 p2 = ld3(X, t.atom[1]);
            const V3 r = sub(ld3(X, t.atom[0]), add(p2, {t.shift[0], t.shift[1], t.shift[2]}));
            V3 f;
            vdw(t.kind == MAFM_FX_LJ ? 0 : 1, t.p[0], t.p[1], t.p[2], r, f, e_out);
            put(R, 0, f);
            put(R, 1, neg(f));
            return;
        }
        case MAFM_FX_COULOMB: {   // interactions.cpp:42-51
            V3 f;
            coulomb(t.p[0], sub(ld3(X, t.atom[0]), ld3(X, t.atom[1])), f, e_out);
            put(R, 0, f);
            put(R, 1, neg(f));
            return;
        }
        case MAFM_FX_TIP_HARMONIC: {   // interactions.cpp:252-266
            const V3 r = sub(ld3(X, t.atom[0]), ld3(X, t.atom[1]));
            const double rl = sqrt(sq(r.x) + sq(r.y));
            const double dr = rl - t.p[1];
            const double e = t.p[0] * dr * dr;
            V3 f = {0, 0, 0};
            if (rl > TOL) {
                const double s = -2 * t.p[0] * dr / rl;
                f = {s * r.x, s * r.y, 0};
            }
            e_out = e;
            put(R, 0, f);
            put(R, 1, neg(f));
            return;
        }
        case MAFM_FX_BOND: {   // HarmonicInteraction::eval, interactions.cpp:296-306
            const V3 r = sub(ld3(X, t.atom[0]), ld3(X, t.atom[1]));
            const double rl = len3(r);
            const double dr = rl - t.p[1];
            const double e = t.p[0] * sq(dr);
            const V3 f = scl(-2 * t.p[0] * dr / rl, r);
            e_out = e;
            put(R, 0, f);
            put(R, 1, neg(f));
            return;
        }
        case MAFM_FX_ANGLE: {   // HarmonicAngleInteraction::eval, interactions.cpp:308-332
            const V3 ps = ld3(X, t.atom[0]);
            const V3 r1v = sub(ld3(X, t.atom[1]), ps), r2v = sub(ld3(X, t.atom[2]), ps);
            const double r1 = len3(r1v), r2 = len3(r2v);
            double cos_t = dot(r1v, r2v) / (r1 * r2);
            if (cos_t > 1) cos_t = 1;
            else if (cos_t < -1) cos_t = -1;
            const double theta = acos(cos_t);
            const double sin_t = sin(theta);
            const double d_theta = theta - t.p[1];
            const double fm = -2 * t.p[0] * d_theta / sin_t;
            const V3 f1 = scl(-fm / r1, sub(dvs(r2v, r2), dvs(scr(r1v, cos_t), r1)));
            const V3 f2 = scl(-fm / r2, sub(dvs(r1v, r1), dvs(scr(r2v, cos_t), r2)));
            e_out = t.p[0] * sq(d_theta);
            put(R, 0, neg(add(f1, f2)));   // forces[shared] -= f1 + f2
            put(R, 1, f1);
            put(R, 2, f2);
            return;
        }
        case MAFM_FX_DIHEDRAL: {   // HarmonicDihedralInteraction::eval, interactions.cpp:334-362
            const V3 p2 = ld3(X, t.atom[1]), p3 = ld3(X, t.atom[2]);
            const V3 r12 = sub(ld3(X, t.atom[0]), p2), r23 = sub(p2, p3), r43 = sub(ld3(X, t.atom[3]), p3);
            const double r23l = len3(r23);
            const V3 mv = cross(r12, r23), nv = cross(r43, r23);
            const double m = len3(mv), n = len3(nv);
            const double tan_s = dot(nv, r12) * r23l / dot(mv, nv);
            const double sigma = atan(tan_s);
            const double d_sigma = sigma - t.p[1];
            const double fm = 2 * t.p[0] * d_sigma;
            e_out = t.p[0] * sq(d_sigma);
            put(R, 0, scl(-fm * r23l / (m * m), mv));
            put(R, 1, scl(-fm, sub(scl(dot(r43, r23) / (n * n * r23l), nv),
                                   scl(((r23l * r23l) + dot(r12, r23)) / (m * m * r23l), mv))));
            put(R, 2, scl(-fm, add(scl(dot(r12, r23) / (m * m * r23l), mv),
                                   scl(((r23l * r23l) - dot(r43, r23)) / (n * n * r23l), nv))));
            put(R, 3, scl(fm * r23l / (n * n), nv));
            return;
        }
        case MAFM_FX_SUBSTRATE: {   // SubstrateInteraction::eval, interactions.cpp:280-294
            // p: multiplier, ulj, ushift, z0, rc, sigma
            const double dz = X[3 * t.atom[0] + 2] - t.p[3];
            const double sig_z = t.p[5] / dz;
            const double sig_z2 = sig_z * sig_z;
            const double sig_z4 = sig_z2 * sig_z2;
            const double sig_z10 = sig_z4 * sig_z4 * sig_z2;
            double eterm = 0.0, fterm = 0.0;
            if (dz <= t.p[4]) {
                eterm += t.p[0] * (2. / 5 * sig_z10 - sig_z4) + t.p[1] * dz * dz - t.p[2];
                fterm += 4 * t.p[0] / dz * (sig_z10 - sig_z4) + 2 * t.p[1] * dz;
            }
            // forces[i].z += fterm: the +0 added to x and y leaves them as they are
            put(R, 0, {0, 0, fterm});
            e_out = eterm;
            return;
        }
        default: {   // MAFM_FX_XY_HARMONIC, interactions.cpp:268-278: p = k, x0, y0
            const double rx = X[3 * t.atom[0]] - t.p[1], ry = X[3 * t.atom[0] + 1] - t.p[2];
            const double rl = sqrt(sq(rx) + sq(ry));
            const double e = t.p[0] * rl * rl;
            V3 f = {0, 0, 0};
            if (rl > TOL) f = {-2 * t.p[0] * rx, -2 * t.p[0] * ry, 0};
            e_out = e;
            put(R, 0, f);
            return;
        }
    }
}

// checkConvergence (minimiser.cpp:15-37)
__device__ inline bool converged(const FlexArgs& a, V3 f, double ediff) {
    switch (a.minterm) {
        case MAFM_TERM_E: return fabs(ediff) < a.etol;
        case MAFM_TERM_F: return len3(f) < a.ftol;
        default: return fabs(ediff) < a.etol && len3(f) < a.ftol;
    }
}

// checkConvergence (minimiser.cpp:15-37) with a System operator's options
template <typename A>
__device__ inline bool converged_sys(const A& a, V3 f, double ediff) {
    switch (a.minterm) {
        case MAFM_TERM_E: return fabs(ediff) < a.etol;
        case MAFM_TERM_F: return len3(f) < a.ftol;
        default: return fabs(ediff) < a.etol && len3(f) < a.ftol;
    }
}

__host__ __device__ inline size_t block_doubles(int n, int n_slots, int n_eval, bool rk4) {
    return (size_t)(16 + (rk4 ? 6 : 0)) * n + 3 * (size_t)n_slots + n_eval + 8;
}

struct State {
    double *pos, *pev, *vel, *vst, *frc, *pdot, *res, *ener, *bc, *facc, *vacc;
    const int32_t *csr_off, *csr, *tip_eval, *fixed;   // the index tables (LDS copy or global)
};

__device__ inline State layout(double* p, int n, int n_slots, int n_eval, bool rk4) {
    State S;
    S.pos = p; p += 3 * n;
    S.pev = p; p += 3 * n;
    S.vel = p; p += 3 * n;
    S.vst = p; p += 3 * n;
    S.frc = p; p += 3 * n;
    S.pdot = p; p += n;
    S.res = p; p += 3 * n_slots;
    S.ener = p; p += n_eval;
    S.bc = p; p += 8;
    S.facc = p; p += rk4 ? 3 * n : 0;
    S.vacc = p;
    return S;
}

__device__ inline void block_sync() { __syncthreads(); }

// acc += the slots res[idx[q]] for q in [q0, q1), in order; the loads of
// four slots are issued before their adds
__device__ inline V3 sum_slots(const double* res, const int32_t* idx, int q0, int q1) {
    V3 acc = {0, 0, 0};
    int q = q0;
    for (; q + 4 <= q1; q += 4) {
        const int s0 = idx[q], s1 = idx[q + 1], s2 = idx[q + 2], s3 = idx[q + 3];
        const V3 r0 = ld3(res, s0), r1 = ld3(res, s1), r2 = ld3(res, s2), r3 = ld3(res, s3);
        acc = add(acc, r0);
        acc = add(acc, r1);
        acc = add(acc, r2);
        acc = add(acc, r3);
    }
    for (; q < q1; ++q) acc = add(acc, ld3(res, idx[q]));
    return acc;
}

// The long ordered sums -- the tip's list, which holds every tip-surface
// term, and FIRE's power over all atoms -- by one thread: eight values are
// loaded before their eight additions, so the chain waits on the adds, not
// on LDS.
__device__ inline V3 seq_sum_slots(const double* res, const int32_t* idx, int q0, int q1) {
    V3 acc = {0, 0, 0};
    int q = q0;
    for (; q + 8 <= q1; q += 8) {
        V3 r[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) r[j] = ld3(res, idx[q + j]);
#pragma unroll
        for (int j = 0; j < 8; ++j) acc = add(acc, r[j]);
    }
    for (; q < q1; ++q) acc = add(acc, ld3(res, idx[q]));
    return acc;
}

// sum of v[idx[q]] (idx NULL: v[q]) for q in [0, n), in order
__device__ inline double seq_sum(const double* v, const int32_t* idx, int n) {
    double acc = 0;
    int q = 0;
    for (; q + 8 <= n; q += 8) {
        double r[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) r[j] = v[idx ? idx[q + j] : q + j];
#pragma unroll
        for (int j = 0; j < 8; ++j) acc += r[j];
    }
    for (; q < n; ++q) acc += v[idx ? idx[q] : q];
    return acc;
}

// Forces on every free atom at the positions X (into S.frc) and the tip's
// energy (into S.bc[0]): the interaction list evaluated as one integrator
// stage does (integrators.cpp:31-33).  ts_only: only the tip-surface terms,
// for the tip, as System::evalTipSurfaceForces does (system.cpp:41-51).
// t / nthr: this thread and the block's threads; the terms follow the
// schedule a.sched when there is one.  The block's last thread sums the tip.
__device__ void evaluate(const FlexArgs& a, const State& S, const double* X, int t, int nthr, bool ts_only) {
    if (ts_only || !a.sched) {
        const int ne = ts_only ? a.n_ts : a.n_eval;
        for (int k = t; k < ne; k += nthr) {
            const mafm_flex_term& term = a.terms[k];
            eval_term(term, X, S.res + 3 * term.reserved, S.ener[k]);   // reserved: the term's first slot
        }
    } else {
        // the next round's term is loaded while this round's is evaluated
        int k = a.sched[t];
        for (int r = 0; r < a.sched_rounds; ++r) {
            const int next = r + 1 < a.sched_rounds ? a.sched[(r + 1) * nthr + t] : -1;
            if (k >= 0) {
                const mafm_flex_term& term = a.terms[k];
                eval_term(term, X, S.res + 3 * term.reserved, S.ener[k]);
            }
            k = next;
        }
    }
    block_sync();
    if (!ts_only)
        for (int i = t; i < a.n_atoms; i += nthr) {
            if (i == 1) continue;
            if (S.fixed[i] == 1) st3(S.frc, i, {0, 0, 0});
            else st3(S.frc, i, sum_slots(S.res, S.csr, S.csr_off[i], S.csr_off[i + 1]));
        }
    if (t == nthr - 1) {   // the tip: its slots (the tip-surface ones first) and its terms' energies
        const int q0 = S.csr_off[1];
        st3(S.frc, 1, seq_sum_slots(S.res, S.csr, q0, ts_only ? q0 + a.tip_ts_slots : S.csr_off[2]));
        // the tip's energy is read by the output and by the energy convergence test only
        if (ts_only || a.minterm != MAFM_TERM_F) S.bc[0] = seq_sum(S.ener, S.tip_eval, ts_only ? a.tip_ts_terms : a.n_tip);
    }
    block_sync();
}

template <bool IN_LDS>
__global__ void __launch_bounds__(256) k_scan_flex(FlexArgs a) {
    extern __shared__ double lds_flex[];
    __shared__ int col_shared;
    const int lane = threadIdx.x;   // thread of the block (1-4 waves share a column)
    const int nthr = blockDim.x;
    const int n = a.n_atoms;
    const bool rk4 = a.minimiser == MAFM_MIN_FIRE && a.integrator == MAFM_INT_RK4;
    const size_t nd = block_doubles(n, a.n_slots, a.n_eval, rk4);
    double* base = IN_LDS ? lds_flex : a.work + (size_t)blockIdx.x * nd;
    State S = layout(base, n, a.n_slots, a.n_eval, rk4);
    // the index tables: csr_off (n + 1), csr, tip_eval, fixed (n); copied to LDS
    // behind the state when it lives there
    const int n_idx = (n + 1) + a.n_csr + a.n_tip + n;
    const int32_t* idx = a.idx;
    if (IN_LDS) {
        int32_t* l = reinterpret_cast<int32_t*>(lds_flex + nd);
        for (int q = lane; q < n_idx; q += nthr) l[q] = a.idx[q];
        idx = l;
    }
    S.csr_off = idx;
    S.csr = idx + (n + 1);
    S.tip_eval = S.csr + a.n_csr;
    S.fixed = S.tip_eval + a.n_tip;
    block_sync();
    const int n_cols = a.nx * a.ny;
    for (;;) {
        if (lane == 0) col_shared = atomicAdd(a.next_col, 1);
        block_sync();
        const int col = col_shared;
        block_sync();
        if (col >= n_cols) break;
        const int ci = col / a.ny, cj = col - ci * a.ny;
        const double x = ci * a.dx, y = cj * a.dy;   // simulation.cpp:80-83
        // System copy (simulation.cpp:93): initial positions, zero velocities
        for (int i = lane; i < n; i += nthr) {
            st3(S.pos, i, ld3(a.pos0, i));
            st3(S.vel, i, {0, 0, 0});
        }
        block_sync();
        if (lane == 0) {
            // setDummyXY / setDummyZ (system.cpp:139-162, system.hpp:47-50)
            double px = x, py = y;
            if (a.tip_pbc) wrap_xy(a.cell, a.cell_inv, a.offset, x, y, px, py);
            st3(S.pos, 0, {px, py, a.zhigh});
            st3(S.pos, 1, {px, py, a.zhigh - a.tip_dummy_d});
        }
        block_sync();
        double e_tip = 0;   // energies_[1] of the System copy: 0 until its first step
        for (int k = 0; k < a.nz; ++k) {
            // FIREMinimisation / SDMinimisation (minimiser.cpp:39-129)
            double alpha = 0.1, dt = a.dt;
            int nnn = 0, n_tot = 1;
            for (; n_tot < a.maxsteps; ++n_tot) {
                const double prev_e = e_tip;
                if (a.minimiser == MAFM_MIN_SD) {   // evaluate, test, move
                    evaluate(a, S, S.pos, lane, nthr, false);
                    e_tip = S.bc[0];
                    if (converged(a, ld3(S.frc, 1), e_tip - prev_e)) break;
                    for (int i = lane; i < n; i += nthr)
                        if (S.fixed[i] != 1) st3(S.pos, i, add(ld3(S.pos, i), scl(a.dt, ld3(S.frc, i))));
                    block_sync();
                    continue;
                }
                if (a.integrator == MAFM_INT_EULER) {   // eulerStep
                    evaluate(a, S, S.pos, lane, nthr, false);
                    for (int i = lane; i < n; i += nthr)
                        if (S.fixed[i] != 1) {
                            st3(S.pos, i, add(ld3(S.pos, i), scl(dt, ld3(S.vel, i))));
                            st3(S.vel, i, add(ld3(S.vel, i), scl(dt, ld3(S.frc, i))));   // masses are 1
                        }
                    e_tip = S.bc[0];
                } else if (a.integrator == MAFM_INT_MIDPOINT) {   // midpointStep
                    evaluate(a, S, S.pos, lane, nthr, false);
                    for (int i = lane; i < n; i += nthr) {
                        if (S.fixed[i] != 1) {
                            st3(S.pev, i, add(ld3(S.pos, i), scl(dt / 2, ld3(S.vel, i))));
                            st3(S.vst, i, add(ld3(S.vel, i), scl(dt / 2, ld3(S.frc, i))));
                        } else {
                            st3(S.pev, i, ld3(S.pos, i));
                        }
                    }
                    block_sync();
                    evaluate(a, S, S.pev, lane, nthr, false);
                    for (int i = lane; i < n; i += nthr)
                        if (S.fixed[i] != 1) {
                            st3(S.pos, i, add(ld3(S.pos, i), scl(dt, ld3(S.vst, i))));
                            st3(S.vel, i, add(ld3(S.vel, i), scl(dt, ld3(S.frc, i))));
                        }
                    e_tip = S.bc[0];
                } else {   // rk4Step: stages at pos, pos + dt/2 v1, pos + dt/2 v2, pos + dt v3
                    double e_acc = 0;
                    for (int s = 0; s < 4; ++s) {
                        evaluate(a, S, s == 0 ? S.pos : S.pev, lane, nthr, false);
                        const double es = S.bc[0];
                        e_acc = s == 0 ? es : e_acc + (s == 3 ? es : 2 * es);
                        for (int i = lane; i < n; i += nthr) {
                            const V3 fs = ld3(S.frc, i), vs = s == 0 ? ld3(S.vel, i) : ld3(S.vst, i);
                            if (s == 0) {
                                st3(S.facc, i, fs);
                                st3(S.vacc, i, vs);
                            } else if (s < 3) {
                                st3(S.facc, i, add(ld3(S.facc, i), scl(2, fs)));
                                st3(S.vacc, i, add(ld3(S.vacc, i), scl(2, vs)));
                            } else {
                                st3(S.facc, i, add(ld3(S.facc, i), fs));
                                st3(S.vacc, i, add(ld3(S.vacc, i), vs));
                            }
                            if (s < 3) {
                                const double h = s < 2 ? dt / 2 : dt;
                                if (S.fixed[i] != 1) {
                                    st3(S.pev, i, add(ld3(S.pos, i), scl(h, vs)));
                                    st3(S.vst, i, add(ld3(S.vel, i), scl(h, fs)));
                                } else {
                                    st3(S.pev, i, ld3(S.pos, i));
                                    st3(S.vst, i, ld3(S.vel, i));
                                }
                            }
                        }
                        block_sync();
                    }
                    for (int i = lane; i < n; i += nthr) {
                        if (S.fixed[i] != 1) {
                            st3(S.pos, i, add(ld3(S.pos, i), scl(dt / 6, ld3(S.vacc, i))));
                            st3(S.vel, i, add(ld3(S.vel, i), scl(dt / 6, ld3(S.facc, i))));
                        }
                        st3(S.frc, i, dvs(ld3(S.facc, i), 6));
                    }
                    e_tip = e_acc / 6;
                }
                block_sync();
                if (converged(a, ld3(S.frc, 1), e_tip - prev_e)) break;
                // FIRE (minimiser.cpp:107-126): P over every atom in index order
                for (int i = lane; i < n; i += nthr) S.pdot[i] = dot(ld3(S.vel, i), ld3(S.frc, i));
                block_sync();
                if (lane == nthr - 1) S.bc[1] = seq_sum(S.pdot, nullptr, n);
                block_sync();
                const double p = S.bc[1];
                for (int i = lane; i < n; i += nthr) {
                    if (S.fixed[i] == 1) continue;
                    const V3 v = ld3(S.vel, i), f = ld3(S.frc, i);
                    const double fl = len3(f);
                    const V3 fh = fl > 1e-10 ? V3{f.x / fl, f.y / fl, f.z / fl} : V3{0, 0, 0};   // normalized()
                    st3(S.vel, i, add(scl(1 - alpha, v), scr(scl(alpha, fh), len3(v))));
                }
                block_sync();
                if (p < 0) {
                    for (int i = lane; i < n; i += nthr) st3(S.vel, i, {0, 0, 0});
                    block_sync();
                    nnn = 0;
                    dt *= 0.5;
                    alpha = 0.1;
                } else {
                    if (nnn > 5) {
                        dt = fmin(dt * 1.1, 10 * a.dt);
                        alpha *= 0.99;
                    }
                    ++nnn;
                }
            }
            // makeXYZFile of the state column (simulation.cpp:114-116), then
            // getOutput (system.cpp:28-51)
            if (col == a.state_col && a.state)
                for (int i = lane; i < n; i += nthr) st3(a.state + (size_t)k * 3 * n, i, ld3(S.pos, i));
            evaluate(a, S, S.pos, lane, nthr, true);
            if (lane == 0) {
                const V3 d = ld3(S.pos, 0), t = ld3(S.pos, 1), f = ld3(S.frc, 1);
                DevRecord rec;
                rec.x = x; rec.y = y; rec.z = d.z;
                rec.fx = f.x; rec.fy = f.y; rec.fz = f.z;
                rec.rx = t.x - d.x; rec.ry = t.y - d.y; rec.rz = t.z - d.z;
                rec.e = S.bc[0];
                rec.n_steps = n_tot;
                rec.flags = n_tot >= a.maxsteps ? REC_UNCONVERGED : 0;
                a.rec[(size_t)col * a.nz + k] = rec;
                // lowerTip (system.hpp:51-54)
                S.pos[2] -= a.dz;
                S.pos[5] -= a.dz;
            }
            block_sync();
        }
    }
}

// ---------------------------------------------------------------- System operators
// State of one system copy: positions, velocities, forces and energies of
// every atom, the integrator's evaluation point, stage velocities and
// accumulators, FIRE's per-atom powers, and the term results.
__host__ __device__ inline size_t sys_doubles(int n, int n_slots, int n_terms) {
    return (size_t)24 * n + 3 * (size_t)n_slots + n_terms + 8;
}

struct SysState {
    double *pos, *vel, *frc, *ener, *pev, *vst, *facc, *vacc, *eacc, *pdot, *res, *tener, *bc;
    const int32_t *f_off, *f_csr, *e_off, *e_csr, *fixed;
};

__device__ inline SysState sys_layout(double* p, int n, int n_slots, int n_terms) {
    SysState S;
    S.pos = p; p += 3 * n;
    S.vel = p; p += 3 * n;
    S.frc = p; p += 3 * n;
    S.ener = p; p += n;
    S.pev = p; p += 3 * n;
    S.vst = p; p += 3 * n;
    S.facc = p; p += 3 * n;
    S.vacc = p; p += 3 * n;
    S.eacc = p; p += n;
    S.pdot = p; p += n;
    S.res = p; p += 3 * n_slots;
    S.tener = p; p += n_terms;
    S.bc = p;
    return S;
}

// Every term's Interaction::eval at positions X, added in list order onto
// F / E (interactions.cpp:15-51, 252-362): F[i] += each result slot of atom
// i, E[i] += the energy of each term acting on i.
__device__ void eval_list(const SysArgs& a, const SysState& S, const double* X, double* F, double* E, int t,
                          int nthr) {
    for (int k = t; k < a.n_terms; k += nthr) {
        const mafm_flex_term& term = a.terms[k];
        eval_term(term, X, S.res + 3 * term.reserved, S.tener[k]);
    }
    block_sync();
    for (int i = t; i < a.n_atoms; i += nthr) {
        V3 f = ld3(F, i);
        for (int q = S.f_off[i]; q < S.f_off[i + 1]; ++q) f = add(f, ld3(S.res, S.f_csr[q]));
        st3(F, i, f);
        double e = E[i];
        for (int q = S.e_off[i]; q < S.e_off[i + 1]; ++q) e += S.tener[S.e_csr[q]];
        E[i] = e;
    }
    block_sync();
}

__device__ inline void zero_fe(const SysArgs& a, double* F, double* E, int t, int nthr) {
    for (int i = t; i < a.n_atoms; i += nthr) {
        st3(F, i, {0, 0, 0});
        E[i] = 0;
    }
}

// eulerStep / midpointStep / rk4Step (integrators.cpp:8-111), masses 1.
// Euler adds the list onto S.frc / S.ener as they are; the other two leave
// their final-stage (midpoint) or averaged (RK4) forces and energies there.
__device__ void sys_step(const SysArgs& a, const SysState& S, int integrator, double dt, int t, int nthr) {
    const int n = a.n_atoms;
    if (integrator == MAFM_INT_EULER) {
        eval_list(a, S, S.pos, S.frc, S.ener, t, nthr);
        for (int i = t; i < n; i += nthr)
            if (S.fixed[i] != 1) {
                st3(S.pos, i, add(ld3(S.pos, i), scl(dt, ld3(S.vel, i))));
                st3(S.vel, i, add(ld3(S.vel, i), scl(dt, ld3(S.frc, i))));
            }
        block_sync();
        return;
    }
    if (integrator == MAFM_INT_MIDPOINT) {
        zero_fe(a, S.facc, S.eacc, t, nthr);   // f1, e1
        block_sync();
        eval_list(a, S, S.pos, S.facc, S.eacc, t, nthr);
        for (int i = t; i < n; i += nthr) {
            const bool free_ = S.fixed[i] != 1;
            st3(S.pev, i, free_ ? add(ld3(S.pos, i), scl(dt / 2, ld3(S.vel, i))) : ld3(S.pos, i));
            st3(S.vst, i, free_ ? add(ld3(S.vel, i), scl(dt / 2, ld3(S.facc, i))) : ld3(S.vel, i));
        }
        zero_fe(a, S.frc, S.ener, t, nthr);    // f2, e2 become the system's
        block_sync();
        eval_list(a, S, S.pev, S.frc, S.ener, t, nthr);
        for (int i = t; i < n; i += nthr)
            if (S.fixed[i] != 1) {
                st3(S.pos, i, add(ld3(S.pos, i), scl(dt, ld3(S.vst, i))));
                st3(S.vel, i, add(ld3(S.vel, i), scl(dt, ld3(S.frc, i))));
            }
        block_sync();
        return;
    }
    // RK4: stages at pos, pos + dt/2 v1, pos + dt/2 v2, pos + dt v3; the sums
    // f1 + 2 f2 + 2 f3 + f4 (and v, e alike) accumulated left to right
    for (int s = 0; s < 4; ++s) {
        zero_fe(a, S.frc, S.ener, t, nthr);
        block_sync();
        eval_list(a, S, s == 0 ? S.pos : S.pev, S.frc, S.ener, t, nthr);
        for (int i = t; i < n; i += nthr) {
            const V3 fs = ld3(S.frc, i), vs = s == 0 ? ld3(S.vel, i) : ld3(S.vst, i);
            const double es = S.ener[i];
            if (s == 0) {
                st3(S.facc, i, fs);
                st3(S.vacc, i, vs);
                S.eacc[i] = es;
            } else if (s < 3) {
                st3(S.facc, i, add(ld3(S.facc, i), scl(2, fs)));
                st3(S.vacc, i, add(ld3(S.vacc, i), scl(2, vs)));
                S.eacc[i] = S.eacc[i] + 2 * es;
            } else {
                st3(S.facc, i, add(ld3(S.facc, i), fs));
                st3(S.vacc, i, add(ld3(S.vacc, i), vs));
                S.eacc[i] = S.eacc[i] + es;
            }
            if (s < 3) {
                const double h = s < 2 ? dt / 2 : dt;
                if (S.fixed[i] != 1) {
                    st3(S.pev, i, add(ld3(S.pos, i), scl(h, vs)));
                    st3(S.vst, i, add(ld3(S.vel, i), scl(h, fs)));
                } else {
                    st3(S.pev, i, ld3(S.pos, i));
                    st3(S.vst, i, ld3(S.vel, i));
                }
            }
        }
        block_sync();
    }
    for (int i = t; i < n; i += nthr) {
        if (S.fixed[i] != 1) {
            st3(S.pos, i, add(ld3(S.pos, i), scl(dt / 6, ld3(S.vacc, i))));
            st3(S.vel, i, add(ld3(S.vel, i), scl(dt / 6, ld3(S.facc, i))));
        }
        st3(S.frc, i, dvs(ld3(S.facc, i), 6));
        S.ener[i] = S.eacc[i] / 6;
    }
    block_sync();
}

template <bool IN_LDS>
__global__ void __launch_bounds__(256) k_system_op(SysArgs a) {
    extern __shared__ double lds_sys[];
    const int t = threadIdx.x, nthr = blockDim.x;
    const int n = a.n_atoms;
    const size_t nd = sys_doubles(n, a.n_slots, a.n_terms);
    double* base = IN_LDS ? lds_sys : a.work + (size_t)blockIdx.x * nd;
    SysState S = sys_layout(base, n, a.n_slots, a.n_terms);
    const int n_idx = 2 * (n + 1) + a.n_fcsr + a.n_ecsr + n;
    const int32_t* idx = a.idx;
    if (IN_LDS) {
        int32_t* l = reinterpret_cast<int32_t*>(lds_sys + nd);
        for (int q = t; q < n_idx; q += nthr) l[q] = a.idx[q];
        idx = l;
    }
    S.f_off = idx;
    S.f_csr = S.f_off + (n + 1);
    S.e_off = S.f_csr + a.n_fcsr;
    S.e_csr = S.e_off + (n + 1);
    S.fixed = S.e_csr + a.n_ecsr;
    for (int64_t st = blockIdx.x; st < a.n_states; st += gridDim.x) {
        const size_t o3 = (size_t)st * 3 * n, o1 = (size_t)st * n;
        for (int i = t; i < n; i += nthr) {
            st3(S.pos, i, ld3(a.pos + o3, i));
            st3(S.frc, i, ld3(a.frc + o3, i));
            S.ener[i] = a.ener[o1 + i];
            st3(S.vel, i, a.op == SYS_EVAL ? V3{0, 0, 0} : ld3(a.vel + o3, i));
        }
        block_sync();
        int n_tot = 0;
        if (a.op == SYS_EVAL) {
            eval_list(a, S, S.pos, S.frc, S.ener, t, nthr);
        } else if (a.op == SYS_STEP) {
            sys_step(a, S, a.integrator, a.dt, t, nthr);
        } else if (a.minimiser == MAFM_MIN_SD) {   // SDMinimisation (minimiser.cpp:39-65)
            for (n_tot = 1; n_tot < a.maxsteps; ++n_tot) {
                const double prev = S.ener[1];
                block_sync();
                zero_fe(a, S.frc, S.ener, t, nthr);
                block_sync();
                eval_list(a, S, S.pos, S.frc, S.ener, t, nthr);
                if (converged_sys(a, ld3(S.frc, 1), S.ener[1] - prev)) break;
                for (int i = t; i < n; i += nthr)
                    if (S.fixed[i] != 1) st3(S.pos, i, add(ld3(S.pos, i), scl(a.dt, ld3(S.frc, i))));
                block_sync();
            }
        } else {                                   // FIREMinimisation (minimiser.cpp:67-129)
            double alpha = 0.1, dt = a.dt;
            int nnn = 0;
            for (n_tot = 1; n_tot < a.maxsteps; ++n_tot) {
                const double prev = S.ener[1];
                block_sync();
                zero_fe(a, S.frc, S.ener, t, nthr);
                block_sync();
                sys_step(a, S, a.integrator, dt, t, nthr);
                if (converged_sys(a, ld3(S.frc, 1), S.ener[1] - prev)) break;
                // P over every atom in index order, then the velocity mix of the free ones
                for (int i = t; i < n; i += nthr) S.pdot[i] = dot(ld3(S.vel, i), ld3(S.frc, i));
                block_sync();
                if (t == nthr - 1) S.bc[1] = seq_sum(S.pdot, nullptr, n);
                block_sync();
                const double p = S.bc[1];
                for (int i = t; i < n; i += nthr) {
                    if (S.fixed[i] == 1) continue;
                    const V3 v = ld3(S.vel, i), f = ld3(S.frc, i);
                    const double fl = len3(f);
                    const V3 fh = fl > 1e-10 ? V3{f.x / fl, f.y / fl, f.z / fl} : V3{0, 0, 0};   // normalized()
                    st3(S.vel, i, add(scl(1 - alpha, v), scr(scl(alpha, fh), len3(v))));
                }
                block_sync();
                if (p < 0) {
                    for (int i = t; i < n; i += nthr) st3(S.vel, i, {0, 0, 0});
                    nnn = 0;
                    dt *= 0.5;
                    alpha = 0.1;
                } else {
                    if (nnn > 5) {
                        dt = fmin(dt * 1.1, 10 * a.dt);
                        alpha *= 0.99;
                    }
                    ++nnn;
                }
                block_sync();
            }
        }
        block_sync();
        for (int i = t; i < n; i += nthr) {
            st3(a.pos + o3, i, ld3(S.pos, i));
            st3(a.frc + o3, i, ld3(S.frc, i));
            a.ener[o1 + i] = S.ener[i];
            if (a.op != SYS_EVAL) st3(a.vel + o3, i, ld3(S.vel, i));
        }
        if (t == 0 && a.op == SYS_MINIMISE) a.steps[st] = n_tot;
        block_sync();
    }
}

}  // namespace

size_t flex_block_doubles(int n_atoms, int n_slots, int n_eval, bool rk4) {
    return block_doubles(n_atoms, n_slots, n_eval, rk4);
}

size_t flex_block_lds(const FlexArgs& a) {
    const bool rk4 = a.minimiser == MAFM_MIN_FIRE && a.integrator == MAFM_INT_RK4;
    const size_t n_idx = (size_t)(a.n_atoms + 1) + a.n_csr + a.n_tip + a.n_atoms;
    return block_doubles(a.n_atoms, a.n_slots, a.n_eval, rk4) * sizeof(double) + n_idx * sizeof(int32_t);
}

size_t sys_block_bytes(const SysArgs& a) {
    const size_t n_idx = 2 * (size_t)(a.n_atoms + 1) + a.n_fcsr + a.n_ecsr + a.n_atoms;
    return sys_doubles(a.n_atoms, a.n_slots, a.n_terms) * sizeof(double) + n_idx * sizeof(int32_t);
}

hipError_t launch_system_op(const SysArgs& a, int waves, bool in_lds, hipStream_t st) {
    const unsigned blocks = (unsigned)std::min<int64_t>(a.n_states, 1 << 20);
    if (in_lds) {
        const size_t lds = sys_block_bytes(a);
        hipError_t e = hipFuncSetAttribute((const void*)k_system_op<true>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_system_op<true>, dim3(blocks), dim3(64 * waves), lds, st, a);
    } else {
        hipLaunchKernelGGL(k_system_op<false>, dim3(blocks), dim3(64 * waves), 0, st, a);
    }
    return hipGetLastError();
}

hipError_t launch_scan_flex(const FlexArgs& a, int n_blocks, int waves, bool in_lds, hipStream_t st) {
    if (in_lds) {
        const size_t lds = flex_block_lds(a);
        // above 64 KiB a workgroup asks for the larger allocation (gfx950: 160 KiB)
        hipError_t e = hipFuncSetAttribute((const void*)k_scan_flex<true>,
                                           hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_scan_flex<true>, dim3(n_blocks), dim3(64 * waves), lds, st, a);
    } else {
        hipLaunchKernelGGL(k_scan_flex<false>, dim3(n_blocks), dim3(64 * waves), 0, st, a);
    }
    return hipGetLastError();
}

}  // namespace mafm
