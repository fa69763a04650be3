/*
 * TEST INFRASTRUCTURE ONLY -- the oracle.  Only tests/, the smoke check and
 * bench.py's cpu_baseline leg may load this library, and only as the checker.
 * No product path links or calls it.
 *
 * A plain-C FP64 restatement of MechAFM's scan with a flexible surface
 * (flexible on): the whole System state -- positions, velocities, forces and
 * energies of every atom -- as the reference carries it, the interaction
 * list evaluated term by term into the per-atom arrays in list order
 * (interactions.cpp:15-51, 252-362), the integrators with their temporary
 * state vectors (integrators.cpp:8-111), FIRE / steepest descent over every
 * atom (minimiser.cpp:39-129) and the column / height sweep of
 * Simulation::run (simulation.cpp:78-140).  Operation by operation, built
 * with -ffp-contract=off and glibc's libm, so its results agree with the
 * reference's to the last bit.
 *
 * Terms use the layout of mafm_flex_term (include/mafm.h); the kinds and
 * parameter slots are listed there.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct { double x, y, z; } v3;

typedef struct {
    int32_t kind;
    int32_t atom[4];
    int32_t reserved;
    double p[6];
    double shift[3];
} fx_term;

enum { FX_LJ, FX_MORSE, FX_COULOMB, FX_TIP_HARMONIC, FX_BOND, FX_ANGLE, FX_DIHEDRAL, FX_SUBSTRATE, FX_XY };

typedef struct {
    int nx, ny, nz;
    double dx, dy, dz, zhigh, tip_dummy_d;
    int minimiser, integrator, minterm;   /* 0 SD / 1 FIRE; 0 euler 1 midpoint 2 rk4; 0 e 1 f 2 ef */
    double etol, ftol, dt;
    int maxsteps;
    int tip_pbc;
    double cell[9], offset[3];
} fx_opts;

#define TOLERANCE 1e-10                   /* globals.hpp:11 */
#define PI 3.14159265358979323846          /* globals.hpp:13 */

static inline double sq(double a) { return a * a; }   /* pow(a, 2) */
static inline v3 V(double x, double y, double z) { v3 r = {x, y, z}; return r; }
static inline v3 add(v3 a, v3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub(v3 a, v3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 smul(double s, v3 a) { return V(s * a.x, s * a.y, s * a.z); }    /* T * Vector3 */
static inline v3 muls(v3 a, double s) { return V(a.x * s, a.y * s, a.z * s); }    /* Vector3 * T */
static inline v3 divs(v3 a, double s) { return V(a.x / s, a.y / s, a.z / s); }    /* Vector3 / T */
static inline double len3(v3 a) { return sqrt(sq(a.x) + sq(a.y) + sq(a.z)); }    /* vectors.hpp:103-105 */
static inline double dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }  /* vectors.hpp:116-118 */
static inline v3 cross(v3 a, v3 b) {                                               /* vectors.hpp:120-122 */
    return V(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
static inline v3 normalized(v3 a) {                                                /* vectors.hpp:107-114 */
    double l = len3(a);
    return l > 1e-10 ? V(a.x / l, a.y / l, a.z / l) : V(0, 0, 0);
}

/* Interaction::eval of one term: adds to forces / energies (interactions.cpp). */
static void eval_term(const fx_term* t, const v3* P, v3* F, double* E) {
    const int* a = t->atom;
    switch (t->kind) {
    case FX_LJ: {                                   /* :15-27 */
        v3 r = sub(P[a[0]], add(P[a[1]], V(t->shift[0], t->shift[1], t->shift[2])));
        double r_sqr = sq(r.x) + sq(r.y) + sq(r.z);
        double r6 = r_sqr * r_sqr * r_sqr;
        double term_a = t->p[1] / (r6 * r6);
        double term_b = t->p[0] / r6;
        double e = term_a - term_b;
        v3 f = smul((12 * term_a - 6 * term_b) / r_sqr, r);
        E[a[0]] += e; E[a[1]] += e;
        F[a[0]] = add(F[a[0]], f); F[a[1]] = sub(F[a[1]], f);
        return;
    }
    case FX_MORSE: {                                /* :29-40 */
        v3 r = sub(P[a[0]], add(P[a[1]], V(t->shift[0], t->shift[1], t->shift[2])));
        double rl = len3(r);
        double d_exp = exp(-t->p[1] * (rl - t->p[2]));
        double e = t->p[0] * (sq(d_exp) - 2 * d_exp + 1);
        double f_abs = 2 * t->p[0] * t->p[1] * (sq(d_exp) - d_exp);
        v3 f = smul(f_abs / rl, r);
        E[a[0]] += e; E[a[1]] += e;
        F[a[0]] = add(F[a[0]], f); F[a[1]] = sub(F[a[1]], f);
        return;
    }
    case FX_COULOMB: {                              /* :42-51 */
        v3 r = sub(P[a[0]], P[a[1]]);
        double rl = len3(r);
        double e = t->p[0] / rl;
        v3 f = smul(t->p[0] / (rl * rl * rl), r);
        E[a[0]] += e; E[a[1]] += e;
        F[a[0]] = add(F[a[0]], f); F[a[1]] = sub(F[a[1]], f);
        return;
    }
    case FX_TIP_HARMONIC: {                         /* :252-266 */
        v3 r = sub(P[a[0]], P[a[1]]);
        double rl = sqrt(sq(r.x) + sq(r.y));
        double dr = rl - t->p[1];
        double e = t->p[0] * dr * dr;
        v3 f = V(0, 0, 0);
        if (rl > TOLERANCE) {
            double s = -2 * t->p[0] * dr / rl;
            f = V(s * r.x, s * r.y, 0);
        }
        E[a[0]] += e; E[a[1]] += e;
        F[a[0]] = add(F[a[0]], f); F[a[1]] = sub(F[a[1]], f);
        return;
    }
    case FX_XY: {                                   /* :268-278 */
        double rx = P[a[0]].x - t->p[1], ry = P[a[0]].y - t->p[2];
        double rl = sqrt(sq(rx) + sq(ry));
        double e = t->p[0] * rl * rl;
        v3 f = V(0, 0, 0);
        if (rl > TOLERANCE) f = V(-2 * t->p[0] * rx, -2 * t->p[0] * ry, 0);
        E[a[0]] += e;
        F[a[0]] = add(F[a[0]], f);
        return;
    }
    case FX_SUBSTRATE: {                            /* :280-294; p = multiplier ulj ushift z0 rc sig */
        double dz = P[a[0]].z - t->p[3];
        double sig_z = t->p[5] / dz;
        double sig_z2 = sig_z * sig_z;
        double sig_z4 = sig_z2 * sig_z2;
        double sig_z10 = sig_z4 * sig_z4 * sig_z2;
        double eterm = 0.0, fterm = 0.0;
        if (dz <= t->p[4]) {
            eterm += t->p[0] * (2. / 5 * sig_z10 - sig_z4) + t->p[1] * dz * dz - t->p[2];
            fterm += 4 * t->p[0] / dz * (sig_z10 - sig_z4) + 2 * t->p[1] * dz;
        }
        F[a[0]].z += fterm;
        E[a[0]] += eterm;
        return;
    }
    case FX_BOND: {                                 /* :296-306 */
        v3 r = sub(P[a[0]], P[a[1]]);
        double rl = len3(r);
        double dr = rl - t->p[1];
        double e = t->p[0] * sq(dr);
        v3 f = smul(-2 * t->p[0] * dr / rl, r);
        E[a[0]] += e; E[a[1]] += e;
        F[a[0]] = add(F[a[0]], f); F[a[1]] = sub(F[a[1]], f);
        return;
    }
    case FX_ANGLE: {                                /* :308-332; atom = shared, i1, i2 */
        v3 r1v = sub(P[a[1]], P[a[0]]), r2v = sub(P[a[2]], P[a[0]]);
        double r1 = len3(r1v), r2 = len3(r2v);
        double cos_t = dot(r1v, r2v) / (r1 * r2);
        if (cos_t > 1) cos_t = 1;
        else if (cos_t < -1) cos_t = -1;
        double theta = acos(cos_t);
        double sin_t = sin(theta);
        double d_theta = theta - t->p[1];
        double fm = -2 * t->p[0] * d_theta / sin_t;
        v3 f1 = smul(-fm / r1, sub(divs(r2v, r2), divs(muls(r1v, cos_t), r1)));
        v3 f2 = smul(-fm / r2, sub(divs(r1v, r1), divs(muls(r2v, cos_t), r2)));
        double e = t->p[0] * sq(d_theta);
        F[a[1]] = add(F[a[1]], f1);
        F[a[2]] = add(F[a[2]], f2);
        F[a[0]] = sub(F[a[0]], add(f1, f2));
        E[a[1]] += e; E[a[2]] += e; E[a[0]] += e;
        return;
    }
    case FX_DIHEDRAL: {                             /* :334-362 */
        v3 r12 = sub(P[a[0]], P[a[1]]), r23 = sub(P[a[1]], P[a[2]]), r43 = sub(P[a[3]], P[a[2]]);
        double r23l = len3(r23);
        v3 mv = cross(r12, r23), nv = cross(r43, r23);
        double m = len3(mv), n = len3(nv);
        double tan_s = dot(nv, r12) * r23l / dot(mv, nv);
        double sigma = atan(tan_s);
        double d_sigma = sigma - t->p[1];
        double fm = 2 * t->p[0] * d_sigma;
        v3 f1 = smul(-fm * r23l / (m * m), mv);
        v3 f2 = smul(-fm, sub(smul(dot(r43, r23) / (n * n * r23l), nv),
                              smul(((r23l * r23l) + dot(r12, r23)) / (m * m * r23l), mv)));
        v3 f3 = smul(-fm, add(smul(dot(r12, r23) / (m * m * r23l), mv),
                              smul(((r23l * r23l) - dot(r43, r23)) / (n * n * r23l), nv)));
        v3 f4 = smul(fm * r23l / (n * n), nv);
        double e = t->p[0] * sq(d_sigma);
        F[a[0]] = add(F[a[0]], f1);
        F[a[1]] = add(F[a[1]], f2);
        F[a[2]] = add(F[a[2]], f3);
        F[a[3]] = add(F[a[3]], f4);
        E[a[0]] += e; E[a[1]] += e; E[a[2]] += e; E[a[3]] += e;
        return;
    }
    default:
        return;
    }
}

/* isTipSurface (interactions.hpp:79-137) */
static int tip_surface(const fx_term* t) {
    return (t->kind == FX_LJ || t->kind == FX_MORSE || t->kind == FX_COULOMB) && t->atom[0] == 1;
}

static void eval_all(const fx_term* T, int nt, int n, const v3* P, v3* F, double* E, int only_ts) {
    for (int i = 0; i < n; ++i) { F[i] = V(0, 0, 0); E[i] = 0; }
    for (int k = 0; k < nt; ++k)
        if (!only_ts || tip_surface(&T[k])) eval_term(&T[k], P, F, E);
}

/* Whole-list evaluation at positions pos (n x 3): forces (n x 3), energies
 * (n); only_ts = System::evalTipSurfaceForces's subset (system.cpp:41-51). */
int oracle_flex_eval(int n, const fx_term* T, int nt, const double* pos, double* forces, double* energies,
                     int only_ts) {
    if (n <= 0) return 1;
    v3* P = (v3*)calloc((size_t)n, sizeof(v3));
    v3* F = (v3*)calloc((size_t)n, sizeof(v3));
    for (int i = 0; i < n; ++i) P[i] = V(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]);
    eval_all(T, nt, n, P, F, energies, only_ts);
    for (int i = 0; i < n; ++i) { forces[3 * i] = F[i].x; forces[3 * i + 1] = F[i].y; forces[3 * i + 2] = F[i].z; }
    free(P);
    free(F);
    return 0;
}

typedef struct {
    int n;
    const fx_term* T;
    int nt;
    const int* fixed;
    v3 *pos, *vel, *frc;
    double* en;
    /* integrator temporaries */
    v3 *p2, *p3, *p4, *v2, *v3_, *v4, *f1, *f2, *f3, *f4;
    double *e1, *e2, *e3, *e4;
} fx_sys;

static void euler_step(fx_sys* s, double dt) {     /* integrators.cpp:8-22 */
    for (int k = 0; k < s->nt; ++k) eval_term(&s->T[k], s->pos, s->frc, s->en);
    for (int i = 0; i < s->n; ++i)
        if (s->fixed[i] != 1) {
            s->pos[i] = add(s->pos[i], smul(dt, s->vel[i]));
            s->vel[i] = add(s->vel[i], divs(smul(dt, s->frc[i]), 1.0));   /* masses are 1 */
        }
}

static void zero(fx_sys* s, v3* f, double* e) {
    for (int i = 0; i < s->n; ++i) { f[i] = V(0, 0, 0); e[i] = 0; }
}

static void midpoint_step(fx_sys* s, double dt) {  /* integrators.cpp:24-54 */
    const int n = s->n;
    memcpy(s->p2, s->pos, sizeof(v3) * n);
    memcpy(s->v2, s->vel, sizeof(v3) * n);
    zero(s, s->f1, s->e1);
    zero(s, s->f2, s->e2);
    for (int k = 0; k < s->nt; ++k) eval_term(&s->T[k], s->pos, s->f1, s->e1);
    for (int i = 0; i < n; ++i)
        if (s->fixed[i] != 1) {
            s->p2[i] = add(s->p2[i], smul(dt / 2, s->vel[i]));
            s->v2[i] = add(s->v2[i], divs(smul(dt / 2, s->f1[i]), 1.0));
        }
    for (int k = 0; k < s->nt; ++k) eval_term(&s->T[k], s->p2, s->f2, s->e2);
    for (int i = 0; i < n; ++i) {
        if (s->fixed[i] != 1) {
            s->pos[i] = add(s->pos[i], smul(dt, s->v2[i]));
            s->vel[i] = add(s->vel[i], divs(smul(dt, s->f2[i]), 1.0));
        }
        s->frc[i] = s->f2[i];
        s->en[i] = s->e2[i];
    }
}

static void rk4_step(fx_sys* s, double dt) {       /* integrators.cpp:56-111 */
    const int n = s->n;
    memcpy(s->p2, s->pos, sizeof(v3) * n);
    memcpy(s->p3, s->pos, sizeof(v3) * n);
    memcpy(s->p4, s->pos, sizeof(v3) * n);
    memcpy(s->v2, s->vel, sizeof(v3) * n);
    memcpy(s->v3_, s->vel, sizeof(v3) * n);
    memcpy(s->v4, s->vel, sizeof(v3) * n);
    zero(s, s->f1, s->e1);
    zero(s, s->f2, s->e2);
    zero(s, s->f3, s->e3);
    zero(s, s->f4, s->e4);
    for (int k = 0; k < s->nt; ++k) eval_term(&s->T[k], s->pos, s->f1, s->e1);
    for (int i = 0; i < n; ++i)
        if (s->fixed[i] != 1) {
            s->p2[i] = add(s->p2[i], smul(dt / 2, s->vel[i]));
            s->v2[i] = add(s->v2[i], divs(smul(dt / 2, s->f1[i]), 1.0));
        }
    for (int k = 0; k < s->nt; ++k) eval_term(&s->T[k], s->p2, s->f2, s->e2);
    for (int i = 0; i < n; ++i)
        if (s->fixed[i] != 1) {
            s->p3[i] = add(s->p3[i], smul(dt / 2, s->v2[i]));
            s->v3_[i] = add(s->v3_[i], divs(smul(dt / 2, s->f2[i]), 1.0));
        }
    for (int k = 0; k < s->nt; ++k) eval_term(&s->T[k], s->p3, s->f3, s->e3);
    for (int i = 0; i < n; ++i)
        if (s->fixed[i] != 1) {
            s->p4[i] = add(s->p4[i], smul(dt, s->v3_[i]));
            s->v4[i] = add(s->v4[i], divs(smul(dt, s->f3[i]), 1.0));
        }
    for (int k = 0; k < s->nt; ++k) eval_term(&s->T[k], s->p4, s->f4, s->e4);
    for (int i = 0; i < n; ++i) {
        v3 fsum = add(add(add(s->f1[i], smul(2, s->f2[i])), smul(2, s->f3[i])), s->f4[i]);
        if (s->fixed[i] != 1) {
            v3 vsum = add(add(add(s->vel[i], smul(2, s->v2[i])), smul(2, s->v3_[i])), s->v4[i]);
            s->pos[i] = add(s->pos[i], smul(dt / 6, vsum));
            s->vel[i] = add(s->vel[i], divs(smul(dt / 6, fsum), 1.0));
        }
        s->frc[i] = divs(fsum, 6);
        s->en[i] = (s->e1[i] + 2 * s->e2[i] + 2 * s->e3[i] + s->e4[i]) / 6;
    }
}

static int converged(const fx_opts* o, v3 f, double ediff) {   /* minimiser.cpp:15-37 */
    switch (o->minterm) {
    case 0: return fabs(ediff) < o->etol;
    case 1: return len3(f) < o->ftol;
    default: return fabs(ediff) < o->etol && len3(f) < o->ftol;
    }
}

static int sd_min(fx_sys* s, const fx_opts* o) {  /* minimiser.cpp:39-62 */
    double prev = -10e6;
    int n = 1;
    for (; n < o->maxsteps; ++n) {
        if (n != 0) prev = s->en[1];
        zero(s, s->frc, s->en);
        for (int k = 0; k < s->nt; ++k) eval_term(&s->T[k], s->pos, s->frc, s->en);
        if (converged(o, s->frc[1], s->en[1] - prev)) break;
        for (int i = 0; i < s->n; ++i)
            if (s->fixed[i] != 1) s->pos[i] = add(s->pos[i], smul(o->dt, s->frc[i]));
    }
    return n;
}

static int fire_min(fx_sys* s, const fx_opts* o) { /* minimiser.cpp:64-129 */
    const int n_min = 5;
    const double f_inc = 1.1, f_dec = 0.5, f_a = 0.99, a_start = 0.1;
    const double dt_max = 10 * o->dt;
    double alpha = a_start, dt = o->dt;
    int n_non_neg = 0;
    double prev = -10e6;
    int n_tot = 1;
    for (; n_tot < o->maxsteps; ++n_tot) {
        if (n_tot != 0) prev = s->en[1];
        zero(s, s->frc, s->en);
        if (o->integrator == 0) euler_step(s, dt);
        else if (o->integrator == 1) midpoint_step(s, dt);
        else rk4_step(s, dt);
        if (converged(o, s->frc[1], s->en[1] - prev)) break;
        double p = 0;
        for (int i = 0; i < s->n; ++i) {
            p += dot(s->vel[i], s->frc[i]);
            if (s->fixed[i] != 1)
                s->vel[i] = add(smul(1 - alpha, s->vel[i]),
                                muls(smul(alpha, normalized(s->frc[i])), len3(s->vel[i])));
        }
        if (p < 0) {
            for (int i = 0; i < s->n; ++i) s->vel[i] = V(0, 0, 0);
            n_non_neg = 0;
            dt *= f_dec;
            alpha = a_start;
        } else {
            if (n_non_neg > n_min) {
                double d = dt * f_inc;
                dt = dt_max < d ? dt_max : d;   /* std::min */
                alpha *= f_a;
            }
            n_non_neg++;
        }
    }
    return n_tot;
}

static void mat_vec(const double* m, v3 r, v3* o) {   /* Matrix3::multiply (matrices.hpp:116-125) */
    o->x = m[0] * r.x + m[1] * r.y + m[2] * r.z;
    o->y = m[3] * r.x + m[4] * r.y + m[5] * r.z;
    o->z = m[6] * r.x + m[7] * r.y + m[8] * r.z;
}

static void inv3(const double* m, double* o) {     /* Matrix3::inverse (matrices.hpp:89-114) */
    double det = m[0] * (m[4] * m[8] - m[5] * m[7]);
    det += m[1] * (m[5] * m[6] - m[3] * m[8]);
    det += m[2] * (m[3] * m[7] - m[4] * m[6]);
    double inv_det = 1 / det;
    double a = m[0], b = m[1], c = m[2], d = m[3], e = m[4], f = m[5], g = m[6], h = m[7], i = m[8];
    o[0] = inv_det * (e * i - f * h);
    o[1] = -inv_det * (b * i - c * h);
    o[2] = inv_det * (b * f - c * e);
    o[3] = -inv_det * (d * i - f * g);
    o[4] = inv_det * (a * i - c * g);
    o[5] = -inv_det * (a * f - c * d);
    o[6] = inv_det * (d * h - e * g);
    o[7] = -inv_det * (a * h - b * g);
    o[8] = inv_det * (a * e - b * d);
}

/* The column sweep of Simulation::run (simulation.cpp:78-140) for columns
 * [col_first, col_first + col_count) of the nx * ny scan.  rec: 16 doubles
 * per position at ((col - col_first) * nz + k) in ref_harness's layout
 * (i j k n x y z fx fy fz rx ry rz r angle e); state (optional): the atoms
 * of column state_col after each height's minimisation, nz x n x 3. */
/* The System-level operators on one state (interactions.cpp:15-362,
 * integrators.cpp:8-111, minimiser.cpp:39-129 as the functions above restate
 * them): op 0 adds every term of the list onto forces / energies as given;
 * op 1 is one eulerStep / midpointStep / rk4Step (`integrator`, `dt`); op 2
 * is SDMinimisation / FIREMinimisation with o's minimiser, integrator,
 * minterm, maxsteps, etol, ftol and dt from the state, energies[1] being
 * the tip energy the first test compares with.  pos / vel / forces (n x 3)
 * and energies (n) are read and written; returns the minimiser's step
 * count for op 2, else 0. */
int oracle_flex_system(int n, const fx_term* T, int nt, const int* fixed, int op, const fx_opts* o, int integrator,
                       double dt, double* pos, double* vel, double* forces, double* energies) {
    fx_sys s;
    memset(&s, 0, sizeof s);
    s.n = n; s.T = T; s.nt = nt; s.fixed = fixed;
    v3* buf = (v3*)calloc((size_t)13 * n, sizeof(v3));
    double* ebuf = (double*)calloc((size_t)5 * n, sizeof(double));
    s.pos = buf; s.vel = buf + n; s.frc = buf + 2 * n;
    s.p2 = buf + 3 * n; s.p3 = buf + 4 * n; s.p4 = buf + 5 * n;
    s.v2 = buf + 6 * n; s.v3_ = buf + 7 * n; s.v4 = buf + 8 * n;
    s.f1 = buf + 9 * n; s.f2 = buf + 10 * n; s.f3 = buf + 11 * n; s.f4 = buf + 12 * n;
    s.en = ebuf; s.e1 = ebuf + n; s.e2 = ebuf + 2 * n; s.e3 = ebuf + 3 * n; s.e4 = ebuf + 4 * n;
    for (int a = 0; a < n; ++a) {
        s.pos[a] = V(pos[3 * a], pos[3 * a + 1], pos[3 * a + 2]);
        s.vel[a] = vel ? V(vel[3 * a], vel[3 * a + 1], vel[3 * a + 2]) : V(0, 0, 0);
        s.frc[a] = V(forces[3 * a], forces[3 * a + 1], forces[3 * a + 2]);
        s.en[a] = energies[a];
    }
    int ret = 0;
    if (op == 0) {
        for (int k = 0; k < nt; ++k) eval_term(&T[k], s.pos, s.frc, s.en);
    } else if (op == 1) {
        if (integrator == 0) euler_step(&s, dt);
        else if (integrator == 1) midpoint_step(&s, dt);
        else rk4_step(&s, dt);
    } else {
        ret = o->minimiser == 0 ? sd_min(&s, o) : fire_min(&s, o);
    }
    for (int a = 0; a < n; ++a) {
        pos[3 * a] = s.pos[a].x; pos[3 * a + 1] = s.pos[a].y; pos[3 * a + 2] = s.pos[a].z;
        if (vel) { vel[3 * a] = s.vel[a].x; vel[3 * a + 1] = s.vel[a].y; vel[3 * a + 2] = s.vel[a].z; }
        forces[3 * a] = s.frc[a].x; forces[3 * a + 1] = s.frc[a].y; forces[3 * a + 2] = s.frc[a].z;
        energies[a] = s.en[a];
    }
    free(buf);
    free(ebuf);
    return ret;
}

int oracle_flex_scan(int n, const fx_term* T, int nt, const double* pos0, const int* fixed, const fx_opts* o,
                     int col_first, int col_count, double* rec, int state_col, double* state, int64_t* steps) {
    int64_t total = 0;
    double inv[9];
    if (o->tip_pbc) inv3(o->cell, inv);
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : total)
    for (int cc = 0; cc < col_count; ++cc) {
        const int col = col_first + cc;
        const int i = col / o->ny, j = col % o->ny;
        const double x = i * o->dx, y = j * o->dy;
        fx_sys s;
        memset(&s, 0, sizeof s);
        s.n = n; s.T = T; s.nt = nt; s.fixed = fixed;
        v3* buf = (v3*)calloc((size_t)13 * n, sizeof(v3));
        double* ebuf = (double*)calloc((size_t)5 * n, sizeof(double));
        s.pos = buf; s.vel = buf + n; s.frc = buf + 2 * n;
        s.p2 = buf + 3 * n; s.p3 = buf + 4 * n; s.p4 = buf + 5 * n;
        s.v2 = buf + 6 * n; s.v3_ = buf + 7 * n; s.v4 = buf + 8 * n;
        s.f1 = buf + 9 * n; s.f2 = buf + 10 * n; s.f3 = buf + 11 * n; s.f4 = buf + 12 * n;
        s.en = ebuf; s.e1 = ebuf + n; s.e2 = ebuf + 2 * n; s.e3 = ebuf + 3 * n; s.e4 = ebuf + 4 * n;
        for (int a = 0; a < n; ++a) s.pos[a] = V(pos0[3 * a], pos0[3 * a + 1], pos0[3 * a + 2]);
        /* setDummyXY (system.cpp:139-162) and setDummyZ (system.hpp:47-50) */
        double px = x, py = y;
        if (o->tip_pbc) {
            v3 q, w;
            mat_vec(inv, V(x - o->offset[0], y - o->offset[1], 0 - o->offset[2]), &q);
            q.x -= floor(q.x);
            q.y -= floor(q.y);
            mat_vec(o->cell, q, &w);
            px = w.x + o->offset[0];
            py = w.y + o->offset[1];
        }
        s.pos[0] = V(px, py, o->zhigh);
        s.pos[1] = V(px, py, o->zhigh - o->tip_dummy_d);
        v3* tf = (v3*)calloc((size_t)n, sizeof(v3));
        double* te = (double*)calloc((size_t)n, sizeof(double));
        for (int k = 0; k < o->nz; ++k) {
            int nn = o->minimiser == 0 ? sd_min(&s, o) : fire_min(&s, o);
            total += nn;
            if (state && col == state_col)
                for (int a = 0; a < n; ++a) {
                    double* q = state + ((size_t)k * n + a) * 3;
                    q[0] = s.pos[a].x; q[1] = s.pos[a].y; q[2] = s.pos[a].z;
                }
            /* getOutput (system.cpp:28-39) */
            eval_all(T, nt, n, s.pos, tf, te, 1);
            double* r = rec + ((size_t)cc * o->nz + k) * 16;
            v3 rv = sub(s.pos[1], s.pos[0]);
            r[0] = i; r[1] = j; r[2] = k; r[3] = nn;
            r[4] = x; r[5] = y; r[6] = s.pos[0].z;
            r[7] = tf[1].x; r[8] = tf[1].y; r[9] = tf[1].z;
            r[10] = rv.x; r[11] = rv.y; r[12] = rv.z;
            r[13] = len3(rv);
            r[14] = atan2(sqrt(sq(rv.x) + sq(rv.y)), rv.z) * (180.0 / PI);
            r[15] = te[1];
            /* lowerTip (system.hpp:51-54) */
            s.pos[0].z -= o->dz;
            s.pos[1].z -= o->dz;
        }
        free(tf);
        free(te);
        free(buf);
        free(ebuf);
    }
    if (steps) *steps = total;
    return 0;
}
