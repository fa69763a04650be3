/*
 * TEST INFRASTRUCTURE ONLY -- the oracle.  Only tests/, the smoke check and
 * bench.py's cpu_baseline leg may load this library, and only as the checker.
 * No product path links or calls it.
 *
 * A plain-C FP64 restatement of MechAFM's rigid probe-particle scan.  Every
 * function follows the reference operation by operation (same evaluation
 * order, no FMA contraction: built with -ffp-contract=off) so that its results
 * agree with the reference to rounding and its FIRE step counts agree exactly.
 *
 * Simplification that is exact for rigid scans: the dummy (index 0) and every
 * surface atom (index >= 2) are fixed (parse.cpp:563, system.cpp:23-24) and
 * keep zero velocity, so only the tip's force / energy / velocity are carried.
 * The FIRE power P = sum_i v_i . f_i (minimiser.cpp:107-109) reduces to
 * v_tip . f_tip because every other term is +0.0.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef struct { double x, y, z; } v3;

/* Interaction kinds (interactions.hpp:79-196). */
enum { OR_LJ = 0, OR_MORSE = 1, OR_COUL = 2, OR_HARM = 3, OR_GRID = 4 };

typedef struct {
    int kind;
    int i1, i2;        /* atom indices: 0 dummy, 1 tip, >= 2 surface */
    int grid;          /* OR_GRID: index into the grid table */
    double p[3];       /* LJ: es6 es12 | Morse: de a re | Coulomb: qq | Harm: k r0 */
    double shift[3];   /* pbc shift of atom i2 (LJ, Morse) */
} or_inter;

typedef struct {
    int n[3];
    int periodic;
    int orthogonal;
    double basis[9];   /* row-major, basis[3*i+j] = Mat3d::at(i, j) */
    double offset[3];
    const double* fe;  /* (nx*ny*nz) x 4: fx fy fz e, index ix*ny*nz + iy*nz + iz */
} or_grid;

typedef struct {
    int nx, ny, nz;
    double dx, dy, dz, zhigh;
    double tip_dummy_d;
    int minimiser;     /* 0 SD, 1 FIRE */
    int integrator;    /* 0 euler, 1 midpoint, 2 rk4 */
    int minterm;       /* 0 e, 1 f, 2 ef */
    double etol, ftol, dt;
    int maxsteps;
    int tip_pbc;       /* System::setDummyXY wraps into the cell (system.cpp:139-162) */
    double cell[9];
    double offset[3];
} or_scan_opts;

typedef struct {
    const or_inter* inter;
    int n_inter;
    const or_grid* grids;
    const double* atoms;   /* (n_atoms) x 3 positions, index 0/1 unused */
    int64_t warnings;      /* grid clamp warnings (force_grid.cpp:181-217) */
    int error;             /* grid index out of range (force_grid.cpp:236-240) */
} or_ctx;

static inline double sq(double a) { return a * a; }          /* pow(a, 2): exact square */
static inline double len3(v3 a) { return sqrt((sq(a.x) + sq(a.y)) + sq(a.z)); }  /* vectors.hpp:99-101 */

/* Matrix3::determinant / inverse / multiply (matrices.hpp:89-153). */
static double det3(const double* m) {
    double det = m[0] * (m[4] * m[8] - m[5] * m[7]);
    det += m[1] * (m[5] * m[6] - m[3] * m[8]);
    det += m[2] * (m[3] * m[7] - m[4] * m[6]);
    return det;
}
static void inv3(const double* m, double* o) {
    double inv_det = 1 / det3(m);
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
static v3 mul3(const double* m, v3 r) {
    v3 o;
    o.x = m[0] * r.x + m[1] * r.y + m[2] * r.z;
    o.y = m[3] * r.x + m[4] * r.y + m[5] * r.z;
    o.z = m[6] * r.x + m[7] * r.y + m[8] * r.z;
    return o;
}

/* ForceGrid::getGridPoint / getGridPointIndex / pbcGridPoint / interpolate
 * (force_grid.cpp:88-275). */
static int pbc_wrap(int g, int n) {
    if (g < 0 || g >= n) {
        int k = (int)floor((double)g / n);
        g -= k * n;
    }
    return g;
}

static void grid_interpolate(or_ctx* c, const or_grid* gr, v3 pos, v3* force, double* energy) {
    int gp[3];
    double pg[3];
    if (gr->orthogonal) {
        gp[0] = (int)floor((pos.x - gr->offset[0]) / gr->basis[0]);
        gp[1] = (int)floor((pos.y - gr->offset[1]) / gr->basis[4]);
        gp[2] = (int)floor((pos.z - gr->offset[2]) / gr->basis[8]);
    } else {
        double inv[9];
        inv3(gr->basis, inv);
        v3 rel = {pos.x - gr->offset[0], pos.y - gr->offset[1], pos.z - gr->offset[2]};
        v3 q = mul3(inv, rel);
        pg[0] = q.x; pg[1] = q.y; pg[2] = q.z;
        gp[0] = (int)floor(q.x); gp[1] = (int)floor(q.y); gp[2] = (int)floor(q.z);
    }
    if (!gr->periodic) {
        for (int a = 0; a < 3; ++a) {
            if (gp[a] < 0) { gp[a] = 0; c->warnings++; }
            else if (gp[a] >= gr->n[a]) { gp[a] = gr->n[a] - 1; c->warnings++; }
        }
    }
    double fs[8][3], es[8];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int k = 0; k < 2; ++k) {
                int q[3] = {gp[0] + i, gp[1] + j, gp[2] + k};
                int out = q[0] < 0 || q[1] < 0 || q[2] < 0 || q[0] >= gr->n[0] || q[1] >= gr->n[1] || q[2] >= gr->n[2];
                if (out) {
                    if (gr->periodic) {
                        q[0] = pbc_wrap(q[0], gr->n[0]);
                        q[1] = pbc_wrap(q[1], gr->n[1]);
                        q[2] = pbc_wrap(q[2], gr->n[2]);
                    } else {
                        c->error = 1;
                        q[0] = q[0] < gr->n[0] ? q[0] : gr->n[0] - 1;
                        q[1] = q[1] < gr->n[1] ? q[1] : gr->n[1] - 1;
                        q[2] = q[2] < gr->n[2] ? q[2] : gr->n[2] - 1;
                    }
                }
                long idx = (long)q[0] * gr->n[1] * gr->n[2] + (long)q[1] * gr->n[2] + q[2];
                int s = i * 4 + j * 2 + k;
                fs[s][0] = gr->fe[4 * idx]; fs[s][1] = gr->fe[4 * idx + 1]; fs[s][2] = gr->fe[4 * idx + 2];
                es[s] = gr->fe[4 * idx + 3];
            }
    double d[3];
    if (gr->orthogonal) {
        d[0] = (pos.x - gr->offset[0]) / gr->basis[0] - gp[0];
        d[1] = (pos.y - gr->offset[1]) / gr->basis[4] - gp[1];
        d[2] = (pos.z - gr->offset[2]) / gr->basis[8] - gp[2];
    } else {
        d[0] = pg[0] - gp[0]; d[1] = pg[1] - gp[1]; d[2] = pg[2] - gp[2];
    }
    double out[3];
    for (int a = 0; a < 3; ++a) {
        double f00 = fs[0][a] * (1 - d[0]) + fs[4][a] * d[0];
        double f01 = fs[1][a] * (1 - d[0]) + fs[5][a] * d[0];
        double f10 = fs[2][a] * (1 - d[0]) + fs[6][a] * d[0];
        double f11 = fs[3][a] * (1 - d[0]) + fs[7][a] * d[0];
        double f0 = f00 * (1 - d[1]) + f10 * d[1];
        double f1 = f01 * (1 - d[1]) + f11 * d[1];
        out[a] = f0 * (1 - d[2]) + f1 * (d[2]);
    }
    force->x = out[0]; force->y = out[1]; force->z = out[2];
    double e00 = es[0] * (1 - d[0]) + es[4] * d[0];
    double e01 = es[1] * (1 - d[0]) + es[5] * d[0];
    double e10 = es[2] * (1 - d[0]) + es[6] * d[0];
    double e11 = es[3] * (1 - d[0]) + es[7] * d[0];
    double e0 = e00 * (1 - d[1]) + e10 * d[1];
    double e1 = e01 * (1 - d[1]) + e11 * d[1];
    *energy = e0 * (1 - d[2]) + e1 * (d[2]);
}

/* Evaluate every interaction touching the tip, in list order, accumulating
 * the tip's force and energy as Interaction::eval does (interactions.cpp:15-51,
 * 236-266).  only_tip_surface selects System::evalTipSurfaceForces
 * (system.cpp:41-51): isTipSurface() is i1 == 1 for pair terms and true for
 * grids. */
static void eval_tip(or_ctx* c, v3 dummy, v3 tip, int only_tip_surface, v3* F, double* E) {
    v3 f = {0, 0, 0};
    double e = 0;
    for (int n = 0; n < c->n_inter; ++n) {
        const or_inter* it = &c->inter[n];
        if (it->kind == OR_GRID) {
            v3 gf; double ge;
            grid_interpolate(c, &c->grids[it->grid], tip, &gf, &ge);
            f.x += gf.x; f.y += gf.y; f.z += gf.z;
            e += ge;
            continue;
        }
        int tip_surface = (it->kind != OR_HARM) && it->i1 == 1;
        if (only_tip_surface && !tip_surface) continue;
        v3 p1, p2;
        p1 = it->i1 == 1 ? tip : (it->i1 == 0 ? dummy : *(const v3*)&c->atoms[3 * it->i1]);
        p2 = it->i2 == 1 ? tip : (it->i2 == 0 ? dummy : *(const v3*)&c->atoms[3 * it->i2]);
        v3 fv = {0, 0, 0};
        double ev = 0;
        if (it->kind == OR_LJ) {               /* interactions.cpp:15-27 */
            v3 r = {p1.x - (p2.x + it->shift[0]), p1.y - (p2.y + it->shift[1]), p1.z - (p2.z + it->shift[2])};
            double r_sqr = (sq(r.x) + sq(r.y)) + sq(r.z);
            double r6 = r_sqr * r_sqr * r_sqr;
            double term_a = it->p[1] / (r6 * r6);
            double term_b = it->p[0] / r6;
            ev = term_a - term_b;
            double s = (12 * term_a - 6 * term_b) / r_sqr;
            fv.x = s * r.x; fv.y = s * r.y; fv.z = s * r.z;
        } else if (it->kind == OR_MORSE) {     /* interactions.cpp:29-40 */
            v3 r = {p1.x - (p2.x + it->shift[0]), p1.y - (p2.y + it->shift[1]), p1.z - (p2.z + it->shift[2])};
            double rl = len3(r);
            double d_exp = exp(-it->p[1] * (rl - it->p[2]));
            ev = it->p[0] * (sq(d_exp) - 2 * d_exp + 1);
            double f_abs = 2 * it->p[0] * it->p[1] * (sq(d_exp) - d_exp);
            double s = f_abs / rl;
            fv.x = s * r.x; fv.y = s * r.y; fv.z = s * r.z;
        } else if (it->kind == OR_COUL) {      /* interactions.cpp:42-51 */
            v3 r = {p1.x - p2.x, p1.y - p2.y, p1.z - p2.z};
            double rl = len3(r);
            ev = it->p[0] / rl;
            double s = it->p[0] / (rl * rl * rl);
            fv.x = s * r.x; fv.y = s * r.y; fv.z = s * r.z;
        } else if (it->kind == OR_HARM) {      /* interactions.cpp:252-266 */
            v3 r = {p1.x - p2.x, p1.y - p2.y, p1.z - p2.z};
            double rl = sqrt(sq(r.x) + sq(r.y));
            double dr = rl - it->p[1];
            ev = it->p[0] * dr * dr;
            if (rl > 1e-10) {
                double s = -2 * it->p[0] * dr / rl;
                fv.x = s * r.x; fv.y = s * r.y; fv.z = 0;
            }
        }
        /* forces[i1] += f; forces[i2] -= f; energies[i1], energies[i2] += e */
        if (it->i1 == 1) { f.x += fv.x; f.y += fv.y; f.z += fv.z; e += ev; }
        else if (it->i2 == 1) { f.x -= fv.x; f.y -= fv.y; f.z -= fv.z; e += ev; }
    }
    *F = f;
    *E = e;
}

typedef struct {
    v3 dummy, tip, v, f;
    double e;
} or_state;

/* checkConvergence (minimiser.cpp:15-37). */
static int converged(const or_scan_opts* o, v3 f, double ediff) {
    switch (o->minterm) {
        case 0: return fabs(ediff) < o->etol;
        case 1: return len3(f) < o->ftol;
        default: return fabs(ediff) < o->etol && len3(f) < o->ftol;
    }
}

/* eulerStep / midpointStep / rk4Step (integrators.cpp:8-111), tip only,
 * masses are 1 (system.cpp:18). */
static void step(or_ctx* c, const or_scan_opts* o, or_state* s, double dt) {
    const double m = 1.0;
    if (o->integrator == 0) {
        v3 f; double e;
        eval_tip(c, s->dummy, s->tip, 0, &f, &e);
        s->tip.x += dt * s->v.x; s->tip.y += dt * s->v.y; s->tip.z += dt * s->v.z;
        s->v.x += dt * f.x / m; s->v.y += dt * f.y / m; s->v.z += dt * f.z / m;
        s->f = f; s->e = e;
    } else if (o->integrator == 1) {
        v3 f1, f2; double e1, e2;
        eval_tip(c, s->dummy, s->tip, 0, &f1, &e1);
        v3 p2 = s->tip, v2 = s->v;
        p2.x += dt / 2 * s->v.x; p2.y += dt / 2 * s->v.y; p2.z += dt / 2 * s->v.z;
        v2.x += dt / 2 * f1.x / m; v2.y += dt / 2 * f1.y / m; v2.z += dt / 2 * f1.z / m;
        eval_tip(c, s->dummy, p2, 0, &f2, &e2);
        s->tip.x += dt * v2.x; s->tip.y += dt * v2.y; s->tip.z += dt * v2.z;
        s->v.x += dt * f2.x / m; s->v.y += dt * f2.y / m; s->v.z += dt * f2.z / m;
        s->f = f2; s->e = e2;
    } else {
        v3 f1, f2, f3, f4; double e1, e2, e3, e4;
        eval_tip(c, s->dummy, s->tip, 0, &f1, &e1);
        v3 p2 = s->tip, v2 = s->v, p3 = s->tip, v3_ = s->v, p4 = s->tip, v4 = s->v;
        p2.x += dt / 2 * s->v.x; p2.y += dt / 2 * s->v.y; p2.z += dt / 2 * s->v.z;
        v2.x += dt / 2 * f1.x / m; v2.y += dt / 2 * f1.y / m; v2.z += dt / 2 * f1.z / m;
        eval_tip(c, s->dummy, p2, 0, &f2, &e2);
        p3.x += dt / 2 * v2.x; p3.y += dt / 2 * v2.y; p3.z += dt / 2 * v2.z;
        v3_.x += dt / 2 * f2.x / m; v3_.y += dt / 2 * f2.y / m; v3_.z += dt / 2 * f2.z / m;
        eval_tip(c, s->dummy, p3, 0, &f3, &e3);
        p4.x += dt * v3_.x; p4.y += dt * v3_.y; p4.z += dt * v3_.z;
        v4.x += dt * f3.x / m; v4.y += dt * f3.y / m; v4.z += dt * f3.z / m;
        eval_tip(c, s->dummy, p4, 0, &f4, &e4);
        s->tip.x += dt / 6 * (s->v.x + 2 * v2.x + 2 * v3_.x + v4.x);
        s->tip.y += dt / 6 * (s->v.y + 2 * v2.y + 2 * v3_.y + v4.y);
        s->tip.z += dt / 6 * (s->v.z + 2 * v2.z + 2 * v3_.z + v4.z);
        s->v.x += dt / 6 * (f1.x + 2 * f2.x + 2 * f3.x + f4.x) / m;
        s->v.y += dt / 6 * (f1.y + 2 * f2.y + 2 * f3.y + f4.y) / m;
        s->v.z += dt / 6 * (f1.z + 2 * f2.z + 2 * f3.z + f4.z) / m;
        s->f.x = (f1.x + 2 * f2.x + 2 * f3.x + f4.x) / 6;
        s->f.y = (f1.y + 2 * f2.y + 2 * f3.y + f4.y) / 6;
        s->f.z = (f1.z + 2 * f2.z + 2 * f3.z + f4.z) / 6;
        s->e = (e1 + 2 * e2 + 2 * e3 + e4) / 6;
    }
}

/* FIREMinimisation (minimiser.cpp:67-129). */
static int fire(or_ctx* c, const or_scan_opts* o, or_state* s) {
    const int n_min = 5;
    const double f_inc = 1.1, f_dec = 0.5, f_a = 0.99, a_start = 0.1;
    const double dt_max = 10 * o->dt;
    double alpha = a_start, dt = o->dt;
    int n_non_neg = 0;
    double prev_tip_e;
    int n_tot = 1;
    for (; n_tot < o->maxsteps; ++n_tot) {
        prev_tip_e = s->e;
        step(c, o, s, dt);
        if (converged(o, s->f, s->e - prev_tip_e)) break;
        double p = 0;
        p += s->v.x * s->f.x + s->v.y * s->f.y + s->v.z * s->f.z;
        {
            double l = len3(s->f);
            v3 fh = {0, 0, 0};
            if (l > 1e-10) { fh.x = s->f.x / l; fh.y = s->f.y / l; fh.z = s->f.z / l; }
            double vl = len3(s->v);
            s->v.x = (1 - alpha) * s->v.x + alpha * fh.x * vl;
            s->v.y = (1 - alpha) * s->v.y + alpha * fh.y * vl;
            s->v.z = (1 - alpha) * s->v.z + alpha * fh.z * vl;
        }
        if (p < 0) {
            s->v.x = 0; s->v.y = 0; s->v.z = 0;
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

/* SDMinimisation (minimiser.cpp:39-65). */
static int sd(or_ctx* c, const or_scan_opts* o, or_state* s) {
    double prev_tip_e;
    int n = 1;
    for (; n < o->maxsteps; ++n) {
        prev_tip_e = s->e;
        eval_tip(c, s->dummy, s->tip, 0, &s->f, &s->e);
        if (converged(o, s->f, s->e - prev_tip_e)) break;
        s->tip.x += o->dt * s->f.x; s->tip.y += o->dt * s->f.y; s->tip.z += o->dt * s->f.z;
    }
    return n;
}

/* System::setDummyXY (system.cpp:139-162). */
static void dummy_xy(const or_scan_opts* o, double x, double y, double* px, double* py) {
    if (o->tip_pbc) {
        double inv[9];
        inv3(o->cell, inv);
        v3 rel = {x - o->offset[0], y - o->offset[1], 0 - o->offset[2]};
        v3 q = mul3(inv, rel);
        q.x -= floor(q.x);
        q.y -= floor(q.y);
        v3 p = mul3(o->cell, q);
        *px = p.x + o->offset[0];
        *py = p.y + o->offset[1];
    } else {
        *px = x;
        *py = y;
    }
}

/* The pixel loop of Simulation::run (simulation.cpp:67-143) plus
 * System::getOutput (system.cpp:28-38).  Writes 16 doubles per record at
 * ((i*ny + j)*nz + k)*16: i j k n x y z fx fy fz rx ry rz r angle e.
 * Returns 0, or -1 when a non-periodic grid lookup left the grid. */
int oracle_scan(const or_inter* inter, int n_inter, const or_grid* grids, const double* atoms,
                const or_scan_opts* o, double* rec, int64_t* warnings, int64_t* total_steps) {
    int64_t warn = 0, steps = 0;
    int err = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : warn, steps) reduction(| : err)
    for (int i = 0; i < o->nx; ++i) {
        or_ctx c = {inter, n_inter, grids, atoms, 0, 0};
        double x = i * o->dx;
        for (int j = 0; j < o->ny; ++j) {
            double y = j * o->dy;
            or_state s;
            memset(&s, 0, sizeof s);
            double px, py;
            dummy_xy(o, x, y, &px, &py);
            s.dummy.x = px; s.dummy.y = py; s.tip.x = px; s.tip.y = py;
            s.dummy.z = o->zhigh;
            s.tip.z = o->zhigh - o->tip_dummy_d;
            for (int k = 0; k < o->nz; ++k) {
                int n = o->minimiser == 0 ? sd(&c, o, &s) : fire(&c, o, &s);
                double* r = &rec[(((size_t)i * o->ny + j) * o->nz + k) * 16];
                v3 rv = {s.tip.x - s.dummy.x, s.tip.y - s.dummy.y, s.tip.z - s.dummy.z};
                v3 tf; double te;
                eval_tip(&c, s.dummy, s.tip, 1, &tf, &te);
                r[0] = i; r[1] = j; r[2] = k; r[3] = n;
                r[4] = x; r[5] = y; r[6] = s.dummy.z;
                r[7] = tf.x; r[8] = tf.y; r[9] = tf.z;
                r[10] = rv.x; r[11] = rv.y; r[12] = rv.z;
                r[13] = len3(rv);
                r[14] = atan2(sqrt(sq(rv.x) + sq(rv.y)), rv.z) * (180.0 / 3.14159265358979323846);
                r[15] = te;
                steps += n;
                s.dummy.z -= o->dz;          /* System::lowerTip (system.hpp:51-54) */
                s.tip.z -= o->dz;
            }
        }
        warn += c.warnings;
        err |= c.error;
    }
    if (warnings) *warnings = warn;
    if (total_steps) *total_steps = steps;
    return err ? -1 : 0;
}

/* Force-grid sampling of Simulation::buildTipGridInteractions
 * (simulation.cpp:420-479): tip on the dummy (setTipDummyDistance(0)), every
 * interaction of the (tip-surface) list evaluated at
 * (i*sx + ox, j*sy + oy, k*sz + oz), xy wrapped by setDummyXY when the system
 * carries a periodic cell.  fe receives 4 doubles per point, z fastest. */
int oracle_build_grid(const or_inter* inter, int n_inter, const or_grid* grids, const double* atoms,
                      const or_scan_opts* o, const int* n, const double* spacing, const double* offset,
                      double* fe) {
    int err = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(| : err)
    for (int i = 0; i < n[0]; ++i) {
        or_ctx c = {inter, n_inter, grids, atoms, 0, 0};
        double x = i * spacing[0] + offset[0];
        for (int j = 0; j < n[1]; ++j) {
            double y = j * spacing[1] + offset[1];
            double px, py;
            dummy_xy(o, x, y, &px, &py);
            for (int k = 0; k < n[2]; ++k) {
                double z = k * spacing[2] + offset[2];
                v3 p = {px, py, z};
                v3 f; double e;
                eval_tip(&c, p, p, 0, &f, &e);
                long idx = ((long)i * n[1] + j) * n[2] + k;
                fe[4 * idx] = f.x; fe[4 * idx + 1] = f.y; fe[4 * idx + 2] = f.z; fe[4 * idx + 3] = e;
            }
        }
        err |= c.error;
    }
    return err ? -1 : 0;
}

/* System::evalTipSurfaceForces at arbitrary tip positions (system.cpp:41-51). */
int oracle_eval_tip_surface(const or_inter* inter, int n_inter, const or_grid* grids, const double* atoms,
                            long n, const double* pos, double* out) {
    int err = 0;
#pragma omp parallel for reduction(| : err)
    for (long p = 0; p < n; ++p) {
        or_ctx c = {inter, n_inter, grids, atoms, 0, 0};
        v3 t = {pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]};
        v3 f; double e;
        eval_tip(&c, t, t, 1, &f, &e);
        out[4 * p] = f.x; out[4 * p + 1] = f.y; out[4 * p + 2] = f.z; out[4 * p + 3] = e;
        err |= c.error;
    }
    return err ? -1 : 0;
}

/* Total tip force/energy (all interactions) at given dummy/tip pairs: the
 * quantity the minimisers drive to zero. */
int oracle_eval_total(const or_inter* inter, int n_inter, const or_grid* grids, const double* atoms,
                      long n, const double* dummy, const double* tip, double* out) {
    int err = 0;
#pragma omp parallel for reduction(| : err)
    for (long p = 0; p < n; ++p) {
        or_ctx c = {inter, n_inter, grids, atoms, 0, 0};
        v3 d = {dummy[3 * p], dummy[3 * p + 1], dummy[3 * p + 2]};
        v3 t = {tip[3 * p], tip[3 * p + 1], tip[3 * p + 2]};
        v3 f; double e;
        eval_tip(&c, d, t, 0, &f, &e);
        out[4 * p] = f.x; out[4 * p + 1] = f.y; out[4 * p + 2] = f.z; out[4 * p + 3] = e;
        err |= c.error;
    }
    return err ? -1 : 0;
}
