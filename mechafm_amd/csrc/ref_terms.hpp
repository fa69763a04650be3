// The reference's pair-term arithmetic in its own operation order, for the
// kernels built with -ffp-contract=off (plain scan, flexible scan): every
// operation rounds as the reference's -O3 x86-64 build rounds it, and the
// divisions and square roots are IEEE.
#pragma once

#include "device_common.hpp"

namespace mafm {

struct V3 {
    double x, y, z;
};

__device__ inline double sq(double a) { return a * a; }   // pow(a, 2): the exact square
__device__ inline double len3(V3 a) { return sqrt((sq(a.x) + sq(a.y)) + sq(a.z)); }   // vectors.hpp:99-101

// LJInteraction / MorseInteraction::eval (interactions.cpp:15-40) with
// r = p1 - p2, returning the force on atom i1 and the pair energy.
__device__ inline void vdw(int kind, double p0, double p1, double p2, V3 r, V3& f, double& e) {
    if (kind == 0) {
        double r_sqr = (sq(r.x) + sq(r.y)) + sq(r.z);
        double r6 = r_sqr * r_sqr * r_sqr;
        double term_a = p1 / (r6 * r6);
        double term_b = p0 / r6;
        e = term_a - term_b;
        double s = (12 * term_a - 6 * term_b) / r_sqr;
        f.x = s * r.x; f.y = s * r.y; f.z = s * r.z;
    } else {
        double rl = len3(r);
        double d_exp = exp(-p1 * (rl - p2));
        e = p0 * (sq(d_exp) - 2 * d_exp + 1);
        double f_abs = 2 * p0 * p1 * (sq(d_exp) - d_exp);
        double s = f_abs / rl;
        f.x = s * r.x; f.y = s * r.y; f.z = s * r.z;
    }
}

// CoulombInteraction::eval (interactions.cpp:42-51).
__device__ inline void coulomb(double qq, V3 r, V3& f, double& e) {
    double rl = len3(r);
    e = qq / rl;
    double s = qq / (rl * rl * rl);
    f.x = s * r.x; f.y = s * r.y; f.z = s * r.z;
}

}  // namespace mafm
