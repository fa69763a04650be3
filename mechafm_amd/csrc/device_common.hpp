// Device-side data layout shared by the scan, grid and eval kernels.
//
// HBM layout (all per context, uploaded once by mafm_set_systems):
//   DevMol[n_mols]        one record per molecule, read with wave-uniform
//                         indices so the compiler keeps it in SGPRs
//   Site64 sites_ref[]    tip-surface sites in the reference's list order
//                         (VDW(1,i) then Coulomb(1,i), simulation.cpp:297-312)
//   SiteF64 segf64[]      the same sites regrouped per molecule into four
//                         homogeneous segments: LJ+Coulomb, LJ, Morse+Coulomb,
//                         Morse, so the hot loops carry no per-site branch
//   float4 seg32a[], seg32b[]  FP32 copy of segf64 as two arrays, coordinates relative to the
//                         molecule's origin (keeps FP32 rounding of the
//                         tip-atom separation at ~1e-6 A)
//   grids                 double4 / float4 {fx, fy, fz, e} per grid point,
//                         index ix*ny*nz + iy*nz + iz (simulation.cpp:474)
//   DevRecord[P*Z]        one per scan position
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mafm {

struct Site64 {
    double x, y, z;
    double p0, p1, p2;
    double qq;
    int32_t vdw;      // 0 LJ, 1 Morse
    int32_t coul;     // Coulomb term present
};

// Fast-path FP64 site, constants pre-scaled for pair64 (p0 = 6 es6 | 12 De,
// p1 = 12 es12 | a, p2 = - | re, p3 = - | 2 De a).
struct SiteF64 {
    double x, y, z;
    double p0, p1, p2;
    double qq;
    double p3;
};

struct Site32 {
    float4 a;         // x, y, z (relative to DevMol::origin), 6 es6 | 12 De
    float4 b;         // 12 es12 | a log2(e), qq, - | a log2(e) re, - | 2 De a
};

struct DevGrid {
    int32_t n[3];
    int32_t periodic;
    int32_t orthogonal;
    int32_t pad;
    double basis[9];
    double inv[9];
    double offset[3];
    const double4* fe64;
    const float4* fe32;
};

struct DevMol {
    int32_t ref_off, n_sites;              // sites_ref range
    int32_t seg_off;                       // seg64 / seg32 start
    int32_t n_seg[4];                      // LJ+C, LJ, Morse+C, Morse
    int32_t force_grid, potential_grid;    // indices into the grid table or -1
    int32_t tip_pbc;
    int32_t td_vdw, td_coul;
    double td_p0, td_p1, td_p2, td_qq, td_k, td_r0;
    double tip_dummy_d;
    double cell[9], cell_inv[9], offset[3];
    double origin[3];
    int64_t rec_off;                       // first record of this molecule
    int32_t n_images;                      // >= 1; image 0 of a non-periodic system is a zero shift
    int32_t planar;                        // 1: every FP32 image shift has z = 0 (in-plane images)
    double shift[27][3];                   // periodic image shifts, evaluated shift-major
    float shift32[27][4];                  // the same shifts in FP32 (x, y, z, 0)
};

struct DevRecord {
    double x, y, z;
    double fx, fy, fz;
    double rx, ry, rz;
    double e;
    int32_t n_steps;
    int32_t flags;
};

struct ScanArgs {
    int32_t nx, ny, nz;
    int32_t n_mols;
    double dx, dy, dz, zhigh;
    int32_t minimiser, integrator, minterm, maxsteps;
    double etol, ftol, dt;
    const DevMol* mols;
    const Site64* sites_ref;
    const SiteF64* segf64;
    const float4* seg32a;     // FP32 sites: {x, y, z, 6 es6 | 12 De}
    const float4* seg32b;     //             {12 es12 | a log2(e), qq, - | a log2(e) re, - | 2 De a}
    const DevGrid* grids;
    DevRecord* rec;
    // work-stealing state (FAST kernels)
    int32_t* next_col;        // per molecule: next column to hand out
    const float* mol_cum;     // initial share of the workgroups, cumulative over ranks (n_mols + 1)
    const int32_t* mol_order; // molecule of each rank: decreasing work per column
    int64_t* counters;        // [0] total steps [1] unconverged [2] grid warnings [3] out-of-grid
    int32_t wide_tail;        // > 0: regroup a wave's last groups over 64 lanes when that cuts
                              // the sites per lane by this factor; 0: never
    // Decision margins (FAST kernels, diagnostics).  Per minimiser
    // iteration: the force's distance from ftol (|f2 - ftol^2| / (2 ftol)),
    // counted when the energy part of the test passes, the energy
    // difference's distance from etol, counted when the force part passes,
    // and FIRE's power P = v.f over |v|; per position the height's minima.
    float4* margins;          // null: not recorded
    double inv_2ftol;         // 1 / (2 ftol)
    // Streamed scans: records go to pinned host memory, and after a pixel
    // column's record of height k is stored, the column's counter (index
    // molecule * nx * ny + column) becomes k + 1 with a system-scope release,
    // so the host may read the column's first k + 1 records as soon as it
    // sees the count.  Null: no counters.
    int32_t* col_done;
};

// Heights [0, heights) of column `gcol` are in host memory (ScanArgs::col_done).
__device__ inline void column_progress(const ScanArgs& a, long gcol, int heights) {
    __hip_atomic_store(a.col_done + gcol, heights, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

enum { REC_UNCONVERGED = 1, REC_GRID_CLAMP = 2, REC_OUT_OF_GRID = 4 };

// Work-stealing state and counters of one scan in a fixed layout: the
// workspace of the allocation-free scan entry point (mafm_scan_noalloc), kept
// in a file-scope device array of this size (scan_noalloc.hip).
constexpr int NOALLOC_MAX_SYSTEMS = 4096;
struct ScanScratch {
    int64_t counters[8];
    int32_t next_col[NOALLOC_MAX_SYSTEMS];
    float cum[NOALLOC_MAX_SYSTEMS + 1];
    int32_t order[NOALLOC_MAX_SYSTEMS];
};

// Read-only tables are read through the constant address space: with a
// wave-uniform index the compiler then emits scalar loads into SGPRs (a plain
// global pointer would get per-lane vector loads, because the kernels also
// store records and the compiler cannot prove the tables unaliased).
template <typename T>
using cptr = const __attribute__((address_space(4))) T*;
using cptr_mol = cptr<DevMol>;

template <typename T>
__device__ inline cptr<T> as_const(const T* p) {
    return (cptr<T>)(p);
}

// Matrix3::multiply for a row-major 3x3 (matrices.hpp:146-153), every product
// and sum rounded on its own as in the reference.
template <typename P>
__device__ __host__ inline void mat_vec(P m, double x, double y, double z, double& ox, double& oy, double& oz) {
#pragma clang fp contract(off)
    ox = m[0] * x + m[1] * y + m[2] * z;
    oy = m[3] * x + m[4] * y + m[5] * z;
    oz = m[6] * x + m[7] * y + m[8] * z;
}

// System::setDummyXY's wrap of a pixel into the unit cell (system.cpp:139-162):
// cell-frame coordinates of (x, y, 0) - offset, the fractional parts of the
// first two, back to Cartesian.  Each product and sum is rounded on its own
// (no fused multiply-add), as the reference computes it, so a pixel on a cell
// face wraps the same way in every kernel and in host code that restates it.
template <typename P>
__device__ __host__ inline void wrap_xy(P cell, P inv, P off, double x, double y, double& px, double& py) {
#pragma clang fp contract(off)
    double qx, qy, qz, ox, oy, oz;
    mat_vec(inv, x - off[0], y - off[1], 0 - off[2], qx, qy, qz);
    qx -= floor(qx);
    qy -= floor(qy);
    mat_vec(cell, qx, qy, qz, ox, oy, oz);
    px = ox + off[0];
    py = oy + off[1];
}

}  // namespace mafm
