/*
 * mafm.h -- C ABI of libmafm, the MI355X (gfx950) probe-particle scan (rigid
 * and flexible surfaces).
 *
 * Plain C types only: host pointers are borrowed for the duration of a call,
 * every entry point returns mafm_status (no exit(), no exceptions cross the
 * ABI), one context per host thread, calls on one context are serialised on
 * the context's own HIP stream.
 *
 * Reference interfaces each entry point replaces (paths relative to the
 * MechAFM source tree):
 *   mafm_set_systems       the tip-surface / tip-dummy part of the Interaction
 *                          list built by Simulation::buildInteractions
 *                          (simulation.cpp:201-216, 238-418, 503-515) and held
 *                          as vector<unique_ptr<Interaction>> (simulation.hpp:76)
 *   mafm_scan              the pixel loop of Simulation::run
 *                          (simulation.cpp:78-140) with FIREMinimisation /
 *                          SDMinimisation (minimiser.cpp:39-129),
 *                          euler/midpoint/rk4Step (integrators.cpp:8-111) and
 *                          System::getOutput (system.cpp:28-51)
 *   mafm_records           the OutputData buffer drained by writeOutput
 *                          (simulation.cpp:145-199; globals.hpp:27-32)
 *   mafm_scan_host         mafm_scan with the records handed to the host as
 *                          the scan runs, as the reference flushes its buffer
 *                          to the writer inside the scan (simulation.cpp:123-138)
 *   mafm_eval_tip_surface  Interaction::eval over the tip-surface list
 *                          (interactions.hpp:68-76, interactions.cpp:15-51,
 *                          236-250) as System::evalTipSurfaceForces calls it
 *   mafm_grid_build        the grid sampling loop of buildTipGridInteractions
 *                          (simulation.cpp:420-479)
 *   mafm_grid_upload /     ForceGrid set-up and swapForceValues /
 *   mafm_grid_download     swapEnergyValues (force_grid.hpp:20-30)
 *   mafm_grid_interpolate  ForceGrid::interpolate (force_grid.hpp:33)
 *   mafm_potential_grid    ElectrostaticPotentialInteraction's constructor
 *                          (interactions.cpp:53-234)
 *   mafm_fft3d             fft_data_grid / ffti_data_grid (fft.hpp:25, 34)
 *   mafm_scan_flexible     Simulation::run with flexible on (simulation.cpp:67-143):
 *                          the whole interaction list of buildInteractions
 *                          (simulation.cpp:201-216, 503-693) relaxed per
 *                          position; mafm_flex_state keeps what
 *                          System::makeXYZFile writes (system.cpp:54-65)
 *   mafm_system_eval /     Interaction::eval over a System's list, eulerStep /
 *   mafm_system_step /     midpointStep / rk4Step and FIREMinimisation /
 *   mafm_system_minimise   SDMinimisation (interactions.hpp:68-76,
 *                          integrators.hpp:13-17, minimiser.hpp:20-23) on many
 *                          System states at once
 *   mafm_adjoint           no counterpart (gradients of an image loss)
 *   mafm_deltaf_weights /  the frequency-shift post-processing of
 *   mafm_delta_f           simple-vafm-image.py:76-204 (Giessibl weighting of
 *                          each pixel's Fz(z) over the oscillation window),
 *                          applied to the scan's records on the device
 */
#ifndef MAFM_H
#define MAFM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    MAFM_OK = 0,
    MAFM_ERR_INVALID = 1,      /* bad argument / descriptor */
    MAFM_ERR_HIP = 2,          /* a HIP runtime call failed */
    MAFM_ERR_OUT_OF_GRID = 3,  /* non-periodic grid lookup needed a point past the grid
                                  (force_grid.cpp:236-240 calls error() there) */
    MAFM_ERR_STATE = 4,        /* call order: e.g. scan before set_systems */
    MAFM_ERR_FFT_IMAG = 5,     /* inverse FFT imaginary part above 1e-8 (fft.cpp:79-83) */
    MAFM_ERR_NO_DEVICE = 6
} mafm_status;

typedef struct mafm_ctx mafm_ctx;

/* ---- interaction tables ------------------------------------------------ */
enum { MAFM_VDW_LJ = 0, MAFM_VDW_MORSE = 1 };

/* One tip-surface site: the VDW term (LJInteraction / MorseInteraction,
 * interactions.hpp:79-119) against one surface atom (or one periodic image of
 * it, simulation.cpp:298-312) plus the optional Coulomb term on the same atom
 * (CoulombInteraction, interactions.hpp:121-137).  Sites are given in the
 * reference's list order: VDW(1,i) then Coulomb(1,i), atom by atom. */
typedef struct {
    double x, y, z;    /* atom position + pbc shift, Angstrom */
    int32_t vdw;       /* MAFM_VDW_LJ or MAFM_VDW_MORSE */
    int32_t has_coulomb;
    double p0, p1, p2; /* LJ: es6, es12, unused | Morse: De, a, re */
    double qq;         /* qbase*q_tip*q_atom when has_coulomb */
} mafm_site;

/* Tip-dummy terms (simulation.cpp:503-515): VDW(0,1), optional Coulomb(0,1)
 * and the lateral TipHarmonicInteraction (interactions.cpp:252-266). */
typedef struct {
    int32_t vdw;
    int32_t has_coulomb;
    double p0, p1, p2;
    double qq;
    double k, r0;
} mafm_tip_dummy;

/* One molecule / surface to scan. */
typedef struct {
    int32_t n_sites;
    int32_t force_grid;     /* handle of a non-periodic force grid replacing the
                               sites (rigidgrid on), or -1 */
    const mafm_site* sites;
    int32_t potential_grid; /* handle of a periodic external-potential grid
                               evaluated after the sites, or -1 */
    int32_t tip_pbc;        /* System::setDummyXY wraps the dummy into the cell */
    mafm_tip_dummy tip_dummy;
    double tip_dummy_d;     /* initial tip-dummy distance (simulation.cpp:218-225) */
    double cell[9];         /* row-major cell matrix (columns = cell vectors) */
    double offset[3];       /* System offset (system.cpp:103-136) */
    /* Periodic images (vdw_pbc, simulation.cpp:298-312): when n_images > 0
     * the site table is evaluated once per shift, shift-major, each site at
     * its position + shift (the reference's order).  0 = no images. */
    int32_t n_images;       /* 0..27 */
    int32_t reserved;
    const double* image_shifts;   /* n_images x 3 */
} mafm_system_desc;

/* ---- grids ------------------------------------------------------------- */
typedef struct {
    int32_t n[3];
    int32_t periodic;
    double basis[9];        /* row-major, columns = grid step vectors */
    double offset[3];       /* position of grid point (0,0,0) */
    const double* fe;       /* n0*n1*n2 x 4 doubles (fx fy fz e), z fastest;
                               NULL for mafm_grid_build */
} mafm_grid_desc;

/* ---- scan -------------------------------------------------------------- */
enum { MAFM_MIN_SD = 0, MAFM_MIN_FIRE = 1 };
enum { MAFM_INT_EULER = 0, MAFM_INT_MIDPOINT = 1, MAFM_INT_RK4 = 2 };
enum { MAFM_TERM_E = 0, MAFM_TERM_F = 1, MAFM_TERM_EF = 2 };
/* Arithmetic of the pair terms.  PLAIN64: one lane per pixel, IEEE divisions,
 * operation order of the reference (the bit-parity path and the "plain FP64"
 * baseline).  FAST64: work-stealing persistent kernel, FP64 with refined
 * reciprocals.  FAST32: the same kernel with FP32 pair terms and compensated
 * accumulation; minimiser state and convergence tests stay FP64. */
enum { MAFM_PREC_PLAIN64 = 0, MAFM_PREC_FAST64 = 1, MAFM_PREC_FAST32 = 2 };

/* mafm_scan_desc::options bits (FAST kernels).
 * REPRODUCIBLE: every position's site sums follow one fixed order, so two
 * runs with the same tuning give identical records.  Without it, a wave
 * whose pixel groups have mostly finished evaluates its remaining groups
 * one after another with all 64 lanes (shorter tail, same arithmetic, but
 * the summation order of a position then depends on timing: results agree
 * to FP32 rounding, not bit for bit). */
/* MARGINS: keep every position's decision margins for mafm_margins: how
 * near the height's minimiser decisions -- the convergence test
 * (minimiser.cpp:15-37) and FIRE's velocity reset on P = v.f < 0
 * (minimiser.cpp:107-126) -- came to coming out the other way.  A decision
 * taken on FP32 pair sums flips against the FP64 trajectory when its
 * margin is below the FP32 error of the tested quantity; the margins show
 * where. */
/* STATE_IN_HBM (mafm_scan_flexible): keep each column's state in HBM even
 * where it fits one workgroup's LDS -- the path systems too large for LDS
 * take; leaves the LDS to other work. */
enum { MAFM_SCAN_REPRODUCIBLE = 1, MAFM_SCAN_MARGINS = 4, MAFM_SCAN_STATE_IN_HBM = 8 };

typedef struct {
    int32_t nx, ny, nz;
    double dx, dy, dz;
    double zhigh;
    int32_t minimiser, integrator, minterm;
    double etol, ftol, dt;
    int32_t maxsteps;
    int32_t precision;
    /* tuning, 0 = automatic */
    int32_t lanes_per_pixel;   /* 1,2,4,...,64 lanes share one pixel column */
    int32_t waves_per_simd;    /* persistent grid size */
    int32_t compensation;      /* FAST32 sums: 0 = 3 = FP32 blocks of 4-6 sites folded into FP64
                                  (default), 1 = Kahan-compensated FP32, 2 = plain FP32 */
    int32_t waves_per_group;   /* 1 .. 8 waves per workgroup sharing one LDS site copy */
    uint32_t options;          /* MAFM_SCAN_* bits */
    int32_t reserved;
} mafm_scan_desc;

/* Per scan position (OutputData, globals.hpp:27-32, minus the derived r and
 * angle which the writer computes from r). */
typedef struct {
    double x, y, z;            /* requested pixel x, y; dummy z */
    double fx, fy, fz;         /* tip-surface force at the relaxed tip */
    double rx, ry, rz;         /* tip - dummy */
    double e;                  /* tip-surface energy */
    int32_t n_steps;           /* minimiser iterations (maxsteps if not converged) */
    int32_t flags;             /* MAFM_REC_* */
} mafm_record;

enum { MAFM_REC_UNCONVERGED = 1, MAFM_REC_GRID_CLAMP = 2, MAFM_REC_OUT_OF_GRID = 4 };

typedef struct {
    double kernel_ms;          /* device time of the scan kernel(s), HIP events */
    int64_t positions;
    int64_t total_steps;
    int64_t unconverged;
    int64_t grid_warnings;
    int64_t pair_evals;        /* sites x evaluations (algorithmic work) */
    int32_t lanes_per_pixel;
    int32_t waves_per_simd;
    int32_t grid_blocks;
    int32_t block_threads;
    int64_t group_trips;        /* FAST kernels: pixel-group slots x loop trips */
    int64_t active_group_trips; /* ... of which the group had a column (utilisation) */
    int64_t wide_trips;         /* wave trips that evaluated their groups with all 64 lanes */
} mafm_scan_stats;

typedef struct {
    char name[64];
    int32_t compute_units;
    int32_t clock_mhz;
    int64_t total_mem;
    int32_t lds_per_block;
    int32_t warp_size;
} mafm_device_props;

/* External electrostatic potential on a periodic grid (the cube data after the
 * unit scaling of simulation.cpp:348-355 and the origin shift of :365). */
typedef struct {
    int32_t n[3];
    double basis[9];          /* row-major, columns = voxel vectors, Angstrom */
    double origin[3];
    const double* potential;  /* n0*n1*n2 values, z fastest */
    double tip_charge;        /* charge of the tip atom (parameter file) */
    double gaussian_width;    /* width of the tip's Gaussian charge, 0.5 A in the reference */
} mafm_potential_desc;

/* Adjoint of the scan (no reference counterpart: MechAFM has no gradients).
 * For an image loss L(Fz) with g = dL/dFz at every relaxed position, the
 * implicit function theorem at each relaxed tip (total tip force = 0) gives
 *   dL/dtheta = sum_pos v . dF_ts/dtheta,  v = g (e_z - w),  J w = grad_p Fz_ts,
 * J the 3x3 Jacobian of the total tip force.  Returned per site of the
 * system's table (LJ + Coulomb sites; Morse sites and grids are rejected). */
typedef struct {
    int64_t n_positions;
    const double* tip;        /* (n, 3) relaxed tip positions */
    const double* dummy;      /* (n, 3) dummy positions */
    const double* g;          /* (n) dL/dFz */
    int32_t precision;        /* MAFM_PREC_FAST32 or MAFM_PREC_FAST64 pair arithmetic */
    int32_t reserved;
} mafm_adjoint_desc;

typedef struct {
    double* d_es6;            /* (n_sites) dL/d es6 */
    double* d_es12;           /* (n_sites) dL/d es12 */
    double* d_qq;             /* (n_sites) dL/d qq (Coulomb prefactor of the site) */
    double* d_pos;            /* (n_sites, 3) dL/d site position */
} mafm_adjoint_out;

/* ---- flexible surface (flexible on) ------------------------------------ */
/* With flexible on every free surface atom relaxes with the tip: the
 * interaction list of Simulation::buildInteractions (simulation.cpp:201-216)
 * then holds, in this order, the tip-surface terms (simulation.cpp:297-318),
 * the tip-dummy terms (:503-515), the bonds, angles, dihedrals and
 * non-bonded pairs of buildSurfaceSurfaceInteractions (:557-693) and the
 * substrate wall and xy tethers of buildSubstrateInteractions (:517-531).
 * mafm_flex_term is one Interaction of that list; the forces on each atom
 * are summed in list order, as Interaction::eval adds them.  Atom 0 is the
 * dummy, 1 the tip. */
enum {
    MAFM_FX_LJ = 0,            /* LJInteraction: atom[0..1], p = es6, es12; shift added to atom[1] */
    MAFM_FX_MORSE = 1,         /* MorseInteraction: atom[0..1], p = De, a, re; shift */
    MAFM_FX_COULOMB = 2,       /* CoulombInteraction: atom[0..1], p = qq */
    MAFM_FX_TIP_HARMONIC = 3,  /* TipHarmonicInteraction (lateral): atom[0..1], p = k, r0 */
    MAFM_FX_BOND = 4,          /* HarmonicInteraction: atom[0..1], p = k, r0 */
    MAFM_FX_ANGLE = 5,         /* HarmonicAngleInteraction: atom = shared, i1, i2, p = k, theta0 */
    MAFM_FX_DIHEDRAL = 6,      /* HarmonicDihedralInteraction: atom[0..3], p = k, sigma0 */
    MAFM_FX_SUBSTRATE = 7,     /* SubstrateInteraction: atom[0], p = multiplier, ulj, ushift, z0, rc, sigma
                                  (the constructor's derived constants, interactions.hpp:222-229) */
    MAFM_FX_XY_HARMONIC = 8    /* XYHarmonicInteraction: atom[0], p = k, x0, y0 */
};

typedef struct {
    int32_t kind;              /* MAFM_FX_* */
    int32_t atom[4];           /* atom indices; unused slots ignored */
    int32_t reserved;
    double p[6];
    double shift[3];           /* LJ / Morse periodic image shift */
} mafm_flex_term;

typedef struct {
    int32_t n_atoms;           /* dummy + tip + surface atoms */
    int32_t n_terms;
    const double* positions;   /* n_atoms x 3 after centring; rows 0 and 1 are set per column */
    const int32_t* fixed;      /* n_atoms: 1 = never moves (fixed[0] must be 1, fixed[1] 0) */
    const mafm_flex_term* terms;   /* the list, tip-surface terms (atom[0] = 1) first */
    int32_t tip_pbc;           /* System::setDummyXY wraps the dummy into the cell */
    int32_t state_column;      /* column i*ny + j whose relaxed atoms are kept at every height
                                  (makeXYZFile, simulation.cpp:114-116), or -1 */
    double tip_dummy_d;
    double cell[9];            /* row-major, columns = cell vectors */
    double offset[3];
} mafm_flex_desc;

/* ---- entry points ------------------------------------------------------ */
mafm_status mafm_create(int device, mafm_ctx** out);
/* Frees the context and its device memory. Simulations opened on it
 * (mafm_sim_open) keep it alive: while any is open, the destruction is
 * completed by the last mafm_sim_close. */
void mafm_destroy(mafm_ctx* ctx);
const char* mafm_last_error(const mafm_ctx* ctx);
mafm_status mafm_device_info(mafm_ctx* ctx, mafm_device_props* props);

mafm_status mafm_grid_upload(mafm_ctx* ctx, const mafm_grid_desc* grid, int* handle);
mafm_status mafm_grid_build(mafm_ctx* ctx, const mafm_system_desc* sys, const mafm_grid_desc* grid,
                            int* handle);
mafm_status mafm_grid_download(mafm_ctx* ctx, int handle, double* fe);
mafm_status mafm_grid_free(mafm_ctx* ctx, int handle);
/* ForceGrid::interpolate (force_grid.hpp:33, force_grid.cpp:88-155) of grid
 * `handle` at n positions (n x 3): fe = n x 4 (fx fy fz e), the reference's
 * trilinear weights and operation order.  MAFM_ERR_OUT_OF_GRID (values
 * still written, clamped) when a position needs a point past a non-periodic
 * grid, where the reference stops (force_grid.cpp:236-240). */
mafm_status mafm_grid_interpolate(mafm_ctx* ctx, int handle, size_t n, const double* positions, double* fe);
/* Energy and force of the tip's Gaussian charge in the potential, by FFT
 * convolution and spectral gradient, stored as a periodic grid. */
mafm_status mafm_potential_grid(mafm_ctx* ctx, const mafm_potential_desc* pot, int* handle);
/* 3D complex DFT, unnormalised forward (e^{-2 pi i jk/n}) or inverse scaled
 * by 1/(n0 n1 n2); interleaved re/im doubles, z fastest. */
mafm_status mafm_fft3d(mafm_ctx* ctx, const int32_t n[3], int inverse, const double* in, double* out);

mafm_status mafm_set_systems(mafm_ctx* ctx, int n_systems, const mafm_system_desc* systems);
mafm_status mafm_scan(mafm_ctx* ctx, const mafm_scan_desc* desc, mafm_scan_stats* stats);
/* mafm_scan without any device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory during a scan
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The systems come from an earlier mafm_set_systems (set-up
 * allocates, this call does not); the records go to the caller's device
 * buffer `records` of `capacity` records, in mafm_records' order; the
 * work-stealing state and counters use a fixed device workspace of the
 * library (one per GPU, about 48 KiB, kept until the process ends), so at
 * most 4096 systems per call; MAFM_SCAN_MARGINS is refused.  Calls on one
 * context are serialised on its stream; calls from two contexts on one GPU
 * that overlap in time race on the shared workspace.  Afterwards the
 * context holds no records of its own (mafm_records and mafm_delta_f
 * report MAFM_ERR_STATE / an empty range until the next mafm_scan). */
mafm_status mafm_scan_noalloc(mafm_ctx* ctx, const mafm_scan_desc* desc, mafm_record* records, size_t capacity,
                              mafm_scan_stats* stats);
/* Streamed scan: the records the reference's scan loop hands to its writer
 * while it runs (simulation.cpp:123-138, flushed every bufsize columns).
 * Launches mafm_scan's kernels and returns at once; the kernels store the
 * records straight into pinned host memory owned by the context
 * (*records, in mafm_records' order), and column_done[system*nx*ny + i*ny
 * + j] counts the heights of that column whose records are there (nz once
 * the column is finished; heights are stored from the top, k = 0 first), so
 * the caller may read records as soon as the count covers them (read the
 * count with acquire ordering).  mafm_scan_host_query reports whether the launch
 * has finished without waiting; mafm_scan_host_wait ends it and fills the
 * statistics (MAFM_ERR_OUT_OF_GRID as mafm_scan).  Until then every other
 * scan on the context (mafm_scan_flexible included) and mafm_set_systems are
 * refused with MAFM_ERR_STATE; other calls may run and wait behind the
 * launch, the scan's statistics are not affected.  The host memory stays
 * valid until the next streamed scan or mafm_destroy; the context holds no
 * device records afterwards (mafm_records / mafm_delta_f refuse).
 * MAFM_SCAN_MARGINS is refused. */
mafm_status mafm_scan_host(mafm_ctx* ctx, const mafm_scan_desc* desc, const mafm_record** records,
                           const int32_t** column_done);
mafm_status mafm_scan_host_query(mafm_ctx* ctx, int* finished);
mafm_status mafm_scan_host_wait(mafm_ctx* ctx, mafm_scan_stats* stats);
/* records of all systems: index ((system*nx + i)*ny + j)*nz + k */
mafm_status mafm_records(mafm_ctx* ctx, size_t first, size_t count, mafm_record* out);
/* Decision margins of the last scan (FAST kernels run with
 * MAFM_SCAN_MARGINS), 4 floats per position in record order: the smallest
 * force margin (|F^2 - ftol^2| / (2 ftol), counted where the energy part of
 * the test passed), energy margin (||E - E_prev| - etol|, counted where the
 * force part passed) and power margin (|P| / |v|) of the height's minimiser
 * iterations (+inf where none was counted), then 0.  No reference
 * counterpart: a diagnostic of the FP32 arithmetic. */
mafm_status mafm_margins(mafm_ctx* ctx, size_t first, size_t count, float* out);
/* VALU issue-rate calibration.  No reference counterpart: a measurement aid
 * for the device's performance counters.  One launch of 32 768 wave64s, each
 * lane running 8 independent chains of `iterations` steps: kind 0 FP32 FMAs,
 * 1 FP64 FMAs, 2 FP64 v_rsq_f64 each followed by an addition.  *ms is the
 * launch's time (HIP events), *wave_instructions the VALU wave instructions
 * of its chains (known exactly, to set against SQ_INSTS_VALU). */
mafm_status mafm_valu_calibration(mafm_ctx* ctx, int kind, int iterations, double* ms, int64_t* wave_instructions);
mafm_status mafm_eval_tip_surface(mafm_ctx* ctx, int system, size_t n, const double* pos, double* fe);
/* The scan with a flexible surface: the pixel loop of Simulation::run
 * (simulation.cpp:78-140) over one System copy per column, whose surface
 * atoms stay relaxed from height to height, with FIREMinimisation /
 * SDMinimisation and the integrators over the whole atom list
 * (minimiser.cpp:39-129, integrators.cpp:8-111).  FP64 in the reference's
 * operation order (desc->precision is not used).  The records replace the
 * context's, in mafm_records' order for one system; the systems of
 * mafm_set_systems are kept. */
mafm_status mafm_scan_flexible(mafm_ctx* ctx, const mafm_flex_desc* sys, const mafm_scan_desc* desc,
                               mafm_scan_stats* stats);
/* Relaxed atom positions of the state column of the last flexible scan,
 * n_atoms x 3 doubles per height for heights [first, first + count). */
mafm_status mafm_flex_state(mafm_ctx* ctx, int32_t first, int32_t count, double* out);

/* ---- System-level operators --------------------------------------------
 * The reference's per-System operators, on the device, for n_states
 * independent copies of one System at once (one workgroup each).  A System
 * is its interaction list (Simulation::interactions_, in list order, every
 * term kept -- also those between fixed atoms) and its fixed flags
 * (System::fixed_, 1 = never moves); the state arrays are n_states x
 * n_atoms x 3 doubles (positions_, velocities_, forces_) and n_states x
 * n_atoms (energies_), read and written in place.  Atom 1 is the tip, whose
 * force and energy the convergence tests read.  Masses are 1, as
 * System::initialize sets them (system.cpp:18). */
typedef struct {
    int32_t n_atoms;
    int32_t n_terms;
    const mafm_flex_term* terms;   /* reserved is ignored */
    const int32_t* fixed;          /* n_atoms */
} mafm_interaction_list;

typedef struct {                   /* the InputOptions fields minimiser.cpp reads */
    int32_t minimiser;             /* MAFM_MIN_* */
    int32_t integrator;            /* MAFM_INT_* (FIRE) */
    int32_t minterm;               /* MAFM_TERM_* */
    int32_t maxsteps;
    double etol, ftol, dt;
} mafm_minimiser_opts;

/* Interaction::eval of every term in list order (interactions.hpp:68-76,
 * interactions.cpp:15-362): each term's force and energy are added onto
 * forces / energies as they are given (zero them first for the plain
 * values), as System code calling eval over interactions_ does. */
mafm_status mafm_system_eval(mafm_ctx* ctx, const mafm_interaction_list* list, int64_t n_states,
                             const double* positions, double* forces, double* energies);
/* eulerStep / midpointStep / rk4Step (integrators.hpp:13-17,
 * integrators.cpp:8-111): Euler adds the list's forces and energies onto
 * forces / energies as given (FIREMinimisation zeroes them first); midpoint
 * and RK4 overwrite them with the last stage's / the averaged values. */
mafm_status mafm_system_step(mafm_ctx* ctx, const mafm_interaction_list* list, int32_t integrator, double dt,
                             int64_t n_states, double* positions, double* velocities, double* forces,
                             double* energies);
/* FIREMinimisation / SDMinimisation (minimiser.hpp:20-23,
 * minimiser.cpp:39-129) from each given state: velocities carry in and out,
 * energies[1] given is the tip energy the first convergence test compares
 * with; steps[s] is the function's return value. */
mafm_status mafm_system_minimise(mafm_ctx* ctx, const mafm_interaction_list* list, const mafm_minimiser_opts* opts,
                                 int64_t n_states, double* positions, double* velocities, double* forces,
                                 double* energies, int32_t* steps);
mafm_status mafm_adjoint(mafm_ctx* ctx, int system, const mafm_adjoint_desc* desc, mafm_adjoint_out* out);

/* ---- frequency shift (simple-vafm-image.py:76-204) ----
 * Forces are taken in kcal/mol/A, as the script assumes (its spring constant
 * is converted with 1 N/m = 1.438978 kcal/mol/A^2). Zero fields take the
 * script's values. */
typedef struct {
    double amplitude;      /* oscillation amplitude, A (0.50, :81) */
    double k_cantilever;   /* cantilever spring constant, N/m (1800, :82) */
    double f0;             /* resonance frequency, Hz (25e3, :83) */
    int32_t fz_from_text;  /* 1: Fz as the script reads it back from the per-height
                              files (:147, :183): printed "%8.4f" (the exact value
                              rounded to 4 decimals, ties to even) and parsed by
                              float(); 0: the records' Fz as they are */
    int32_t reserved;
} mafm_deltaf_desc;

typedef struct {
    int32_t n_weights;     /* heights per image (window length) */
    int32_t n_images;      /* nz - n_weights: images p = 0 .. n_images-1, ascending height */
    double dz;             /* height step as the script derives it from the file names */
    double amplitude;      /* amplitude after the script's half-step adjustment (:87-88) */
    double prefactor;      /* f0 / (2k) * 2 / (pi A^2) with the unadjusted A (:86) */
    double weights[128];   /* W[t], t = 0 .. n_weights-1, lowest height first (:89-92) */
} mafm_deltaf_weights_out;

/* The window and constants of a scan with nz heights from zhigh down in
 * steps of dz (host only; ctx may be NULL and only receives the message). */
mafm_status mafm_deltaf_weights(mafm_ctx* ctx, const mafm_deltaf_desc* desc, int32_t nz, double zhigh, double dz,
                                mafm_deltaf_weights_out* out);
/* Delta f images (Hz) of `system` from the records of the last mafm_scan:
 * df[(p * nx + i) * ny + j] = -prefactor * sum_t Fz(i, j, height p + t) W[t] * dz
 * for p < n_images (mafm_deltaf_weights gives n_images), and z_low[p] = the
 * lowest height of image p's window (the height the script names it by).
 * z_low may be NULL. */
mafm_status mafm_delta_f(mafm_ctx* ctx, int system, const mafm_deltaf_desc* desc, double* df, double* z_low);

#ifdef __cplusplus
}
#endif
#endif
