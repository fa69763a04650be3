// Device arguments of the flexible scan (scan_flex.hip), shared with the
// host side that fills them (api.cpp).
#pragma once
#include "../../include/mafm.h"
#include "device_common.hpp"

namespace mafm {

// Every term is evaluated once per integrator stage into result slots, one
// slot {fx, fy, fz} per atom the term acts on, holding the signed force that
// atom receives (the reference's `forces[i] -= f` is `+= -f`, the same
// rounding), plus the term's energy.  Each free atom then sums its slots in
// the interaction list's order; the tip sums the energies of its terms.
struct FlexArgs {
    int32_t n_atoms, n_eval, n_ts;    // eval entries [0, n_ts) are the tip-surface terms
    int32_t n_slots, ts_slots;        // slots [0, ts_slots) belong to the tip-surface terms
    int32_t nx, ny, nz;
    double dx, dy, dz, zhigh;
    int32_t minimiser, integrator, minterm, maxsteps;
    double etol, ftol, dt;
    // the terms that act on a free atom, in list order, each with its first
    // result slot in `reserved`
    const mafm_flex_term* terms;
    // index tables, back to back: csr_off (n_atoms + 1, empty ranges for fixed
    // atoms), csr (n_csr: the result slots of each free atom in list order),
    // tip_eval (n_tip: the terms acting on the tip, in list order), fixed (n_atoms)
    const int32_t* idx;
    int32_t n_csr, n_tip;
    int32_t tip_ts_slots, tip_ts_terms;   // the tip's slots / terms from tip-surface terms
    // which terms each thread evaluates in one integrator stage: thread l of
    // the block takes sched[r * block threads + l] (-1: none) in rounds
    // r < sched_rounds; NULL: terms l, l + T, ...
    // The schedule only places the work: every term still writes its own
    // result slots, which are summed in list order.
    const int32_t* sched;
    int32_t sched_rounds;
    const double* pos0;               // n_atoms x 3
    int32_t tip_pbc;
    int32_t state_col;                // column whose positions are kept per height, -1 none
    double cell[9], cell_inv[9], offset[3];
    double tip_dummy_d;
    DevRecord* rec;
    int32_t* next_col;                // work counter, zero at launch
    double* state;                    // nz x n_atoms x 3
    double* work;                     // per-block state when it does not fit LDS
};

// The reference's System-level operators on n_states independent copies of
// one system (one workgroup each): the whole interaction list evaluated into
// every atom's force and energy, one integrator step, or a whole FIRE / SD
// minimisation from the given state.
enum { SYS_EVAL = 0, SYS_STEP = 1, SYS_MINIMISE = 2 };
struct SysArgs {
    int32_t n_atoms, n_terms, n_slots, op;
    int32_t minimiser, integrator, minterm, maxsteps;
    double etol, ftol, dt;
    int64_t n_states;
    // every term, in list order, its first result slot in `reserved`
    const mafm_flex_term* terms;
    // back to back: f_off (n_atoms + 1), f_csr (n_fcsr: each atom's result
    // slots in list order), e_off (n_atoms + 1), e_csr (n_ecsr: each atom's
    // terms in list order, for its energy), fixed (n_atoms)
    const int32_t* idx;
    int32_t n_fcsr, n_ecsr;
    double* pos;       // n_states x n_atoms x 3, in and out
    double* vel;       // n_states x n_atoms x 3, in and out (step, minimise)
    double* frc;       // n_states x n_atoms x 3: in (eval and the Euler step add to it), out
    double* ener;      // n_states x n_atoms: as frc; minimise reads energies[1] as the last tip energy
    int32_t* steps;    // n_states (minimise)
    double* work;      // per-block state when it does not fit LDS
};
size_t sys_block_bytes(const SysArgs& a);
hipError_t launch_system_op(const SysArgs& a, int waves, bool in_lds, hipStream_t st);

// Doubles of per-column state one block keeps (positions, integrator
// stages, velocities, forces, term results).
size_t flex_block_doubles(int n_atoms, int n_slots, int n_eval, bool rk4);
// LDS bytes of one block when the state lives there (state + index tables).
size_t flex_block_lds(const FlexArgs& a);
// waves: wavefronts per workgroup (1-4); in_lds: the state lives in LDS (else
// in a.work, one slice per block).
hipError_t launch_scan_flex(const FlexArgs& a, int n_blocks, int waves, bool in_lds, hipStream_t st);

}  // namespace mafm
