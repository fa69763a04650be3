/*
 * mafm_system.hpp -- the reference's System-level operator API in C++, over
 * the C ABI of mafm.h (header only; link libmafm).
 *
 * Same names and argument meaning as MechAFM's:
 *   System                 the state vectors of system.hpp:57-66 (positions_,
 *                          velocities_, forces_, energies_, fixed_) and the
 *                          interaction list it points to, as device terms
 *   FIREMinimisation /     minimiser.hpp:20-23 (minimiser.cpp:39-129): return
 *   SDMinimisation         the iteration count, leave the relaxed state
 *   eulerStep /            integrators.hpp:13-17 (integrators.cpp:8-111)
 *   midpointStep / rk4Step
 *   evalInteractions       `for (interaction : *interactions_) interaction->
 *                          eval(positions_, forces_, energies_)`
 *                          (interactions.hpp:68-76)
 * The arithmetic runs on the GPU of the System's context.  Where the
 * reference calls error() and exits, these throw std::runtime_error with the
 * library's message.  For many states at once call mafm_system_* directly.
 */
#ifndef MAFM_SYSTEM_HPP
#define MAFM_SYSTEM_HPP

#include <array>
#include <stdexcept>
#include <string>
#include <vector>

#include "mafm.h"

namespace mafm {

using Vec3d = std::array<double, 3>;

// The fields of InputOptions that the minimisers read, with the defaults of
// readInputFile (parse.cpp:101-134); the reference resolves an unset minterm
// while reading the input, this struct starts from MAFM_TERM_F.
struct InputOptions {
    int minimiser_type = MAFM_MIN_FIRE;
    int integrator_type = MAFM_INT_MIDPOINT;
    int minterm = MAFM_TERM_F;
    int maxsteps = 5000;
    double etol = 0.01;
    double ftol = 0.01;
    double dt = 0.001;
};

struct System {
    explicit System(mafm_ctx* ctx) : ctx_(ctx) {}
    // sizes every state vector to n atoms, zeroed (System::initialize,
    // system.cpp:10-25); masses are 1 there and here
    void initialize(int n_atoms) {
        n_atoms_ = n_atoms;
        positions_.assign(n_atoms, Vec3d{0, 0, 0});
        velocities_.assign(n_atoms, Vec3d{0, 0, 0});
        forces_.assign(n_atoms, Vec3d{0, 0, 0});
        energies_.assign(n_atoms, 0.0);
        fixed_.assign(n_atoms, 0);
    }
    mafm_ctx* ctx_;
    int n_atoms_ = 0;   // dummy (0), tip (1) and surface atoms
    std::vector<Vec3d> positions_, velocities_, forces_;
    std::vector<double> energies_;
    std::vector<int32_t> fixed_;
    std::vector<mafm_flex_term> interactions_;   // the interaction list, in order
};

namespace detail {
inline mafm_interaction_list list_of(const System& s) {
    if ((int)s.positions_.size() != s.n_atoms_ || (int)s.velocities_.size() != s.n_atoms_ ||
        (int)s.forces_.size() != s.n_atoms_ || (int)s.energies_.size() != s.n_atoms_ ||
        (int)s.fixed_.size() != s.n_atoms_)
        throw std::runtime_error("System state vectors do not hold n_atoms_ entries (call initialize)");
    mafm_interaction_list l;
    l.n_atoms = s.n_atoms_;
    l.n_terms = (int32_t)s.interactions_.size();
    l.terms = s.interactions_.data();
    l.fixed = s.fixed_.data();
    return l;
}
inline void check(const System& s, mafm_status st) {
    if (st != MAFM_OK) throw std::runtime_error(mafm_last_error(s.ctx_));
}
inline double* d(std::vector<Vec3d>& v) { return v.empty() ? nullptr : v[0].data(); }
inline int minimise(System& s, const InputOptions& o, int minimiser) {
    const mafm_interaction_list l = list_of(s);
    mafm_minimiser_opts m;
    m.minimiser = minimiser;
    m.integrator = o.integrator_type;
    m.minterm = o.minterm;
    m.maxsteps = o.maxsteps;
    m.etol = o.etol;
    m.ftol = o.ftol;
    m.dt = o.dt;
    int32_t steps = 0;
    check(s, mafm_system_minimise(s.ctx_, &l, &m, 1, d(s.positions_), d(s.velocities_), d(s.forces_),
                                  s.energies_.data(), &steps));
    return steps;
}
inline void step(System& s, int integrator, double dt) {
    const mafm_interaction_list l = list_of(s);
    check(s, mafm_system_step(s.ctx_, &l, integrator, dt, 1, d(s.positions_), d(s.velocities_), d(s.forces_),
                              s.energies_.data()));
}
}  // namespace detail

inline int SDMinimisation(System& system, const InputOptions& options) {
    return detail::minimise(system, options, MAFM_MIN_SD);
}
inline int FIREMinimisation(System& system, const InputOptions& options) {
    return detail::minimise(system, options, MAFM_MIN_FIRE);
}
inline void eulerStep(System& system, const double dt) { detail::step(system, MAFM_INT_EULER, dt); }
inline void midpointStep(System& system, const double dt) { detail::step(system, MAFM_INT_MIDPOINT, dt); }
inline void rk4Step(System& system, const double dt) { detail::step(system, MAFM_INT_RK4, dt); }

inline void evalInteractions(System& system) {
    const mafm_interaction_list l = detail::list_of(system);
    detail::check(system, mafm_system_eval(system.ctx_, &l, 1, detail::d(system.positions_),
                                           detail::d(system.forces_), system.energies_.data()));
}

}  // namespace mafm

#endif
