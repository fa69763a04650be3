// Drives the C++ System-level API (include/mafm_system.hpp) the way the
// reference's Simulation::run drives its own (simulation.cpp:93-140): one
// System copy of a flexible job, dummy and tip placed at pixel (0, 0), then
// for every height FIREMinimisation (or SDMinimisation) and lowerTip.
// Prints "k iterations rx ry rz" per height (tip minus dummy, %.17g), for
// comparison with the reference's records of that column.
//   system_api_check INPUT.scan
#include <cstdio>

#include "../include/mafm_sim.h"
#include "../include/mafm_system.hpp"

int main(int argc, char** argv) {
    if (argc < 2) {
        fprintf(stderr, "usage: %s INPUT.scan\n", argv[0]);
        return 2;
    }
    mafm_ctx* ctx = nullptr;
    if (mafm_create(0, &ctx) != MAFM_OK) {
        fprintf(stderr, "no device\n");
        return 3;
    }
    mafm_sim* sim = nullptr;
    mafm_flex_desc fd;
    mafm_scan_desc sd;
    if (mafm_sim_open(argv[1], nullptr, ctx, &sim) != MAFM_OK || mafm_sim_initialize(sim) != MAFM_OK ||
        mafm_sim_flex_system(sim, &fd) != MAFM_OK || mafm_sim_scan_desc(sim, nullptr, &sd) != MAFM_OK) {
        fprintf(stderr, "%s\n", sim ? mafm_sim_last_error(sim) : "open failed");
        return 1;
    }
    mafm::System system(ctx);
    system.initialize(fd.n_atoms);
    for (int i = 0; i < fd.n_atoms; ++i) {
        system.positions_[i] = {fd.positions[3 * i], fd.positions[3 * i + 1], fd.positions[3 * i + 2]};
        system.fixed_[i] = fd.fixed[i];
    }
    system.interactions_.assign(fd.terms, fd.terms + fd.n_terms);
    mafm::InputOptions options;
    options.minimiser_type = sd.minimiser;
    options.integrator_type = sd.integrator;
    options.minterm = sd.minterm;
    options.maxsteps = sd.maxsteps;
    options.etol = sd.etol;
    options.ftol = sd.ftol;
    options.dt = sd.dt;
    // setDummyXY(0, 0) and setDummyZ(zhigh) (system.cpp:139-162, system.hpp:47-50)
    system.positions_[0] = {0.0, 0.0, sd.zhigh};
    system.positions_[1] = {0.0, 0.0, sd.zhigh - fd.tip_dummy_d};
    try {
        for (int k = 0; k < sd.nz; ++k) {
            const int n = options.minimiser_type == MAFM_MIN_SD ? mafm::SDMinimisation(system, options)
                                                               : mafm::FIREMinimisation(system, options);
            const mafm::Vec3d& d = system.positions_[0];
            const mafm::Vec3d& t = system.positions_[1];
            printf("%d %d %.17g %.17g %.17g\n", k, n, t[0] - d[0], t[1] - d[1], t[2] - d[2]);
            system.positions_[0][2] -= sd.dz;   // lowerTip (system.hpp:51-54)
            system.positions_[1][2] -= sd.dz;
        }
        // the other operators once, so every wrapper is exercised
        mafm::midpointStep(system, options.dt);
        mafm::eulerStep(system, options.dt);
        mafm::rk4Step(system, options.dt);
        mafm::evalInteractions(system);
    } catch (const std::exception& e) {
        fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    mafm_sim_close(sim);
    mafm_destroy(ctx);
    return 0;
}
