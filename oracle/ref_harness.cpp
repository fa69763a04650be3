// Test infrastructure only (never shipped, never on a product path).
//
// Full-precision driver around the MechAFM reference objects compiled from
// /root/reference/src by oracle/Makefile.  The reference prints only 3-4
// decimals (simulation.cpp:180-195), which cannot pin 1e-4 A / 1e-3 relative
// parity, so this harness re-runs the reference's own public API and dumps
// every record as raw doubles:
//
//   ref_harness scan INPUT OUT.bin
//       readInputFile / readXYZFile / readParameterFile (parse.cpp:83,477,588),
//       the n_points formula of openUniverse (mechafm.cpp:47-49),
//       Simulation::initialize (simulation.cpp:31-65), then the pixel loop of
//       Simulation::run (simulation.cpp:67-143) with the reference's own
//       FIREMinimisation / SDMinimisation, System::getOutput and lowerTip.
//       Records are stored at index (i*ny + j)*nz + k, so the file is
//       independent of OpenMP scheduling.
//       With flexible on and STATE.bin given, the centre column's atom
//       positions after each height's minimisation (what System::makeXYZFile
//       writes, simulation.cpp:114-116) go to STATE.bin: nz x n_atoms x 3.
//   ref_harness forces INPUT POS.bin OUT.bin
//       after initialize(), the whole interaction list evaluated as one
//       integrator stage does (integrators.cpp:8-13) at each configuration
//       in POS.bin (n_atoms x 3 doubles each, dummy and tip included):
//       forces (n_atoms x 3) then energies (n_atoms) per configuration.
//   ref_harness eval INPUT POS.bin OUT.bin
//       after initialize(), System::evalTipSurfaceForces (system.cpp:41-51)
//       at each tip position in POS.bin (n x 3 doubles): the force-grid and
//       external-potential lookups are checked through this.
//
// Record layout (16 doubles): i j k n_steps  x y z  fx fy fz  rx ry rz  r angle e
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "globals.hpp"
#include "minimiser.hpp"
#include "parse.hpp"
#include "simulation.hpp"

static void setup(Simulation& sim, const char* input) {
    sim.root_process_ = 0;
    sim.current_process_ = 0;
    sim.n_processes_ = 1;
    sim.points_per_process_.assign(1, 0);
    char prog[] = "ref_harness";
    std::vector<char> in(input, input + strlen(input) + 1);
    char* argv[] = {prog, in.data()};
    parseCommandLine(2, argv, sim);
    readInputFile(sim);
    readXYZFile(sim);
    readParameterFile(sim);
    InputOptions& o = sim.options_;
    sim.n_total_ = 0;
    sim.n_points_.x = floor(o.area.x / o.dx) + 1;
    sim.n_points_.y = floor(o.area.y / o.dy) + 1;
    sim.n_points_.z = floor((o.zhigh - o.zlow) / o.dz) + 1;
    sim.initialize();
}

static int run_scan(const char* input, const char* out, const char* state_out) {
    Simulation sim;
    setup(sim, input);
    const InputOptions& o = sim.options_;
    const int nx = sim.n_points_.x, ny = sim.n_points_.y, nz = sim.n_points_.z;
    const int na = sim.system.n_atoms_;
    std::vector<double> rec((size_t)nx * ny * nz * 16, 0.0);
    std::vector<double> state((size_t)nz * na * 3, 0.0);
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < nx; ++i) {
        double x = i * o.dx;
        for (int j = 0; j < ny; ++j) {
            double y = j * o.dy;
            System s = sim.system;
            if (o.rigidgrid) s.setTipPbc(false);
            s.setDummyXY(x, y);
            s.setDummyZ(o.zhigh);
            for (int k = 0; k < nz; ++k) {
                int n = 0;
                if (o.minimiser_type == STEEPEST_DESCENT) n = SDMinimisation(s, o);
                else n = FIREMinimisation(s, o);
                if (o.flexible && i * ny + j == nx * ny / 2)
                    for (int a = 0; a < na; ++a) {
                        double* q = &state[((size_t)k * na + a) * 3];
                        q[0] = s.positions_[a].x; q[1] = s.positions_[a].y; q[2] = s.positions_[a].z;
                    }
                OutputData d = s.getOutput();
                double* r = &rec[(((size_t)i * ny + j) * nz + k) * 16];
                r[0] = i; r[1] = j; r[2] = k; r[3] = n;
                r[4] = d.position.x; r[5] = d.position.y; r[6] = d.position.z;
                r[7] = d.tip_force.x; r[8] = d.tip_force.y; r[9] = d.tip_force.z;
                r[10] = d.r_vec.x; r[11] = d.r_vec.y; r[12] = d.r_vec.z;
                r[13] = d.r; r[14] = d.angle; r[15] = d.tip_energy;
                s.lowerTip(o.dz);
            }
        }
    }
    FILE* f = fopen(out, "wb");
    if (!f) return 2;
    int hdr[4] = {nx, ny, nz, 16};
    fwrite(hdr, sizeof(int), 4, f);
    fwrite(rec.data(), sizeof(double), rec.size(), f);
    fclose(f);
    if (state_out) {
        f = fopen(state_out, "wb");
        if (!f) return 2;
        fwrite(state.data(), sizeof(double), state.size(), f);
        fclose(f);
    }
    printf("HARNESS nx=%d ny=%d nz=%d tipdummy=%.17g offset=%.17g,%.17g,%.17g\n", nx, ny, nz,
           sim.system.getTipDummyDistance(), sim.system.getOffset().x,
           sim.system.getOffset().y, sim.system.getOffset().z);
    return 0;
}

static int run_forces(const char* input, const char* posfile, const char* out) {
    Simulation sim;
    setup(sim, input);
    const int na = sim.system.n_atoms_;
    FILE* f = fopen(posfile, "rb");
    if (!f) return 2;
    fseek(f, 0, SEEK_END);
    long bytes = ftell(f);
    fseek(f, 0, SEEK_SET);
    size_t nc = bytes / (3 * na * sizeof(double));
    std::vector<double> pos(3 * na * nc), res(4 * na * nc);
    if (fread(pos.data(), sizeof(double), pos.size(), f) != pos.size()) return 3;
    fclose(f);
    for (size_t c = 0; c < nc; ++c) {
        System s = sim.system;
        for (int a = 0; a < na; ++a) s.positions_[a] = Vec3d(pos[(c * na + a) * 3], pos[(c * na + a) * 3 + 1],
                                                            pos[(c * na + a) * 3 + 2]);
        fill(s.forces_.begin(), s.forces_.end(), Vec3d(0));
        fill(s.energies_.begin(), s.energies_.end(), 0);
        for (const auto& it : *s.interactions_) it->eval(s.positions_, s.forces_, s.energies_);
        double* r = &res[c * 4 * na];
        for (int a = 0; a < na; ++a) {
            r[3 * a] = s.forces_[a].x; r[3 * a + 1] = s.forces_[a].y; r[3 * a + 2] = s.forces_[a].z;
            r[3 * na + a] = s.energies_[a];
        }
    }
    f = fopen(out, "wb");
    if (!f) return 2;
    fwrite(res.data(), sizeof(double), res.size(), f);
    fclose(f);
    printf("HARNESS n_atoms=%d interactions=%zu\n", na, sim.system.interactions_->size());
    return 0;
}

static int run_eval(const char* input, const char* posfile, const char* out) {
    Simulation sim;
    setup(sim, input);
    FILE* f = fopen(posfile, "rb");
    if (!f) return 2;
    fseek(f, 0, SEEK_END);
    long bytes = ftell(f);
    fseek(f, 0, SEEK_SET);
    size_t n = bytes / (3 * sizeof(double));
    std::vector<double> pos(3 * n), res(4 * n);
    if (fread(pos.data(), sizeof(double), 3 * n, f) != 3 * n) return 3;
    fclose(f);
#pragma omp parallel for
    for (long p = 0; p < (long)n; ++p) {
        System s = sim.system;
        s.positions_[1] = Vec3d(pos[3 * p], pos[3 * p + 1], pos[3 * p + 2]);
        Vec3d fo;
        double e;
        s.evalTipSurfaceForces(fo, e);
        res[4 * p] = fo.x; res[4 * p + 1] = fo.y; res[4 * p + 2] = fo.z; res[4 * p + 3] = e;
    }
    f = fopen(out, "wb");
    if (!f) return 2;
    fwrite(res.data(), sizeof(double), res.size(), f);
    fclose(f);
    return 0;
}

int main(int argc, char** argv) {
    if (argc >= 4 && strcmp(argv[1], "scan") == 0) return run_scan(argv[2], argv[3], argc >= 5 ? argv[4] : nullptr);
    if (argc >= 5 && strcmp(argv[1], "forces") == 0) return run_forces(argv[2], argv[3], argv[4]);
    if (argc >= 5 && strcmp(argv[1], "eval") == 0) return run_eval(argv[2], argv[3], argv[4]);
    fprintf(stderr, "usage: ref_harness scan INPUT OUT.bin [STATE.bin] | forces INPUT POS.bin OUT.bin | "
                    "eval INPUT POS.bin OUT.bin\n");
    return 1;
}
