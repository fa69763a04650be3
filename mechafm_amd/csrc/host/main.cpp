// mechafm: the reference's command line (mechafm.cpp:209-233),
//   mechafm INPUT [OUTDIR] [--precision plain64|fast64|fast32] [--device N] [--deltaf]
// parse -> initialize (interactions, force grid / external potential on the
// device) -> one device scan (rigid or flexible) -> per-height output files.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "../../../include/mafm.h"
#include "../../../include/mafm_sim.h"

// error() of the reference (messages.cpp:12-27): the message on stderr, exit code 1
static int die(const char* msg) {
    fprintf(stderr, "+- ERROR (on process 0): %s\n", msg);
    return 1;
}

int main(int argc, char** argv) {
    std::string input, outdir;
    mafm_run_opts ro{};
    ro.precision = MAFM_PREC_FAST64;
    int device = 0;
    bool deltaf = false;
    for (int i = 1; i < argc; ++i) {
        if (!strcmp(argv[i], "--precision") && i + 1 < argc) {
            std::string p = argv[++i];
            if (p == "plain64") ro.precision = MAFM_PREC_PLAIN64;
            else if (p == "fast64") ro.precision = MAFM_PREC_FAST64;
            else if (p == "fast32") ro.precision = MAFM_PREC_FAST32;
            else return die("--precision expects plain64, fast64 or fast32");
        } else if (!strcmp(argv[i], "--device") && i + 1 < argc) {
            device = atoi(argv[++i]);
        } else if (!strcmp(argv[i], "--deltaf")) {
            deltaf = true;
        } else if (input.empty()) {
            input = argv[i];
        } else if (outdir.empty()) {
            outdir = argv[i];
        } else {
            return die("Too many command line arguments!");
        }
    }
    if (input.empty()) return die("Specify an input file to be read!");
    mafm_ctx* ctx = nullptr;
    mafm_status st = mafm_create(device, &ctx);
    if (st != MAFM_OK) return die("no usable HIP device");
    mafm_sim* sim = nullptr;
    auto t0 = std::chrono::steady_clock::now();
    printf("+- MechAFM probe-particle scan on gfx950 (libmafm)\n");
    st = mafm_sim_open(input.c_str(), outdir.c_str(), ctx, &sim);
    if (st == MAFM_OK) {
        // input warnings (stderr) and the settings echo (stdout), as the
        // reference prints them (messages.cpp:29-56, parse.cpp:407-473)
        size_t need = 0;
        mafm_sim_settings(sim, nullptr, 0, &need);
        std::string text(need, '\0');
        mafm_sim_settings(sim, &text[0], need, nullptr);
        text.resize(strlen(text.c_str()));
        const size_t echo = text.find("+- \n+- Input settings");
        fputs(text.substr(0, echo).c_str(), stderr);
        fflush(stderr);
        if (echo != std::string::npos) fputs(text.substr(echo).c_str(), stdout);
        st = mafm_sim_initialize(sim);
    }
    if (st != MAFM_OK) {
        int rc = die(mafm_sim_last_error(sim));
        mafm_sim_close(sim);
        mafm_destroy(ctx);
        return rc;
    }
    mafm_sim_info info;
    mafm_sim_info_get(sim, &info);
    printf("+- 3D data grid is: %d x %d x %d (%lld in total)\n", info.nx, info.ny, info.nz,
           (long long)info.nx * info.ny * info.nz);
    printf("+- Starting simulation\n");
    fflush(stdout);
    mafm_scan_stats ss{};
    if (!info.flexible && !deltaf) {
        // files written while the scan runs, progress as the reference reports
        // it: "Finished x %" at every tenth (simulation.cpp:131-137)
        double next_report = 0.1;
        auto report = [](double fraction, void* user) {
            double& next = *static_cast<double*>(user);
            while (fraction >= next - 1e-12 && next <= 1.0 + 1e-12) {
                printf("+- Finished %4.1f %% of the simulation\n", 100 * next);
                next += 0.1;
            }
            fflush(stdout);
        };
        st = mafm_sim_run_output(sim, &ro, 0, report, &next_report, &ss);
    } else {
        // flexible jobs and delta f images keep the device records: scan, then write
        st = mafm_sim_run(sim, &ro, &ss);
        if (st == MAFM_OK) st = mafm_sim_write_output(sim, 0);
    }
    if (st == MAFM_OK) {
        // flexible on: the "+- Wrote state_..." lines of makeXYZFile (system.cpp:63)
        size_t need = 0;
        mafm_sim_log(sim, nullptr, 0, &need);
        std::string log(need, '\0');
        mafm_sim_log(sim, &log[0], need, nullptr);
        fputs(log.c_str(), stdout);
    }
    if (st == MAFM_OK && deltaf) st = mafm_sim_write_deltaf(sim, nullptr);
    auto t1 = std::chrono::steady_clock::now();
    if (st != MAFM_OK) {
        int rc = die(mafm_sim_last_error(sim));
        mafm_sim_close(sim);
        mafm_destroy(ctx);
        return rc;
    }
    // the reference's closing lines (mechafm.cpp:130-139, 197-200)
    long long n = (long long)info.nx * info.ny * info.nz;
    const double wall = std::chrono::duration<double>(t1 - t0).count();
    printf("+- Simulation run finished\n+- Statistics:\n");
    printf("+-     Computed %lld tip positions\n", n);
    printf("+-     Needed %lld minimization steps in total\n", (long long)ss.total_steps);
    printf("+-     Which means approximately %.2f minimization steps per tip position\n",
           (double)ss.total_steps / (double)n);
    printf("+-     The simulation wall time is %.2f seconds\n", wall);
    printf("+-     The entire simulation took %.2f seconds\n", wall);
    printf("+- \n");
    printf("+- How many x,y points did each process handle:\n");
    printf("+-     Process %2d: %6d x,y points\n", 0, info.nx * info.ny);
    printf("+- The scan kernel took %.3f seconds (%.0f positions/s)\n", ss.kernel_ms / 1e3, n / (ss.kernel_ms / 1e3));
    mafm_sim_close(sim);
    mafm_destroy(ctx);
    return 0;
}
