/*
 * mafm_sim.h -- C ABI over the C++ Simulation API of libmafm (host side).
 *
 * A mafm_sim is one MechAFM job: the input.scan / parameter / xyz (and cube)
 * files are parsed exactly as the reference does, the system is centred and
 * lowered to device tables, and the scan runs through mafm.h on the context
 * given at open.  The per-height output files are those of the reference.
 *
 * Reference call sites each entry replaces:
 *   mafm_sim_open          parseCommandLine / readInputFile / readXYZFile /
 *                          readParameterFile (parse.cpp:30-697) and the
 *                          scan-grid size of openUniverse (mechafm.cpp:47-49)
 *   mafm_sim_initialize    Simulation::initialize (simulation.cpp:31-65),
 *                          including buildInteractions and the force-grid /
 *                          external-potential construction
 *   mafm_sim_run           Simulation::run (simulation.cpp:67-143)
 *   mafm_batch_run         many jobs in one launch (no reference counterpart:
 *                          the reference runs one molecule per process)
 *   mafm_batch_run_output  the same, writing the files during the scan
 *                          (simulation.cpp:123-138 flushes during the run)
 *   mafm_sim_run_output    Simulation::run with the files written and the
 *                          progress reported during the scan (simulation.cpp:123-138)
 *   mafm_batch_write_output  writeOutput for many jobs on one thread pool
 *                          (simulation.cpp:145-199)
 *   mafm_sim_flex_system   the flexible interaction list of buildInteractions
 *                          (simulation.cpp:201-216, 503-693)
 *   mafm_sim_write_output  openUniverse / writeOutput / closeUniverse
 *                          (mechafm.cpp:59-101, simulation.cpp:145-199) and
 *                          the statistics block of finalize (mechafm.cpp:103-154)
 */
#ifndef MAFM_SIM_H
#define MAFM_SIM_H

#include "mafm.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct mafm_sim mafm_sim;

typedef struct {
    int32_t nx, ny, nz;           /* scan positions (mechafm.cpp:47-49) */
    int32_t n_atoms;              /* surface atoms */
    int32_t n_sites;              /* lowered tip-surface sites (pbc images included) */
    int32_t rigidgrid, use_external_potential, vdw_pbc, coulomb;
    int32_t force_grid_n[3];      /* force-grid dimensions when rigidgrid is on */
    double tip_dummy_d;
    double offset[3];             /* System offset after centring */
    double grid_build_ms;         /* device time spent building grids / potential */
    int32_t flexible;             /* flexible on: the job scans with mafm_scan_flexible */
    int32_t n_terms;              /* flexible on: terms of the interaction list */
} mafm_sim_info;

typedef struct {
    int32_t precision;            /* MAFM_PREC_* */
    int32_t lanes_per_pixel;      /* 0 = automatic */
    int32_t waves_per_simd;       /* 0 = automatic */
    int32_t compensation;         /* FAST32 sums (mafm_scan_desc::compensation) */
    int32_t waves_per_group;      /* 0 = automatic (mafm_scan_desc::waves_per_group) */
    uint32_t options;             /* mafm_scan_desc::options */
} mafm_run_opts;

/* ctx may be NULL: the job is then parsed and, where its set-up needs no
 * device (no rigidgrid, no external potential), initialised on the host
 * only -- enough for mafm_sim_system / mafm_sim_flex_system; running it
 * needs a context. */
mafm_status mafm_sim_open(const char* input_file, const char* output_dir, mafm_ctx* ctx, mafm_sim** out);
mafm_status mafm_sim_initialize(mafm_sim* sim);
mafm_status mafm_sim_info_get(mafm_sim* sim, mafm_sim_info* info);
/* the lowered system (valid until the sim is closed); rigid jobs */
mafm_status mafm_sim_system(mafm_sim* sim, mafm_system_desc* desc);
/* flexible jobs: the whole interaction list as device terms, with the
 * surface after centring and the fixed flags of the xyz file's sixth column
 * (valid until the sim is closed) */
mafm_status mafm_sim_flex_system(mafm_sim* sim, mafm_flex_desc* desc);
/* the scan descriptor this job's options give (for callers that drive
 * mafm_scan themselves, e.g. with modified site tables) */
mafm_status mafm_sim_scan_desc(mafm_sim* sim, const mafm_run_opts* opts, mafm_scan_desc* desc);
mafm_status mafm_sim_run(mafm_sim* sim, const mafm_run_opts* opts, mafm_scan_stats* stats);
mafm_status mafm_batch_run(int n, mafm_sim* const* sims, const mafm_run_opts* opts, mafm_scan_stats* stats);
/* A batch run that writes every job's output files while the batch scans
 * (the reference writes its buffer out during the scan,
 * simulation.cpp:123-138): one launch (mafm_scan_host) stores the records
 * in pinned host memory as pixel columns finish; as soon as every column of
 * a job is there, a writer thread copies the job's records and formats and
 * writes its files (each job's heights on `threads` threads, <= 0: every
 * hardware thread) while the rest of the batch scans.  Afterwards each job
 * holds its records on the host (mafm_sim_records serves them; delta f,
 * which reads the device records, needs another run). */
mafm_status mafm_batch_run_output(int n, mafm_sim* const* sims, const mafm_run_opts* opts, int threads,
                                  mafm_scan_stats* stats);
/* progress of a streamed run: the fraction (0..1) of the positions whose
 * records are in, called from the calling thread */
typedef void (*mafm_progress_fn)(double fraction, void* user);
/* One rigid job scanned with its files written as the records land (the
 * reference's Simulation::run, writing its buffer and reporting progress
 * from inside the scan, simulation.cpp:123-138): mafm_batch_run_output for
 * one job, with progress (may be NULL) called as heights of columns come
 * in and once with 1.0 at the end. */
mafm_status mafm_sim_run_output(mafm_sim* sim, const mafm_run_opts* opts, int threads, mafm_progress_fn progress,
                                void* user, mafm_scan_stats* stats);
/* the per-height files (and statistics.txt) of many rigid jobs that have
 * run, written on one pool of `threads` threads (<= 0: every hardware
 * thread), chunk by chunk of pixel columns, gzip deflated in parallel within
 * a file: mafm_sim_write_output for a batch (simulation.cpp:145-199) */
mafm_status mafm_batch_write_output(int n, mafm_sim* const* sims, int threads);
mafm_status mafm_sim_records(mafm_sim* sim, size_t first, size_t count, mafm_record* out);
/* per-height files scan-%06.3f.dat[.gz] with the reference's line format
 * (simulation.cpp:145-199); threads <= 0 uses every hardware thread */
mafm_status mafm_sim_write_output(mafm_sim* sim, int threads);
/* console lines of the last mafm_sim_write_output: with flexible on, one
 * "+- Wrote state_X-Y-Z.xyz" per height of the centre column, whose relaxed
 * atoms it writes as System::makeXYZFile does (simulation.cpp:114-116,
 * system.cpp:54-65); same buffer convention as mafm_sim_settings */
mafm_status mafm_sim_log(mafm_sim* sim, char* buf, size_t size, size_t* needed);
/* frequency-shift images of the last run (simple-vafm-image.py:76-204) as
 * deltaf-%06.3f.dat files in the output folder; desc NULL = the script's
 * amplitude, cantilever and frequency, and Fz as it reads it from the
 * per-height files (fz_from_text) */
mafm_status mafm_sim_write_deltaf(mafm_sim* sim, const mafm_deltaf_desc* desc);
/* the warnings raised while reading the inputs and the input-settings echo,
 * in the reference's console format (messages.cpp:29-56, parse.cpp:407-473,
 * 575-581); copies at most size - 1 characters and a terminating NUL into
 * buf, sets *needed (if not NULL) to the full length + 1 */
mafm_status mafm_sim_settings(mafm_sim* sim, char* buf, size_t size, size_t* needed);
const char* mafm_sim_last_error(const mafm_sim* sim);
/* frees the simulation and the grids it built; completes a pending
 * mafm_destroy of its context when it is the last simulation open there */
void mafm_sim_close(mafm_sim* sim);

#ifdef __cplusplus
}
#endif
#endif
