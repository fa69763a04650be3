// Context lifetime shared with the job layer (host/simulation.cpp): every
// open mafm_sim holds a reference on its context, so mafm_destroy of a
// context that simulations still use is completed when the last of them is
// closed (their destructors free the grids they built on it).
#pragma once
#include <stdint.h>

struct mafm_ctx;

namespace mafm {
void ctx_retain(mafm_ctx* c);
void ctx_release(mafm_ctx* c);
// Number of scans the context has run: a simulation's records stay valid
// only while this equals the count right after its own scan.
uint64_t ctx_scan_generation(const mafm_ctx* c);
}  // namespace mafm
