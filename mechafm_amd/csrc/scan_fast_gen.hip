// Fast scan kernels for every minimiser / integrator combination (steepest
// descent, FIRE with Euler, midpoint or RK4), and the launcher that picks
// the family.
#include "../../include/mafm.h"
#include "scan_fast.hpp"

namespace mafm {

// the FIRE + midpoint family is instantiated in scan_fast_mid.hip
extern template hipError_t launch_scan_fast_family<true>(const ScanArgs&, bool, int, int, int, int, int, bool,
                                                          hipStream_t);
template hipError_t launch_scan_fast_family<false>(const ScanArgs&, bool, int, int, int, int, int, bool, hipStream_t);

// n_blocks workgroups of W waves (W = 1 .. 8); lds_sites: largest
// per-molecule site count (sizes the LDS copy); scalar_sites: stream the
// sites with scalar loads instead (G = 1, molecules too large for LDS);
// mid_family: the FIRE + midpoint kernels (the host picks them only for
// that minimiser and non-periodic systems)
hipError_t launch_scan_fast(const ScanArgs& a, bool f32, int comp, int G, int n_blocks, int W, int lds_sites,
                            bool scalar_sites, bool mid_family, hipStream_t st) {
    if (mid_family) {
        if (a.minimiser != MAFM_MIN_FIRE || a.integrator != MAFM_INT_MIDPOINT) return hipErrorInvalidValue;
        return launch_scan_fast_family<true>(a, f32, comp, G, n_blocks, W, lds_sites, scalar_sites, st);
    }
    return launch_scan_fast_family<false>(a, f32, comp, G, n_blocks, W, lds_sites, scalar_sites, st);
}

}  // namespace mafm
