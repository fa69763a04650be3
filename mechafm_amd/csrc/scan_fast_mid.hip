// Fast scan kernels for FIRE with the midpoint integrator (the reference's
// default minimiser), compiled apart from the generic family.
#include "scan_fast.hpp"

namespace mafm {

template hipError_t launch_scan_fast_family<true>(const ScanArgs&, bool, int, int, int, int, int, bool, hipStream_t);

}  // namespace mafm
