// placeholder until the FFT kernels land
#include "../../include/mafm.h"
extern "C" mafm_status mafm_potential_grid(mafm_ctx*, const mafm_potential_desc*, int*) { return MAFM_ERR_STATE; }
extern "C" mafm_status mafm_fft3d(mafm_ctx*, const int32_t*, int, const double*, double*) { return MAFM_ERR_STATE; }
