// Device workspace of the allocation-free scan entry point.
//
// mafm_scan_noalloc serves callers that cannot allocate device memory (code
// under HIP graph capture, tight serving loops, harnesses that forbid it):
// its work-stealing counters, rank order, initial shares and totals live in
// this fixed file-scope array instead of buffers the context allocates.
// Every call on one GPU shares that GPU's copy (calls on one context are
// serialised on its stream; calls from two contexts on one GPU that overlap
// in time race on it).  The host takes its address once per context, with
// the context's device current, when the context is created.
#include "device_common.hpp"

namespace mafm {
namespace {

static __device__ __align__(256) unsigned char scratch[sizeof(ScanScratch)];

}  // namespace

// Device address of `scratch` on the current device (the array's name in
// host code is only a stand-in for the symbol, not a device address).
hipError_t noalloc_scratch_address(void** out) {
    return hipGetSymbolAddress(out, HIP_SYMBOL(scratch));
}

}  // namespace mafm
