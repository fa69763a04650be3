"""mafm_scan_noalloc (the allocation-free scan) through the C ABI against
mafm_scan on the same systems: records into a device buffer the caller
allocated, bit for bit equal to mafm_scan's in the reproducible mode, on two
calls in a row, the second on a smaller input (fewer systems and pixels)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HIP_MEMCPY_D2H = 2


def _hip():
    # the HIP runtime libmafm itself is bound to (same soname, already loaded)
    from mechafm_amd import abi
    abi.lib()
    h = ctypes.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipDeviceSynchronize.argtypes = []
    return h


def _noalloc(ctx, desc, dev, capacity, n):
    from mechafm_amd import abi
    st = abi.ScanStats()
    ctx.check(abi.lib().mafm_scan_noalloc(ctx.h, ctypes.byref(desc), dev, capacity, ctypes.byref(st)))
    out = np.zeros(n, dtype=abi.RECORD_DTYPE)
    hip = _hip()
    assert hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), dev, n * abi.RECORD_DTYPE.itemsize,
                         HIP_MEMCPY_D2H) == 0
    return st.as_dict(), out


@pytest.mark.parametrize("precision", ["fast64", "fast32", "plain64"])
def test_noalloc_scan_equals_scan(gpu_ctx, tmp_path, precision):
    from mechafm_amd import abi, workloads as W
    hip = _hip()
    wls = W.make_c3(str(tmp_path), mini=True, count=4)
    sims = [abi.Simulation(w.input_path, gpu_ctx) for w in wls]
    descs = [s.system() for s in sims]
    cases = []
    for systems, shrink in ((descs, 0), (descs[1:3], 2)):   # second call: fewer systems, fewer pixels
        d = sims[0].scan_desc(precision, lanes_per_pixel=2, options=abi.SCAN_REPRODUCIBLE)
        d.nx -= shrink
        d.ny -= shrink // 2
        cases.append((systems, d, len(systems) * d.nx * d.ny * d.nz))
    cap = cases[0][2]
    dev = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(dev), cap * abi.RECORD_DTYPE.itemsize) == 0
    try:
        for k, (systems, d, n) in enumerate(cases):
            gpu_ctx.set_systems(systems)
            want_st = gpu_ctx.scan(d)
            want = gpu_ctx.records(0, n)
            got_st, got = _noalloc(gpu_ctx, d, dev, cap, n)
            print(f"{precision} call {k + 1}: {n} positions, {got_st['total_steps']} steps, "
                  f"{got_st['kernel_ms']:.2f} ms")
            assert got_st["total_steps"] == want_st["total_steps"] and got_st["positions"] == n
            assert (got == want).all()
            # the context holds no records of its own after the allocation-free scan
            with pytest.raises(abi.MafmError):
                gpu_ctx.records(0, 1)
        # a buffer too small for the scan (the systems of the last call) is
        # refused before anything runs
        st = abi.ScanStats()
        rc = abi.lib().mafm_scan_noalloc(gpu_ctx.h, ctypes.byref(cases[1][1]), dev, cases[1][2] - 1,
                                         ctypes.byref(st))
        assert rc == 1 and "record buffer" in abi.lib().mafm_last_error(gpu_ctx.h).decode()
    finally:
        hip.hipDeviceSynchronize()
        hip.hipFree(dev)
