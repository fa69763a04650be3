"""ctypes binding of libmafm (include/mafm.h, include/mafm_sim.h).

The library is the product: every call here goes to the gfx950 kernels.  If
``mechafm_amd/lib/libmafm.so`` is missing this module raises instead of
falling back to anything else.
"""
from __future__ import annotations

import ctypes
import os
from typing import List, Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libmafm.so")

MAFM_OK = 0
STATUS = {0: "OK", 1: "INVALID", 2: "HIP", 3: "OUT_OF_GRID", 4: "STATE", 5: "FFT_IMAG", 6: "NO_DEVICE"}
PREC = {"plain64": 0, "fast64": 1, "fast32": 2}
SCAN_REPRODUCIBLE = 1
SCAN_MARGINS = 4
SCAN_STATE_IN_HBM = 8
VDW_LJ, VDW_MORSE = 0, 1


class Site(ctypes.Structure):
    _fields_ = [("x", ctypes.c_double), ("y", ctypes.c_double), ("z", ctypes.c_double),
                ("vdw", ctypes.c_int32), ("has_coulomb", ctypes.c_int32),
                ("p0", ctypes.c_double), ("p1", ctypes.c_double), ("p2", ctypes.c_double),
                ("qq", ctypes.c_double)]


class TipDummy(ctypes.Structure):
    _fields_ = [("vdw", ctypes.c_int32), ("has_coulomb", ctypes.c_int32),
                ("p0", ctypes.c_double), ("p1", ctypes.c_double), ("p2", ctypes.c_double),
                ("qq", ctypes.c_double), ("k", ctypes.c_double), ("r0", ctypes.c_double)]


class SystemDesc(ctypes.Structure):
    _fields_ = [("n_sites", ctypes.c_int32), ("force_grid", ctypes.c_int32),
                ("sites", ctypes.POINTER(Site)), ("potential_grid", ctypes.c_int32),
                ("tip_pbc", ctypes.c_int32), ("tip_dummy", TipDummy), ("tip_dummy_d", ctypes.c_double),
                ("cell", ctypes.c_double * 9), ("offset", ctypes.c_double * 3),
                ("n_images", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("image_shifts", ctypes.POINTER(ctypes.c_double))]


class GridDesc(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32 * 3), ("periodic", ctypes.c_int32), ("basis", ctypes.c_double * 9),
                ("offset", ctypes.c_double * 3), ("fe", ctypes.POINTER(ctypes.c_double))]


class ScanDesc(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
                ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz", ctypes.c_double),
                ("zhigh", ctypes.c_double), ("minimiser", ctypes.c_int32), ("integrator", ctypes.c_int32),
                ("minterm", ctypes.c_int32), ("etol", ctypes.c_double), ("ftol", ctypes.c_double),
                ("dt", ctypes.c_double), ("maxsteps", ctypes.c_int32), ("precision", ctypes.c_int32),
                ("lanes_per_pixel", ctypes.c_int32), ("waves_per_simd", ctypes.c_int32),
                ("compensation", ctypes.c_int32), ("waves_per_group", ctypes.c_int32),
                ("options", ctypes.c_uint32), ("reserved", ctypes.c_int32)]


RECORD_DTYPE = np.dtype([("x", "f8"), ("y", "f8"), ("z", "f8"), ("fx", "f8"), ("fy", "f8"), ("fz", "f8"),
                         ("rx", "f8"), ("ry", "f8"), ("rz", "f8"), ("e", "f8"),
                         ("n_steps", "i4"), ("flags", "i4")])


class ScanStats(ctypes.Structure):
    _fields_ = [("kernel_ms", ctypes.c_double), ("positions", ctypes.c_int64), ("total_steps", ctypes.c_int64),
                ("unconverged", ctypes.c_int64), ("grid_warnings", ctypes.c_int64), ("pair_evals", ctypes.c_int64),
                ("lanes_per_pixel", ctypes.c_int32), ("waves_per_simd", ctypes.c_int32),
                ("grid_blocks", ctypes.c_int32), ("block_threads", ctypes.c_int32),
                ("group_trips", ctypes.c_int64), ("active_group_trips", ctypes.c_int64),
                ("wide_trips", ctypes.c_int64)]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


class DeviceProps(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 64), ("compute_units", ctypes.c_int32), ("clock_mhz", ctypes.c_int32),
                ("total_mem", ctypes.c_int64), ("lds_per_block", ctypes.c_int32), ("warp_size", ctypes.c_int32)]


class PotentialDesc(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32 * 3), ("basis", ctypes.c_double * 9), ("origin", ctypes.c_double * 3),
                ("potential", ctypes.POINTER(ctypes.c_double)), ("tip_charge", ctypes.c_double),
                ("gaussian_width", ctypes.c_double)]


class AdjointDesc(ctypes.Structure):
    _fields_ = [("n_positions", ctypes.c_int64), ("tip", ctypes.POINTER(ctypes.c_double)),
                ("dummy", ctypes.POINTER(ctypes.c_double)), ("g", ctypes.POINTER(ctypes.c_double)),
                ("precision", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class AdjointOut(ctypes.Structure):
    _fields_ = [("d_es6", ctypes.POINTER(ctypes.c_double)), ("d_es12", ctypes.POINTER(ctypes.c_double)),
                ("d_qq", ctypes.POINTER(ctypes.c_double)), ("d_pos", ctypes.POINTER(ctypes.c_double))]


class SimInfo(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
                ("n_atoms", ctypes.c_int32), ("n_sites", ctypes.c_int32), ("rigidgrid", ctypes.c_int32),
                ("use_external_potential", ctypes.c_int32), ("vdw_pbc", ctypes.c_int32),
                ("coulomb", ctypes.c_int32), ("force_grid_n", ctypes.c_int32 * 3),
                ("tip_dummy_d", ctypes.c_double), ("offset", ctypes.c_double * 3),
                ("grid_build_ms", ctypes.c_double), ("flexible", ctypes.c_int32), ("n_terms", ctypes.c_int32)]


# flexible on: one term of the interaction list (mafm_flex_term)
FX_LJ, FX_MORSE, FX_COULOMB, FX_TIP_HARMONIC, FX_BOND, FX_ANGLE, FX_DIHEDRAL, FX_SUBSTRATE, FX_XY_HARMONIC = range(9)
FLEX_TERM_DTYPE = np.dtype([("kind", "<i4"), ("atom", "<i4", (4,)), ("reserved", "<i4"), ("p", "<f8", (6,)),
                            ("shift", "<f8", (3,))])


class FlexDesc(ctypes.Structure):
    _fields_ = [("n_atoms", ctypes.c_int32), ("n_terms", ctypes.c_int32),
                ("positions", ctypes.POINTER(ctypes.c_double)), ("fixed", ctypes.POINTER(ctypes.c_int32)),
                ("terms", ctypes.c_void_p), ("tip_pbc", ctypes.c_int32), ("state_column", ctypes.c_int32),
                ("tip_dummy_d", ctypes.c_double), ("cell", ctypes.c_double * 9), ("offset", ctypes.c_double * 3)]


class InteractionList(ctypes.Structure):
    _fields_ = [("n_atoms", ctypes.c_int32), ("n_terms", ctypes.c_int32), ("terms", ctypes.c_void_p),
                ("fixed", ctypes.POINTER(ctypes.c_int32))]


class MinimiserOpts(ctypes.Structure):
    _fields_ = [("minimiser", ctypes.c_int32), ("integrator", ctypes.c_int32), ("minterm", ctypes.c_int32),
                ("maxsteps", ctypes.c_int32), ("etol", ctypes.c_double), ("ftol", ctypes.c_double),
                ("dt", ctypes.c_double)]


class RunOpts(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int32), ("lanes_per_pixel", ctypes.c_int32),
                ("waves_per_simd", ctypes.c_int32), ("compensation", ctypes.c_int32),
                ("waves_per_group", ctypes.c_int32), ("options", ctypes.c_uint32)]


# every exported entry point and its signature: (restype, argtypes)
_P = ctypes.POINTER
_vp = ctypes.c_void_p
class DeltaFDesc(ctypes.Structure):
    _fields_ = [("amplitude", ctypes.c_double), ("k_cantilever", ctypes.c_double), ("f0", ctypes.c_double),
                ("fz_from_text", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class DeltaFWeights(ctypes.Structure):
    _fields_ = [("n_weights", ctypes.c_int32), ("n_images", ctypes.c_int32), ("dz", ctypes.c_double),
                ("amplitude", ctypes.c_double), ("prefactor", ctypes.c_double), ("weights", ctypes.c_double * 128)]


SIGNATURES = {
    "mafm_create": (ctypes.c_int, [ctypes.c_int, _P(_vp)]),
    "mafm_destroy": (None, [_vp]),
    "mafm_last_error": (ctypes.c_char_p, [_vp]),
    "mafm_device_info": (ctypes.c_int, [_vp, _P(DeviceProps)]),
    "mafm_grid_upload": (ctypes.c_int, [_vp, _P(GridDesc), _P(ctypes.c_int)]),
    "mafm_grid_build": (ctypes.c_int, [_vp, _P(SystemDesc), _P(GridDesc), _P(ctypes.c_int)]),
    "mafm_grid_download": (ctypes.c_int, [_vp, ctypes.c_int, _P(ctypes.c_double)]),
    "mafm_grid_free": (ctypes.c_int, [_vp, ctypes.c_int]),
    "mafm_potential_grid": (ctypes.c_int, [_vp, _P(PotentialDesc), _P(ctypes.c_int)]),
    "mafm_fft3d": (ctypes.c_int, [_vp, _P(ctypes.c_int32), ctypes.c_int, _P(ctypes.c_double), _P(ctypes.c_double)]),
    "mafm_set_systems": (ctypes.c_int, [_vp, ctypes.c_int, _P(SystemDesc)]),
    "mafm_scan": (ctypes.c_int, [_vp, _P(ScanDesc), _P(ScanStats)]),
    "mafm_scan_noalloc": (ctypes.c_int, [_vp, _P(ScanDesc), _vp, ctypes.c_size_t, _P(ScanStats)]),
    "mafm_scan_host": (ctypes.c_int, [_vp, _P(ScanDesc), _P(_vp), _P(_P(ctypes.c_int32))]),
    "mafm_scan_host_query": (ctypes.c_int, [_vp, _P(ctypes.c_int)]),
    "mafm_scan_host_wait": (ctypes.c_int, [_vp, _P(ScanStats)]),
    "mafm_records": (ctypes.c_int, [_vp, ctypes.c_size_t, ctypes.c_size_t, _vp]),
    "mafm_grid_interpolate": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_size_t, _P(ctypes.c_double),
                                             _P(ctypes.c_double)]),
    "mafm_margins": (ctypes.c_int, [_vp, ctypes.c_size_t, ctypes.c_size_t, _P(ctypes.c_float)]),
    "mafm_valu_calibration": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, _P(ctypes.c_double),
                                             _P(ctypes.c_int64)]),
    "mafm_scan_flexible": (ctypes.c_int, [_vp, _P(FlexDesc), _P(ScanDesc), _P(ScanStats)]),
    "mafm_system_eval": (ctypes.c_int, [_vp, _P(InteractionList), ctypes.c_int64, _P(ctypes.c_double),
                                        _P(ctypes.c_double), _P(ctypes.c_double)]),
    "mafm_system_step": (ctypes.c_int, [_vp, _P(InteractionList), ctypes.c_int32, ctypes.c_double, ctypes.c_int64,
                                        _P(ctypes.c_double), _P(ctypes.c_double), _P(ctypes.c_double),
                                        _P(ctypes.c_double)]),
    "mafm_system_minimise": (ctypes.c_int, [_vp, _P(InteractionList), _P(MinimiserOpts), ctypes.c_int64,
                                            _P(ctypes.c_double), _P(ctypes.c_double), _P(ctypes.c_double),
                                            _P(ctypes.c_double), _P(ctypes.c_int32)]),
    "mafm_flex_state": (ctypes.c_int, [_vp, ctypes.c_int32, ctypes.c_int32, _P(ctypes.c_double)]),
    "mafm_eval_tip_surface": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_size_t, _P(ctypes.c_double),
                                             _P(ctypes.c_double)]),
    "mafm_adjoint": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _vp]),
    "mafm_deltaf_weights": (ctypes.c_int, [_vp, _P(DeltaFDesc), ctypes.c_int32, ctypes.c_double, ctypes.c_double,
                                           _P(DeltaFWeights)]),
    "mafm_delta_f": (ctypes.c_int, [_vp, ctypes.c_int, _P(DeltaFDesc), _P(ctypes.c_double), _P(ctypes.c_double)]),
    "mafm_sim_open": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_char_p, _vp, _P(_vp)]),
    "mafm_sim_initialize": (ctypes.c_int, [_vp]),
    "mafm_sim_info_get": (ctypes.c_int, [_vp, _P(SimInfo)]),
    "mafm_sim_system": (ctypes.c_int, [_vp, _P(SystemDesc)]),
    "mafm_sim_flex_system": (ctypes.c_int, [_vp, _P(FlexDesc)]),
    "mafm_sim_log": (ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.c_size_t, _P(ctypes.c_size_t)]),
    "mafm_sim_scan_desc": (ctypes.c_int, [_vp, _P(RunOpts), _P(ScanDesc)]),
    "mafm_sim_run": (ctypes.c_int, [_vp, _P(RunOpts), _P(ScanStats)]),
    "mafm_batch_run": (ctypes.c_int, [ctypes.c_int, _P(_vp), _P(RunOpts), _P(ScanStats)]),
    "mafm_batch_run_output": (ctypes.c_int, [ctypes.c_int, _P(_vp), _P(RunOpts), ctypes.c_int, _P(ScanStats)]),
    "mafm_batch_write_output": (ctypes.c_int, [ctypes.c_int, _P(_vp), ctypes.c_int]),
    "mafm_sim_run_output": (ctypes.c_int, [_vp, _P(RunOpts), ctypes.c_int, _vp, _vp, _P(ScanStats)]),
    "mafm_sim_settings": (ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.c_size_t, _P(ctypes.c_size_t)]),
    "mafm_sim_records": (ctypes.c_int, [_vp, ctypes.c_size_t, ctypes.c_size_t, _vp]),
    "mafm_sim_write_output": (ctypes.c_int, [_vp, ctypes.c_int]),
    "mafm_sim_write_deltaf": (ctypes.c_int, [_vp, _P(DeltaFDesc)]),
    "mafm_sim_last_error": (ctypes.c_char_p, [_vp]),
    "mafm_sim_close": (None, [_vp]),
}

_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libmafm is not built ({LIB_PATH}); run __graft_entry__.build()")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class MafmError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status


def _dp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Context:
    """A libmafm context on one device (mafm_create / mafm_destroy)."""

    def __init__(self, device: int = 0):
        self.h = ctypes.c_void_p()
        st = lib().mafm_create(device, ctypes.byref(self.h))
        if st != MAFM_OK:
            raise MafmError(st, "mafm_create failed (no HIP device?)")
        self._keep = []

    def check(self, st: int) -> None:
        if st != MAFM_OK:
            raise MafmError(st, lib().mafm_last_error(self.h).decode())

    def close(self) -> None:
        if self.h:
            lib().mafm_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def device_info(self) -> dict:
        p = DeviceProps()
        self.check(lib().mafm_device_info(self.h, ctypes.byref(p)))
        return {"name": p.name.decode(), "compute_units": p.compute_units, "clock_mhz": p.clock_mhz,
                "total_mem": p.total_mem, "lds_per_block": p.lds_per_block, "warp_size": p.warp_size}

    # ---- grids
    def grid_upload(self, n, basis, offset, fe: np.ndarray, periodic: bool) -> int:
        fe = np.ascontiguousarray(fe, dtype=np.float64)
        g = GridDesc()
        for a in range(3):
            g.n[a] = int(n[a])
            g.offset[a] = float(offset[a])
        for a, v in enumerate(np.asarray(basis, dtype=np.float64).ravel()):
            g.basis[a] = v
        g.periodic = int(periodic)
        g.fe = _dp(fe)
        h = ctypes.c_int(-1)
        self.check(lib().mafm_grid_upload(self.h, ctypes.byref(g), ctypes.byref(h)))
        return h.value

    def grid_free(self, handle: int) -> None:
        self.check(lib().mafm_grid_free(self.h, handle))

    def grid_interpolate(self, handle: int, pos: np.ndarray) -> np.ndarray:
        """ForceGrid::interpolate at pos (n, 3): (n, 4) fx fy fz e."""
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        out = np.zeros((len(pos), 4))
        self.check(lib().mafm_grid_interpolate(self.h, handle, len(pos), _dp(pos), _dp(out)))
        return out

    def grid_download(self, handle: int, n) -> np.ndarray:
        out = np.zeros((int(n[0]) * int(n[1]) * int(n[2]), 4))
        self.check(lib().mafm_grid_download(self.h, handle, _dp(out)))
        return out

    def potential_grid(self, pot: np.ndarray, basis, origin, tip_charge: float, width: float = 0.5) -> int:
        pot = np.ascontiguousarray(pot, dtype=np.float64)
        d = PotentialDesc()
        for a in range(3):
            d.n[a] = pot.shape[a]
            d.origin[a] = float(origin[a])
        for a, v in enumerate(np.asarray(basis, dtype=np.float64).ravel()):
            d.basis[a] = v
        d.potential = _dp(pot)
        d.tip_charge = tip_charge
        d.gaussian_width = width
        h = ctypes.c_int(-1)
        self.check(lib().mafm_potential_grid(self.h, ctypes.byref(d), ctypes.byref(h)))
        return h.value

    def fft3d(self, x: np.ndarray, inverse: bool = False) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.complex128)
        n = (ctypes.c_int32 * 3)(*x.shape)
        out = np.zeros_like(x)
        self.check(lib().mafm_fft3d(self.h, n, int(inverse), x.view(np.float64).ctypes.data_as(_P(ctypes.c_double)),
                                    out.view(np.float64).ctypes.data_as(_P(ctypes.c_double))))
        return out

    # ---- systems and scans
    def set_systems(self, descs: Sequence[SystemDesc]) -> None:
        arr = (SystemDesc * len(descs))(*descs)
        self._keep = [arr]
        self.check(lib().mafm_set_systems(self.h, len(descs), arr))

    def scan(self, desc: ScanDesc) -> dict:
        st = ScanStats()
        self.check(lib().mafm_scan(self.h, ctypes.byref(desc), ctypes.byref(st)))
        return st.as_dict()

    def scan_host(self, desc: ScanDesc, n_systems: int):
        """Launch a streamed scan of the n_systems set by set_systems
        (mafm_scan_host); returns numpy views of the pinned host records and
        the per-column flags, valid until the next streamed scan on this
        context.  End it with scan_host_wait()."""
        rec = ctypes.c_void_p()
        done = _P(ctypes.c_int32)()
        self.check(lib().mafm_scan_host(self.h, ctypes.byref(desc), ctypes.byref(rec), ctypes.byref(done)))
        n_cols = desc.nx * desc.ny * n_systems
        n_rec = n_cols * desc.nz
        recs = np.frombuffer((ctypes.c_char * (n_rec * RECORD_DTYPE.itemsize)).from_address(rec.value),
                             dtype=RECORD_DTYPE)
        flags = np.ctypeslib.as_array(done, shape=(n_cols,))
        return recs, flags

    def scan_host_query(self) -> bool:
        fin = ctypes.c_int()
        self.check(lib().mafm_scan_host_query(self.h, ctypes.byref(fin)))
        return bool(fin.value)

    def scan_host_wait(self) -> dict:
        st = ScanStats()
        self.check(lib().mafm_scan_host_wait(self.h, ctypes.byref(st)))
        return st.as_dict()

    def records(self, first: int, count: int) -> np.ndarray:
        out = np.zeros(count, dtype=RECORD_DTYPE)
        self.check(lib().mafm_records(self.h, first, count, out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def margins(self, first: int, count: int) -> np.ndarray:
        """(count, 4) float32 decision margins {f, e, p, 0} of the last scan
        run with SCAN_MARGINS (mafm_margins)."""
        out = np.zeros((count, 4), dtype=np.float32)
        self.check(lib().mafm_margins(self.h, first, count, out.ctypes.data_as(_P(ctypes.c_float))))
        return out

    def valu_calibration(self, kind: int, iterations: int = 4096) -> dict:
        """One launch of the VALU calibration chains (mafm_valu_calibration):
        kind 0 FP32 FMAs, 1 FP64 FMAs, 2 FP64 v_rsq + add."""
        ms = ctypes.c_double()
        insts = ctypes.c_int64()
        self.check(lib().mafm_valu_calibration(self.h, kind, iterations, ctypes.byref(ms), ctypes.byref(insts)))
        return {"ms": ms.value, "wave_instructions": insts.value}

    def adjoint(self, system: int, n_sites: int, tip: np.ndarray, dummy: np.ndarray, g: np.ndarray,
                precision: str = "fast64"):
        """Site gradients (d_es6, d_es12, d_qq, d_pos) of sum g*Fz by the
        implicit function theorem at the relaxed tips (mafm_adjoint)."""
        tip = np.ascontiguousarray(tip, dtype=np.float64)
        dummy = np.ascontiguousarray(dummy, dtype=np.float64)
        g = np.ascontiguousarray(g, dtype=np.float64)
        d = AdjointDesc(len(g), _dp(tip), _dp(dummy), _dp(g), PREC[precision], 0)
        out = [np.zeros(n_sites), np.zeros(n_sites), np.zeros(n_sites), np.zeros((n_sites, 3))]
        o = AdjointOut(*[_dp(a) for a in out])
        self.check(lib().mafm_adjoint(self.h, system, ctypes.byref(d), ctypes.byref(o)))
        return tuple(out)

    def delta_f(self, system: int, nx: int, ny: int, nz: int, zhigh: float, dz: float, amplitude: float = 0.0,
                k_cantilever: float = 0.0, f0: float = 0.0, fz_from_text: bool = False):
        """Frequency-shift images of `system` from the last scan's records
        (mafm_delta_f): (df [n_images, nx, ny] in Hz, z_low [n_images]).
        fz_from_text: Fz as the script reads it from the "%8.4f" files."""
        d = DeltaFDesc(amplitude, k_cantilever, f0, int(fz_from_text), 0)
        w = deltaf_weights(nz, zhigh, dz, amplitude, k_cantilever, f0)
        df = np.zeros((w["n_images"], nx, ny))
        zl = np.zeros(w["n_images"])
        self.check(lib().mafm_delta_f(self.h, system, ctypes.byref(d), _dp(df), _dp(zl)))
        return df, zl

    # ---- System-level operators (mafm_system_eval / _step / _minimise) ----
    @staticmethod
    def _system_list(terms: np.ndarray, fixed: np.ndarray):
        tm = np.ascontiguousarray(terms, dtype=FLEX_TERM_DTYPE)
        fx = np.ascontiguousarray(fixed, dtype=np.int32)
        L = InteractionList(len(fx), len(tm), tm.ctypes.data_as(ctypes.c_void_p), fx.ctypes.data_as(_P(ctypes.c_int32)))
        return L, (tm, fx)

    @staticmethod
    def _state(a, shape):
        out = np.ascontiguousarray(a if a is not None else np.zeros(shape), dtype=np.float64)
        if out.shape != shape:
            raise ValueError(f"state array of shape {out.shape}, expected {shape}")
        return out.copy() if out is a else out

    def system_eval(self, terms, fixed, positions, forces=None, energies=None):
        """Interaction::eval of every term onto forces / energies (zeros when
        not given) for positions (n_states, n_atoms, 3); returns both."""
        L, keep = self._system_list(terms, fixed)
        pos = self._state(positions, np.shape(positions))
        ns, n = pos.shape[0], pos.shape[1]
        frc, ene = self._state(forces, (ns, n, 3)), self._state(energies, (ns, n))
        self.check(lib().mafm_system_eval(self.h, ctypes.byref(L), ns, _dp(pos), _dp(frc), _dp(ene)))
        return frc, ene

    def system_step(self, terms, fixed, integrator: int, dt: float, positions, velocities, forces=None,
                    energies=None):
        """One eulerStep / midpointStep / rk4Step per state; returns new
        (positions, velocities, forces, energies)."""
        L, keep = self._system_list(terms, fixed)
        pos = self._state(positions, np.shape(positions))
        ns, n = pos.shape[0], pos.shape[1]
        vel = self._state(velocities, (ns, n, 3))
        frc, ene = self._state(forces, (ns, n, 3)), self._state(energies, (ns, n))
        self.check(lib().mafm_system_step(self.h, ctypes.byref(L), integrator, dt, ns, _dp(pos), _dp(vel), _dp(frc),
                                          _dp(ene)))
        return pos, vel, frc, ene

    def system_minimise(self, terms, fixed, positions, velocities, energies=None, minimiser: int = 1,
                        integrator: int = 1, minterm: int = 1, maxsteps: int = 5000, etol: float = 0.01,
                        ftol: float = 0.01, dt: float = 0.01):
        """FIREMinimisation / SDMinimisation from each state; returns
        (positions, velocities, forces, energies, steps)."""
        L, keep = self._system_list(terms, fixed)
        pos = self._state(positions, np.shape(positions))
        ns, n = pos.shape[0], pos.shape[1]
        vel = self._state(velocities, (ns, n, 3))
        frc, ene = np.zeros((ns, n, 3)), self._state(energies, (ns, n))
        steps = np.zeros(ns, dtype=np.int32)
        o = MinimiserOpts(minimiser, integrator, minterm, maxsteps, etol, ftol, dt)
        self.check(lib().mafm_system_minimise(self.h, ctypes.byref(L), ctypes.byref(o), ns, _dp(pos), _dp(vel),
                                              _dp(frc), _dp(ene), steps.ctypes.data_as(_P(ctypes.c_int32))))
        return pos, vel, frc, ene, steps

    def scan_flexible(self, n_atoms: int, positions: np.ndarray, fixed: np.ndarray, terms: np.ndarray,
                      desc: ScanDesc, tip_pbc: bool = False, state_column: int = -1, tip_dummy_d: float = 0.0,
                      cell=None, offset=None) -> dict:
        """The flexible scan over an interaction list (mafm_scan_flexible)."""
        pos = np.ascontiguousarray(positions, dtype=np.float64)
        fx = np.ascontiguousarray(fixed, dtype=np.int32)
        tm = np.ascontiguousarray(terms, dtype=FLEX_TERM_DTYPE)
        d = FlexDesc()
        d.n_atoms, d.n_terms = n_atoms, len(tm)
        d.positions = _dp(pos)
        d.fixed = fx.ctypes.data_as(_P(ctypes.c_int32))
        d.terms = tm.ctypes.data_as(ctypes.c_void_p)
        d.tip_pbc, d.state_column, d.tip_dummy_d = int(tip_pbc), state_column, tip_dummy_d
        for a, v in enumerate(np.asarray(cell if cell is not None else np.zeros(9), dtype=np.float64).ravel()):
            d.cell[a] = v
        for a, v in enumerate(np.asarray(offset if offset is not None else np.zeros(3), dtype=np.float64)):
            d.offset[a] = v
        st = ScanStats()
        self.check(lib().mafm_scan_flexible(self.h, ctypes.byref(d), ctypes.byref(desc), ctypes.byref(st)))
        return st.as_dict()

    def flex_state(self, n_atoms: int, first: int, count: int) -> np.ndarray:
        """(count, n_atoms, 3) relaxed atoms of the state column (mafm_flex_state)."""
        out = np.zeros((count, n_atoms, 3))
        self.check(lib().mafm_flex_state(self.h, first, count, _dp(out)))
        return out

    def eval_tip_surface(self, system: int, pos: np.ndarray) -> np.ndarray:
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        out = np.zeros((len(pos), 4))
        self.check(lib().mafm_eval_tip_surface(self.h, system, len(pos), _dp(pos), _dp(out)))
        return out


class Simulation:
    """One MechAFM job (mafm_sim_*): input files -> device scan -> records / files."""

    def __init__(self, input_path: str, ctx: Optional[Context], outdir: Optional[str] = None,
                 initialize: bool = True):
        """ctx None: a host-only job (parse + interaction list, no device)."""
        self.ctx = ctx
        self.h = ctypes.c_void_p()
        st = lib().mafm_sim_open(input_path.encode(), (outdir or "").encode(), ctx.h if ctx else None,
                                 ctypes.byref(self.h))
        self.check(st)
        if initialize:
            self.check(lib().mafm_sim_initialize(self.h))

    def check(self, st: int) -> None:
        if st != MAFM_OK:
            msg = lib().mafm_sim_last_error(self.h).decode() if self.h else "open failed"
            raise MafmError(st, msg)

    def close(self) -> None:
        if self.h:
            lib().mafm_sim_close(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self) -> dict:
        i = SimInfo()
        self.check(lib().mafm_sim_info_get(self.h, ctypes.byref(i)))
        d = {k: getattr(i, k) for k, _ in i._fields_}
        d["force_grid_n"] = tuple(i.force_grid_n)
        d["offset"] = tuple(i.offset)
        return d

    def system(self) -> SystemDesc:
        d = SystemDesc()
        self.check(lib().mafm_sim_system(self.h, ctypes.byref(d)))
        return d

    def flex_system(self) -> dict:
        """Flexible jobs: the interaction list and the surface
        (mafm_sim_flex_system) as numpy copies."""
        d = FlexDesc()
        self.check(lib().mafm_sim_flex_system(self.h, ctypes.byref(d)))
        n, nt = d.n_atoms, d.n_terms
        terms = np.ctypeslib.as_array(ctypes.cast(d.terms, _P(ctypes.c_uint8)), shape=(nt * FLEX_TERM_DTYPE.itemsize,))
        return {"n_atoms": n, "terms": terms.view(FLEX_TERM_DTYPE).copy(),
                "positions": np.ctypeslib.as_array(d.positions, shape=(n * 3,)).reshape(n, 3).copy(),
                "fixed": np.ctypeslib.as_array(d.fixed, shape=(n,)).copy(), "tip_pbc": d.tip_pbc,
                "state_column": d.state_column, "tip_dummy_d": d.tip_dummy_d, "cell": np.array(d.cell[:]),
                "offset": np.array(d.offset[:])}

    def log(self) -> str:
        """Console lines of the last write_output (mafm_sim_log)."""
        need = ctypes.c_size_t()
        self.check(lib().mafm_sim_log(self.h, None, 0, ctypes.byref(need)))
        buf = ctypes.create_string_buffer(need.value)
        self.check(lib().mafm_sim_log(self.h, buf, need.value, None))
        return buf.value.decode()

    def sites(self) -> List[Site]:
        d = self.system()
        return [d.sites[k] for k in range(d.n_sites)]

    def scan_desc(self, precision: str = "fast64", lanes_per_pixel: int = 0, waves_per_simd: int = 0,
                  compensation: int = 0, waves_per_group: int = 0, options: int = 0) -> ScanDesc:
        o = RunOpts(PREC[precision], lanes_per_pixel, waves_per_simd, compensation, waves_per_group, options)
        d = ScanDesc()
        self.check(lib().mafm_sim_scan_desc(self.h, ctypes.byref(o), ctypes.byref(d)))
        return d

    def run(self, precision: str = "fast64", lanes_per_pixel: int = 0, waves_per_simd: int = 0,
            compensation: int = 0, waves_per_group: int = 0, options: int = 0) -> dict:
        o = RunOpts(PREC[precision], lanes_per_pixel, waves_per_simd, compensation, waves_per_group, options)
        st = ScanStats()
        self.check(lib().mafm_sim_run(self.h, ctypes.byref(o), ctypes.byref(st)))
        return st.as_dict()

    def settings(self) -> str:
        """Input warnings and the settings echo, as the reference prints them."""
        need = ctypes.c_size_t()
        self.check(lib().mafm_sim_settings(self.h, None, 0, ctypes.byref(need)))
        buf = ctypes.create_string_buffer(need.value)
        self.check(lib().mafm_sim_settings(self.h, buf, need.value, None))
        return buf.value.decode()

    def records(self) -> np.ndarray:
        i = self.info()
        n = i["nx"] * i["ny"] * i["nz"]
        out = np.zeros(n, dtype=RECORD_DTYPE)
        self.check(lib().mafm_sim_records(self.h, 0, n, out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def run_output(self, precision: str = "fast64", threads: int = 0, progress=None, lanes_per_pixel: int = 0,
                   waves_per_simd: int = 0, options: int = 0) -> dict:
        """Scan with the files written as the records land
        (mafm_sim_run_output); progress(fraction) is called as they come."""
        o = RunOpts(PREC[precision], lanes_per_pixel, waves_per_simd, 0, 0, options)
        st = ScanStats()
        cb = ctypes.CFUNCTYPE(None, ctypes.c_double, ctypes.c_void_p)(
            (lambda f, _u: progress(f)) if progress else (lambda f, _u: None))
        self.check(lib().mafm_sim_run_output(self.h, ctypes.byref(o), threads, ctypes.cast(cb, ctypes.c_void_p),
                                             None, ctypes.byref(st)))
        return st.as_dict()

    def write_output(self, threads: int = 8) -> None:
        self.check(lib().mafm_sim_write_output(self.h, threads))

    def write_deltaf(self, amplitude: float = 0.0, k_cantilever: float = 0.0, f0: float = 0.0,
                     fz_from_text: bool = True) -> None:
        d = DeltaFDesc(amplitude, k_cantilever, f0, int(fz_from_text), 0)
        self.check(lib().mafm_sim_write_deltaf(self.h, ctypes.byref(d)))


def deltaf_weights(nz: int, zhigh: float, dz: float, amplitude: float = 0.0, k_cantilever: float = 0.0,
                   f0: float = 0.0) -> dict:
    """Window and constants of the frequency-shift images (mafm_deltaf_weights;
    host only, no device needed)."""
    d = DeltaFDesc(amplitude, k_cantilever, f0, 0, 0)
    o = DeltaFWeights()
    rc = lib().mafm_deltaf_weights(None, ctypes.byref(d), nz, zhigh, dz, ctypes.byref(o))
    if rc != MAFM_OK:
        raise MafmError(rc, "mafm_deltaf_weights rejected the window")
    return {"n_weights": o.n_weights, "n_images": o.n_images, "dz": o.dz, "amplitude": o.amplitude,
            "prefactor": o.prefactor, "weights": np.array(o.weights[:o.n_weights])}


def batch_run(sims: Sequence[Simulation], precision: str = "fast32", lanes_per_pixel: int = 0,
              waves_per_simd: int = 0, compensation: int = 0, waves_per_group: int = 0,
              options: int = 0) -> dict:
    arr = (ctypes.c_void_p * len(sims))(*[s.h.value for s in sims])
    o = RunOpts(PREC[precision], lanes_per_pixel, waves_per_simd, compensation, waves_per_group, options)
    st = ScanStats()
    rc = lib().mafm_batch_run(len(sims), arr, ctypes.byref(o), ctypes.byref(st))
    if rc != MAFM_OK:
        raise MafmError(rc, lib().mafm_sim_last_error(sims[0].h).decode())
    return st.as_dict()


def batch_run_output(sims: Sequence[Simulation], precision: str = "fast64", threads: int = 0,
                     lanes_per_pixel: int = 0, waves_per_simd: int = 0, compensation: int = 0,
                     waves_per_group: int = 0, options: int = 0) -> dict:
    """A batch run that writes each job's files as soon as its columns are
    scanned, while the rest of the batch scans (mafm_batch_run_output)."""
    arr = (ctypes.c_void_p * len(sims))(*[s.h.value for s in sims])
    o = RunOpts(PREC[precision], lanes_per_pixel, waves_per_simd, compensation, waves_per_group, options)
    st = ScanStats()
    rc = lib().mafm_batch_run_output(len(sims), arr, ctypes.byref(o), threads, ctypes.byref(st))
    if rc != MAFM_OK:
        raise MafmError(rc, lib().mafm_sim_last_error(sims[0].h).decode())
    return st.as_dict()


def batch_write_output(sims: Sequence[Simulation], threads: int = 0) -> None:
    """The per-height files of many jobs that have run, on one thread pool
    (mafm_batch_write_output)."""
    arr = (ctypes.c_void_p * len(sims))(*[s.h.value for s in sims])
    rc = lib().mafm_batch_write_output(len(sims), arr, threads)
    if rc != MAFM_OK:
        raise MafmError(rc, lib().mafm_sim_last_error(sims[0].h).decode())


def records_to_table(rec: np.ndarray, shape) -> np.ndarray:
    """Expand device records to the 16-column layout of the reference's
    OutputData (i j k n x y z fx fy fz rx ry rz r angle e); r and angle use
    System::getOutput's formulas (system.cpp:28-38)."""
    nx, ny, nz = shape
    idx = np.arange(len(rec))
    k = idx % nz
    j = (idx // nz) % ny
    i = idx // (nz * ny)
    t = np.zeros((len(rec), 16))
    t[:, 0], t[:, 1], t[:, 2], t[:, 3] = i, j, k, rec["n_steps"]
    t[:, 4], t[:, 5], t[:, 6] = rec["x"], rec["y"], rec["z"]
    t[:, 7], t[:, 8], t[:, 9] = rec["fx"], rec["fy"], rec["fz"]
    t[:, 10], t[:, 11], t[:, 12] = rec["rx"], rec["ry"], rec["rz"]
    t[:, 13] = np.sqrt((rec["rx"] ** 2 + rec["ry"] ** 2) + rec["rz"] ** 2)
    t[:, 14] = np.arctan2(np.sqrt(rec["rx"] ** 2 + rec["ry"] ** 2), rec["rz"]) * (180.0 / 3.14159265358979323846)
    t[:, 15] = rec["e"]
    return t
