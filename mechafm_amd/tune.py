"""Launch-shape autotuner for the fast scan kernels.

The fast kernels take three launch parameters (mafm_scan_desc):
lanes_per_pixel G (lanes that split one pixel's sites), waves_per_simd (size
of the persistent grid) and waves_per_group (waves sharing one LDS copy of a
molecule's sites). The best G depends on the workload: few pixel columns or
large molecules want more lanes per pixel, many small molecules fewer. The
tuner measures candidates on the workload itself (whole launches, kernel
time from HIP events) and caches the winner under a signature of the
workload and device, so a repeated workload is tuned once.

Search: every G jointly with 2 and 3 waves per SIMD (the FIRE + midpoint
kernels are built for three resident waves, the others for two), then one
wave per SIMD for the best G (11 launches by default); waves_per_group stays
automatic.
"""
from __future__ import annotations

import json
import math
import os
from typing import Callable, Dict, Iterable, List, Optional, Sequence, Tuple

from . import abi

DEFAULT_LANES = (1, 2, 4, 8, 16)
WAVES = (2, 3)
LDS_BYTES = 160 * 1024                     # LDS one workgroup may take (gfx950), when the device is not given
SITE_BYTES = {"fast32": 32, "fast64": 64}  # bytes per site in LDS
CACHE_ENV = "MAFM_TUNE_CACHE"


def _default_cache() -> str:
    return os.environ.get(CACHE_ENV) or os.path.join(os.path.expanduser("~"), ".cache", "mechafm_amd", "tune.json")


def _round2(x: float) -> float:
    """Two significant digits: workloads of similar size share an entry."""
    if x <= 0:
        return 0.0
    e = math.floor(math.log10(x)) - 1
    return round(x / 10 ** e) * 10 ** e


def signature(sims: Sequence[abi.Simulation], precision: str, compensation: int = 0, options: int = 0,
              device: Optional[dict] = None) -> str:
    """Key of a workload: device, arithmetic, scan geometry, minimiser, the
    size distribution of the systems (sites per system rounded to two
    significant digits) and whether every system's sites fit the LDS copy a
    workgroup may take (the device's per-workgroup LDS less 1 KiB; otherwise
    the kernel streams them with one lane per pixel, so a shape tuned on a
    system that fits does not carry over)."""
    info = sims[0].info()
    d = sims[0].scan_desc(precision)
    systems = []
    lds_fit = True
    lds_limit = int((device or {}).get("lds_per_block") or LDS_BYTES) - 1024
    for s in sims:
        sd = s.system()
        lds_fit &= sd.force_grid >= 0 or sd.n_sites * SITE_BYTES.get(precision, 64) <= lds_limit
        systems.append((int(_round2(sd.n_sites)), int(sd.n_images or 1), int(sd.force_grid >= 0),
                        int(sd.potential_grid >= 0)))
    systems.sort()
    key = {
        "device": (device or {}).get("name", ""), "cus": (device or {}).get("compute_units", 0),
        "precision": precision, "compensation": compensation, "options": options,
        "grid": [info["nx"], info["ny"], info["nz"]], "minimiser": d.minimiser, "integrator": d.integrator,
        "systems": len(sims), "sites_total": _round2(sum(x[0] for x in systems)),
        "sites_max": max(x[0] for x in systems), "images": max(x[1] for x in systems),
        "force_grid": any(x[2] for x in systems), "potential_grid": any(x[3] for x in systems),
        "lds_fit": lds_fit,
    }
    return json.dumps(key, sort_keys=True)


def _load(path: str) -> Dict[str, dict]:
    try:
        with open(path) as f:
            data = json.load(f)
        return data if isinstance(data, dict) else {}
    except (OSError, ValueError):
        return {}


def _store(path: str, key: str, value: dict) -> None:
    try:
        os.makedirs(os.path.dirname(path), exist_ok=True)
        data = _load(path)
        data[key] = value
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(data, f, indent=1, sort_keys=True)
        os.replace(tmp, path)
    except OSError:
        pass                       # a read-only home only costs a re-tune


def run_batch(sims: Sequence[abi.Simulation], precision: str, **kw) -> dict:
    if len(sims) == 1:
        return sims[0].run(precision, **kw)
    return abi.batch_run(list(sims), precision, **kw)


def tune(sims: Sequence[abi.Simulation], precision: str = "fast32", compensation: int = 0, options: int = 0,
         lanes: Iterable[int] = DEFAULT_LANES, cache: Optional[str] = None, use_cache: bool = True,
         run: Optional[Callable[..., dict]] = None, device: Optional[dict] = None,
         log: Optional[Callable[[str], None]] = None) -> dict:
    """Best {lanes_per_pixel, waves_per_simd, kernel_ms} for this workload.

    ``run(sims, precision, **launch)`` performs one scan and returns its
    statistics (defaults to a real launch); shapes the library rejects (a
    molecule too large for LDS with G > 1) are skipped."""
    if precision == "plain64":
        raise ValueError("the plain kernel has no launch shape to tune")
    run = run or run_batch
    path = cache or _default_cache()
    key = signature(sims, precision, compensation, options, device)
    if use_cache:
        hit = _load(path).get(key)
        if hit:
            return dict(hit, cached=True)
    tried: List[Tuple[int, int, float]] = []

    def measure(g: int, w: int) -> Optional[float]:
        try:
            st = run(sims, precision, lanes_per_pixel=g, waves_per_simd=w, compensation=compensation,
                     options=options)
        except abi.MafmError as e:
            if log:
                log(f"G={g} waves/SIMD={w}: skipped ({e})")
            return None
        ms = float(st["kernel_ms"])
        tried.append((g, w, ms))
        if log:
            log(f"G={g} waves/SIMD={w}: {ms:.1f} ms")
        return ms

    for g in lanes:
        for w in WAVES:
            measure(g, w)
    if not tried:
        raise RuntimeError("no launch shape was accepted for this workload")
    g_best = min(tried, key=lambda t: t[2])[0]
    measure(g_best, 1)
    g, w, ms = min(tried, key=lambda t: t[2])
    best = {"lanes_per_pixel": g, "waves_per_simd": w, "kernel_ms": ms,
            "tried": [{"lanes_per_pixel": a, "waves_per_simd": b, "kernel_ms": c} for a, b, c in tried]}
    _store(path, key, best)
    return dict(best, cached=False)
