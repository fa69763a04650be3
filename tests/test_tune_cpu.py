"""Autotuner logic without a GPU: the search order, the choice, rejected
shapes, and the cache (a fake runner stands in for the scans)."""
import json
from types import SimpleNamespace

import pytest

from mechafm_amd import abi, tune


class FakeSim:
    def __init__(self, n_sites, n_images=1):
        self.n_sites, self.n_images = n_sites, n_images

    def info(self):
        return {"nx": 64, "ny": 64, "nz": 40}

    def scan_desc(self, precision):
        return SimpleNamespace(minimiser=1, integrator=1)

    def system(self):
        return SimpleNamespace(n_sites=self.n_sites, n_images=self.n_images, force_grid=-1, potential_grid=-1)


def fake_runner(times, calls):
    def run(sims, precision, lanes_per_pixel, waves_per_simd, **kw):
        calls.append((lanes_per_pixel, waves_per_simd))
        t = times.get((lanes_per_pixel, waves_per_simd))
        if t is None:
            raise abi.MafmError(1, "rejected shape")
        return {"kernel_ms": t}
    return run


def test_search_picks_fastest_and_caches(tmp_path):
    sims = [FakeSim(120), FakeSim(300)]
    # G and waves per SIMD searched jointly: the best shape (1, 3) is one a
    # search of G at two waves alone would miss (G = 2 wins there)
    times = {(1, 2): 50.0, (1, 3): 36.0, (2, 2): 40.0, (2, 3): 39.0, (4, 2): 45.0, (4, 3): 44.0,
             (8, 2): 60.0, (8, 3): 58.0, (16, 2): 90.0, (16, 3): 85.0, (1, 1): 70.0}
    calls = []
    cache = str(tmp_path / "tune.json")
    best = tune.tune(sims, "fast32", cache=cache, run=fake_runner(times, calls))
    assert (best["lanes_per_pixel"], best["waves_per_simd"]) == (1, 3) and not best["cached"]
    assert calls == [(1, 2), (1, 3), (2, 2), (2, 3), (4, 2), (4, 3), (8, 2), (8, 3), (16, 2), (16, 3), (1, 1)]
    again = tune.tune(sims, "fast32", cache=cache, run=fake_runner(times, calls))
    assert again["cached"] and again["lanes_per_pixel"] == 1 and len(calls) == 11
    assert len(json.load(open(cache))) == 1


def test_rejected_shapes_are_skipped(tmp_path):
    times = {(1, 2): 70.0, (1, 1): 80.0, (1, 3): 75.0}       # G > 1 rejected (molecule beyond LDS)
    best = tune.tune([FakeSim(5000)], "fast32", cache=str(tmp_path / "t.json"), run=fake_runner(times, []))
    assert best["lanes_per_pixel"] == 1 and best["waves_per_simd"] == 2


def test_signature_separates_workloads():
    a = tune.signature([FakeSim(120)], "fast32")
    assert a == tune.signature([FakeSim(121)], "fast32")            # same two significant digits
    assert a != tune.signature([FakeSim(400)], "fast32")
    assert a != tune.signature([FakeSim(120)], "fast64")
    assert a != tune.signature([FakeSim(120, 9)], "fast32")
    # 5080 and 5100 FP32 sites round alike, but only the first fits the LDS a
    # workgroup may take (160 KiB less 1 KiB of static LDS)
    assert tune.signature([FakeSim(5080)], "fast32") != tune.signature([FakeSim(5100)], "fast32")
    # on a device that reports 64 KiB per workgroup the split moves with it
    small = {"name": "x", "compute_units": 1, "lds_per_block": 64 * 1024}
    assert tune.signature([FakeSim(2000)], "fast32", device=small) != \
        tune.signature([FakeSim(2049)], "fast32", device=small)


def test_plain_has_nothing_to_tune():
    with pytest.raises(ValueError):
        tune.tune([FakeSim(10)], "plain64", run=fake_runner({}, []))
