import gzip
import os
import shutil
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


def golden_cases():
    """Rigid cases with full-precision reference records (tests/golden/<case>/records.npz)."""
    return sorted(d for d in os.listdir(GOLDEN)
                  if os.path.exists(os.path.join(GOLDEN, d, "records.npz")) and not d.startswith("fx_"))


def flex_cases():
    """Flexible-on cases (tests/golden/fx_*): records, the centre column's
    relaxed atoms per height and, for some, whole-list forces."""
    return sorted(d for d in os.listdir(GOLDEN) if d.startswith("fx_"))


def flex_golden(case: str):
    d = np.load(os.path.join(GOLDEN, case, "records.npz"))
    return tuple(int(v) for v in d["shape"]), d["records"], d["state"]


def materialize(case: str, dst: str) -> str:
    """Copy a golden case's input files into dst (gunzipping cubes); returns
    the path of its input.scan."""
    src = os.path.join(GOLDEN, case)
    os.makedirs(dst, exist_ok=True)
    for fn in os.listdir(src):
        p = os.path.join(src, fn)
        if fn.endswith(".npz"):
            continue
        if fn.endswith(".gz"):
            with gzip.open(p, "rb") as fi, open(os.path.join(dst, fn[:-3]), "wb") as fo:
                fo.write(fi.read())
        else:
            shutil.copy(p, os.path.join(dst, fn))
    return os.path.join(dst, "input.scan")


def golden_records(case: str):
    d = np.load(os.path.join(GOLDEN, case, "records.npz"))
    return tuple(int(v) for v in d["shape"]), d["records"]


def golden_eval(case: str):
    p = os.path.join(GOLDEN, case, "eval.npz")
    if not os.path.exists(p):
        return None
    d = np.load(p)
    return d["pos"], d["fe"]


@pytest.fixture(scope="session")
def gpu_ctx():
    from mechafm_amd import abi
    ctx = abi.Context(0)
    yield ctx
    ctx.close()


SCALE = os.path.join(GOLDEN, "scale")


def materialize_scale(case: str, dst: str):
    """Regenerate a scale case's inputs with the seeded generator its
    inputs.json names and check every file against the SHA-256 the
    reference read; returns the input.scan paths (one per system)."""
    import hashlib
    import json
    from mechafm_amd import workloads as W
    spec = json.load(open(os.path.join(SCALE, case, "inputs.json")))
    wls = getattr(W, spec["generator"])(dst, **spec["kwargs"])
    wls = wls if isinstance(wls, list) else [wls]
    assert len(wls) == len(spec["systems"])
    for wl, sysd in zip(wls, spec["systems"]):
        folder = os.path.dirname(wl.input_path)
        for fn, digest in sysd["files"].items():
            got = hashlib.sha256(open(os.path.join(folder, fn), "rb").read()).hexdigest()
            assert got == digest, f"{case}: regenerated {sysd['folder']}/{fn} differs from the reference's input"
    return [wl.input_path for wl in wls]


def scale_records(case: str):
    d = np.load(os.path.join(SCALE, case, "records.npz"))
    return [tuple(int(v) for v in s) for s in d["shapes"]], d["records"]
