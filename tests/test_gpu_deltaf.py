"""Frequency-shift images on the GPU (mafm_delta_f) against the oracle's
restatement of simple-vafm-image.py, bit for bit on the same Fz."""
import numpy as np
import pytest

from conftest import golden_records, materialize
from oracle import deltaf as D

pytestmark = pytest.mark.gpu


def _info(sim):
    i = sim.info()
    return i["nx"], i["ny"], i["nz"]


@pytest.mark.parametrize("precision", ["plain64", "fast32"])
def test_delta_f_matches_oracle(gpu_ctx, tmp_path, precision):
    from mechafm_amd import abi
    inp = materialize("c1_mini", str(tmp_path))
    sim = abi.Simulation(inp, gpu_ctx)
    sim.run(precision)
    nx, ny, nz = _info(sim)
    d = sim.scan_desc(precision)
    df, zlow = gpu_ctx.delta_f(0, nx, ny, nz, d.zhigh, d.dz)
    rec = sim.records()
    zpos = D.file_heights(d.zhigh, d.dz, nz)
    want, wz = D.images(D.fz_from_records(rec["fz"], nx, ny, nz), zpos)
    assert df.shape == want.shape and df.shape[0] == nz - 8
    np.testing.assert_array_equal(df, want)
    np.testing.assert_array_equal(zlow, wz)
    if precision == "plain64":
        # the reference's own records give the same images to rounding
        _, ref = golden_records("c1_mini")
        fz = np.zeros(nx * ny * nz)
        fz[((ref[:, 0] * ny + ref[:, 1]) * nz + ref[:, 2]).astype(int)] = ref[:, 9]
        ref_img, _ = D.images(D.fz_from_records(fz, nx, ny, nz), zpos)
        np.testing.assert_allclose(df, ref_img, rtol=1e-9, atol=1e-9)


def test_delta_f_of_batched_systems(gpu_ctx, tmp_path):
    """Images of each system of a batched launch come from its own records;
    other amplitudes and cantilevers change the window as in the oracle."""
    from mechafm_amd import abi, workloads as W
    wls = W.make_c3(str(tmp_path), mini=True, count=3)
    sims = [abi.Simulation(w.input_path, gpu_ctx) for w in wls]
    abi.batch_run(sims, "fast32")
    nx, ny, nz = _info(sims[0])
    d = sims[0].scan_desc("fast32")
    zpos = D.file_heights(d.zhigh, d.dz, nz)
    for m, s in enumerate(sims):
        for amp, k, f0 in ((0.0, 0.0, 0.0), (0.3, 35.0, 150e3)):
            df, _ = gpu_ctx.delta_f(m, nx, ny, nz, d.zhigh, d.dz, amp, k, f0)
            kw = {}
            if amp:
                kw = dict(amplitude=amp, k_cantilever=k, frequency0=f0)
            want, _ = D.images(D.fz_from_records(s.records()["fz"], nx, ny, nz), zpos, **kw)
            np.testing.assert_array_equal(df, want)


def test_deltaf_files(gpu_ctx, tmp_path):
    """deltaf-%06.3f.dat per image, named by the window's lowest height."""
    import os
    from mechafm_amd import abi
    inp = materialize("c1_mini", str(tmp_path / "in"))
    out = str(tmp_path / "out") + "/"
    os.makedirs(out)
    sim = abi.Simulation(inp, gpu_ctx, outdir=out)
    sim.run("fast64")
    sim.write_deltaf()
    nx, ny, nz = _info(sim)
    d = sim.scan_desc("fast64")
    zpos = D.file_heights(d.zhigh, d.dz, nz)
    want, wz = D.images(D.fz_from_records(sim.records()["fz"], nx, ny, nz), zpos)
    files = sorted(f for f in os.listdir(out) if f.startswith("deltaf-"))
    assert files == sorted("deltaf-%06.3f.dat" % z for z in wz)
    rows = np.loadtxt(out + "deltaf-%06.3f.dat" % wz[3], comments="%")
    assert rows.shape == (nx * ny, 5)
    np.testing.assert_allclose(rows[:, 4], want[3].ravel(), rtol=1e-6, atol=1e-12)
