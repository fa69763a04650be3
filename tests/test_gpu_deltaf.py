"""Frequency-shift images on the GPU (mafm_delta_f) against the oracle's
restatement of simple-vafm-image.py, bit for bit on the same Fz -- and, with
fz_from_text, on the script's own input: the Fz column of the reference
program's per-height files for its example job."""
import gzip
import os
import shutil

import numpy as np
import pytest

from conftest import GOLDEN, golden_records, materialize
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
    """deltaf-%06.3f.dat per image, named by the window's lowest height,
    from Fz as the script reads it (the job-level default)."""
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
    want, wz = D.images(D.fz_text(D.fz_from_records(sim.records()["fz"], nx, ny, nz)), zpos)
    files = sorted(f for f in os.listdir(out) if f.startswith("deltaf-"))
    assert files == sorted("deltaf-%06.3f.dat" % z for z in wz)
    rows = np.loadtxt(out + "deltaf-%06.3f.dat" % wz[3], comments="%")
    assert rows.shape == (nx * ny, 5)
    np.testing.assert_allclose(rows[:, 4], want[3].ravel(), rtol=1e-6, atol=1e-12)


def test_delta_f_from_the_reference_files(gpu_ctx, tmp_path, monkeypatch):
    """The script reads Fz from the "%8.4f" files.  On the reference's example
    job (plain64, whose lines equal the reference program's): the records'
    Fz printed and parsed back equals the reference files' Fz column, and the
    device images with fz_from_text equal the oracle's images of the
    reference files bit for bit -- what the script computes from them.  The
    images of the unrounded Fz stay within the bound the 5e-5 rounding of
    each Fz allows."""
    from mechafm_amd import abi
    src = os.path.join(GOLDEN, "example_graphene")
    job = tmp_path / "job"
    job.mkdir()
    for fn in ("input.scan", "graphene.xyz", "parameters.dat"):
        shutil.copy(os.path.join(src, fn), job / fn)
    files, name = {}, None
    with gzip.open(os.path.join(src, "outputs.txt.gz"), "rt") as f:
        for line in f.read().splitlines():
            if line.startswith("## "):
                name = line[3:]
                files[name] = []
            else:
                files[name].append(line)
    monkeypatch.chdir(job)
    sim = abi.Simulation("input.scan", gpu_ctx, outdir=str(tmp_path / "out"))
    sim.run("plain64")
    nx, ny, nz = _info(sim)
    d = sim.scan_desc("plain64")
    fz_ref = D.fz_from_files(files, nx, ny)
    np.testing.assert_array_equal(D.fz_text(D.fz_from_records(sim.records()["fz"], nx, ny, nz)), fz_ref)
    zpos = D.file_heights(d.zhigh, d.dz, nz)
    want, wz = D.images(fz_ref, zpos)
    df, zl = gpu_ctx.delta_f(0, nx, ny, nz, d.zhigh, d.dz, fz_from_text=True)
    np.testing.assert_array_equal(df, want)
    np.testing.assert_array_equal(zl, wz)
    raw, _ = gpu_ctx.delta_f(0, nx, ny, nz, d.zhigh, d.dz)
    w = abi.deltaf_weights(nz, d.zhigh, d.dz)
    bound = w["prefactor"] * w["dz"] * np.abs(w["weights"]).sum() * 0.5e-4
    assert np.abs(raw - want).max() <= bound * (1 + 1e-9)
    print(f"example job: {df.shape[0]} images; unrounded Fz moves them by at most "
          f"{np.abs(raw - want).max():.3g} Hz (bound {bound:.3g} Hz)")
