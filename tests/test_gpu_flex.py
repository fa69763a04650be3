"""Flexible on, on the GPU (mafm_scan_flexible through the C ABI).

Every flexible golden case runs through the job layer (mafm_sim_run) and is
compared with the reference's records and the centre column's relaxed atoms:

* the FIRE / SD iteration count of every position is the reference's;
* forces, energies, tip offsets and relaxed atoms agree to 1e-6 (measured
  ~1e-10: the device's acos / sin / atan / exp differ from glibc's in the
  last place, everything else is the reference's operation order) -- far
  inside the stated bar of 1e-4 A on tip positions and 1e-3 relative on Fz,
  which is asserted as well.

The C ABI is also driven directly with the oracle's list on a perturbed
surface (compared with the oracle's scan), the reference's example job runs
at full size against its golden, and the state files / delta f / error paths
are checked.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN, flex_cases, flex_golden, materialize
from mechafm_amd import abi

pytestmark = pytest.mark.gpu

CASES = flex_cases()
TOL = 1e-6


def _check(rec, ref, state=None, ref_state=None):
    assert rec.shape == ref.shape
    np.testing.assert_array_equal(rec[:, :4], ref[:, :4])          # i j k and the iteration count
    np.testing.assert_array_equal(rec[:, 4:6], ref[:, 4:6])        # pixel x, y
    np.testing.assert_allclose(rec[:, 6:], ref[:, 6:], rtol=0, atol=TOL)
    # the stated bar: tip position 1e-4 A, Fz 1e-3 relative (abs floor 1e-6)
    assert np.abs(rec[:, 10:13] - ref[:, 10:13]).max() <= 1e-4
    assert np.all(np.abs(rec[:, 9] - ref[:, 9]) <= 1e-3 * np.abs(ref[:, 9]) + 1e-6)
    if state is not None:
        np.testing.assert_allclose(state, ref_state, rtol=0, atol=TOL)


@pytest.mark.parametrize("case", CASES)
def test_flexible_scan_matches_the_reference(gpu_ctx, tmp_path, case):
    shape, ref, ref_state = flex_golden(case)
    sim = abi.Simulation(materialize(case, str(tmp_path)), gpu_ctx)
    st = sim.run("plain64")
    rec = abi.records_to_table(sim.records(), shape)
    state = gpu_ctx.flex_state(ref_state.shape[1], 0, shape[2])
    _check(rec, ref, state, ref_state)
    assert st["total_steps"] == int(ref[:, 3].sum())
    assert st["positions"] == len(ref)
    assert st["unconverged"] == int((ref[:, 3] >= int(_opt(tmp_path, "maxsteps", 5000))).sum())
    sim.close()


def _opt(folder, key, default):
    for line in open(os.path.join(str(folder), "input.scan")):
        t = line.split()
        if t and t[0] == key:
            return t[1]
    return default


def test_c_abi_on_a_perturbed_surface_matches_the_oracle(gpu_ctx, tmp_path):
    """mafm_scan_flexible called directly with the oracle's list and a
    seeded perturbation of the surface (new geometry, same list)."""
    from oracle import oracle_flex as OF
    fs = OF.flex_setup(materialize("fx_mini", str(tmp_path)))
    pos = fs.base.system.pos.copy()
    rng = np.random.default_rng(11)
    pos[2:] += rng.normal(0.0, 0.03, pos[2:].shape)
    nx, ny, nz = fs.base.scan_shape
    ref, ref_state, steps = OF.scan(fs, pos=pos)
    sim = abi.Simulation(os.path.join(str(tmp_path), "input.scan"), None)
    desc = sim.scan_desc("plain64")
    s = fs.base.system
    st = gpu_ctx.scan_flexible(len(pos), pos, fs.fixed, fs.terms, desc, tip_pbc=s.tip_pbc,
                               state_column=nx * ny // 2, tip_dummy_d=s.tip_dummy_d, cell=s.cell, offset=s.offset)
    rec = abi.records_to_table(gpu_ctx.records(0, nx * ny * nz), (nx, ny, nz))
    _check(rec, ref, gpu_ctx.flex_state(len(pos), 0, nz), ref_state)
    assert st["total_steps"] == steps


def test_state_in_hbm_matches(gpu_ctx, tmp_path):
    """The column state in HBM (the path of systems too large for one
    workgroup's LDS, forced here by MAFM_SCAN_STATE_IN_HBM) gives the same
    records; a state column of -1 keeps no state."""
    shape, ref, _ = flex_golden("fx_graphene")
    sim = abi.Simulation(materialize("fx_graphene", str(tmp_path)), None)
    fx = sim.flex_system()
    desc = sim.scan_desc("plain64", options=abi.SCAN_STATE_IN_HBM)
    gpu_ctx.scan_flexible(fx["n_atoms"], fx["positions"], fx["fixed"], fx["terms"], desc,
                          state_column=-1, tip_dummy_d=fx["tip_dummy_d"], cell=fx["cell"], offset=fx["offset"])
    rec = abi.records_to_table(gpu_ctx.records(0, len(ref)), shape)
    _check(rec, ref)
    with pytest.raises(abi.MafmError, match="no flexible state"):
        gpu_ctx.flex_state(fx["n_atoms"], 0, 1)


def test_example_at_full_size(gpu_ctx, tmp_path):
    """The reference's example job with flexible on at its own size: 41 x 31
    columns x 41 heights, 12.9 M FIRE iterations.

    Low over the sheet the flexible relaxation is chaotic: the reference's own
    answer moves when one input coordinate changes in its last printed digit
    (1e-5 A) -- a fifth of the positions change their iteration count and Fz
    moves by up to ~0.5 there (the golden's sensitivity profile).  The device's
    last-place libm differences act the same way.  So, height by height: where
    the reference is stable (every count kept under the input change) the
    counts are the reference's and Fz agrees to 1e-6; below, the GPU's
    deviation from the reference is no larger than the reference's own under
    that input change (count agreement within 5 points, |dFz| quantiles within
    2x)."""
    from mechafm_amd import workloads as W
    g = np.load(os.path.join(GOLDEN, "scale", "fx_example_full", "records.npz"))
    sens = g["sensitivity"]
    wl = W.make_fx_example(str(tmp_path))
    sim = abi.Simulation(wl.input_path, gpu_ctx)
    st = sim.run("plain64")
    r = sim.records()
    nz = int(g["shape"][2])
    assert len(r) == len(g["n_steps"]) == st["positions"]
    k = np.arange(len(r)) % nz
    same = r["n_steps"] == g["n_steps"]
    dfz = np.abs(r["fz"] - g["fz"])
    for kk in range(nz):
        m = k == kk
        if sens[kk, 0] == 1.0:
            assert same[m].all(), f"height {kk}"
            assert dfz[m].max() <= TOL, f"height {kk}"
        else:
            assert same[m].mean() >= sens[kk, 0] - 0.05, f"height {kk}"
            q = np.quantile(dfz[m], [0.5, 0.9, 0.99])
            assert np.all(q <= 2 * sens[kk, 1:4] + 1e-9), f"height {kk}: {q} vs {sens[kk, 1:4]}"
    state = gpu_ctx.flex_state(g["state"].shape[1], 0, nz)
    assert np.abs(state - g["state"]).max() <= 1e-4


def test_state_files_and_log(gpu_ctx, tmp_path):
    """System::makeXYZFile (system.cpp:54-65) of the centre column at every
    height, with its console line."""
    shape, ref, ref_state = flex_golden("fx_graphene")
    out = tmp_path / "out"
    sim = abi.Simulation(materialize("fx_graphene", str(tmp_path / "in")), gpu_ctx, outdir=str(out))
    sim.run("plain64")
    sim.write_output(threads=4)
    nx, ny, nz = shape
    col = nx * ny // 2
    lines = sim.log().splitlines()
    assert len(lines) == nz
    types = [l.split()[0] for l in open(tmp_path / "in" / "graphene.xyz").read().splitlines()[2:] if l.split()]
    types = ["X", "O"] + types
    for k in range(nz):
        r = ref[col * nz + k]
        name = f"{out}/state_{r[4]:.1f}-{r[5]:.1f}-{ref_state[k, 0, 2]:.1f}.xyz"
        assert lines[k] == "+- Wrote " + name
    last = ref[col * nz + nz - 1]
    name = f"{out}/state_{last[4]:.1f}-{last[5]:.1f}-{ref_state[-1, 0, 2]:.1f}.xyz"
    body = open(name).read()
    expect = f"{len(types)}\n\n" + "".join(
        f"{t} {p[0]:8.4f} {p[1]:8.4f} {p[2]:8.4f}\n" for t, p in zip(types, ref_state[-1]))
    assert body == expect
    assert sorted(f for f in os.listdir(out) if f.startswith("scan-")) == \
        sorted(f"scan-{9.0 + 0.1 * k:06.3f}.dat" for k in range(nz))


def test_delta_f_of_a_flexible_scan(gpu_ctx, tmp_path):
    from conftest import materialize as mat
    sim = abi.Simulation(mat("fx_graphene", str(tmp_path)), gpu_ctx)
    sim.run("plain64")
    i = sim.info()
    df, zl = gpu_ctx.delta_f(0, i["nx"], i["ny"], i["nz"], 10.0, 0.1, amplitude=0.2)
    assert df.shape[1:] == (i["nx"], i["ny"]) and np.all(np.isfinite(df))
    with pytest.raises(abi.MafmError):
        gpu_ctx.delta_f(1, i["nx"], i["ny"], i["nz"], 10.0, 0.1, amplitude=0.2)


def test_bad_term_lists_are_refused(gpu_ctx, tmp_path):
    sim = abi.Simulation(materialize("fx_mini", str(tmp_path)), None)
    fx = sim.flex_system()
    desc = sim.scan_desc("plain64")

    def run(terms, fixed=None):
        gpu_ctx.scan_flexible(fx["n_atoms"], fx["positions"], fx["fixed"] if fixed is None else fixed, terms, desc,
                              tip_dummy_d=fx["tip_dummy_d"])

    t = fx["terms"].copy()
    t["atom"][100, 1] = fx["n_atoms"]
    with pytest.raises(abi.MafmError, match="outside"):
        run(t)
    t = fx["terms"].copy()
    t["kind"][5] = 42
    with pytest.raises(abi.MafmError, match="unknown kind"):
        run(t)
    t = np.concatenate([fx["terms"], fx["terms"][:1]])
    with pytest.raises(abi.MafmError, match="tip-surface term after"):
        run(t)
    f = fx["fixed"].copy()
    f[1] = 1
    with pytest.raises(abi.MafmError, match="tip"):
        run(fx["terms"], f)
