"""GPU parity of the scan against the reference's golden fixtures and the
oracle, through the C ABI (libmafm).

Bars (stated per precision):
* plain64 -- the reference's arithmetic: identical FIRE step counts, records
  equal to 1e-9 (observed ~1e-14), and bit for bit on the LJ / Coulomb cases;
* fast64  -- FP64 with refined hardware reciprocals: tip positions within
  1e-4 A and Fz within 1e-3 relative (relative to max(|Fz|, 1e-2)) for every
  position;
* fast32  -- FP32 pair terms, FP64 minimiser: the same 1e-4 A / 1e-3 bar on
  at least 99 % of the positions, run in the reproducible mode so the
  fraction is the same on every run.  Not on every position: the FP32 error
  of the tip force (~1e-6 kcal/mol/A, the difference of the large
  tip-surface and tip-dummy forces) is of the order of the distance of a
  FIRE iteration's |F| from ftol, so some convergence tests come out the
  other way than on the FP64 trajectory; such a position moves by one FIRE
  step and, through the warm start, takes the rest of its column along.
"""
import os

import numpy as np
import pytest

from conftest import GOLDEN, golden_cases, golden_eval, golden_records, materialize

pytestmark = pytest.mark.gpu

CASES = golden_cases()


def _run(ctx, case, tmp_path, precision, **kw):
    from mechafm_amd import abi
    inp = materialize(case, str(tmp_path))
    sim = abi.Simulation(inp, ctx)
    stats = sim.run(precision, **kw)
    info = sim.info()
    shape = (info["nx"], info["ny"], info["nz"])
    tab = abi.records_to_table(sim.records(), shape)
    return sim, tab, stats, shape


def _errors(tab, ref):
    dpos = np.abs(tab[:, 10:13] - ref[:, 10:13]).max(axis=1)
    dfz = np.abs(tab[:, 9] - ref[:, 9]) / np.maximum(np.abs(ref[:, 9]), 1e-2)
    return dpos, dfz


def _bitwise_case(case, tmp_path):
    """LJ / Coulomb cases without an external potential: the plain kernel
    uses only IEEE +, -, *, / and sqrt in the reference's order, so its
    records equal the reference's bit for bit.  Morse terms call exp (the
    device's and glibc's differ in the last place) and the external
    potential goes through an FFT, so those cases are held to 1e-9."""
    d = os.path.join(str(tmp_path), "inputs")
    materialize(case, d)
    params = open(os.path.join(d, "parameters.dat")).read().lower()
    scan = open(os.path.join(d, "input.scan")).read().lower().split("\n")
    potential = any(l.split()[:2] == ["use_external_potential", "on"] for l in scan)
    return "morse" not in params and not potential


@pytest.mark.parametrize("case", CASES)
def test_plain64_matches_reference(gpu_ctx, case, tmp_path):
    sim, tab, stats, shape = _run(gpu_ctx, case, tmp_path, "plain64")
    gshape, ref = golden_records(case)
    assert shape == gshape
    np.testing.assert_array_equal(tab[:, 3], ref[:, 3])
    scale = max(1.0, np.abs(ref).max())
    assert np.abs(tab - ref).max() <= 1e-9 * scale
    assert stats["total_steps"] == int(ref[:, 3].sum())
    if _bitwise_case(case, tmp_path):
        # every column the kernel produces (the angle is derived here with numpy)
        cols = [c for c in range(16) if c != 14]
        np.testing.assert_array_equal(tab[:, cols], ref[:, cols])


@pytest.mark.parametrize("precision", ["plain64", "fast64"])
def test_reference_example_files(gpu_ctx, tmp_path, precision, monkeypatch):
    """The reference's own example job (examples/input.scan: 72-atom graphene,
    41 x 31 px x 41 heights, FIRE + midpoint to |F| < 1e-3) through the
    job-level API: every per-height file equals the one the reference program
    writes, line for line once sorted (the reference writes lines in thread
    order), and the step total is exact (the sum of the reference files' step
    column; the reference's own statistics.txt undercounts it through an
    unsynchronised += under OpenMP, simulation.cpp:120)."""
    import gzip
    import shutil
    from mechafm_amd import abi
    src = os.path.join(GOLDEN, "example_graphene")
    job, out = tmp_path / "job", tmp_path / "out"
    job.mkdir()
    for fn in ("input.scan", "graphene.xyz", "parameters.dat"):
        shutil.copy(os.path.join(src, fn), job / fn)
    expected, name = {}, None
    with gzip.open(os.path.join(src, "outputs.txt.gz"), "rt") as f:
        for line in f.read().splitlines():
            if line.startswith("## "):
                name = line[3:]
                expected[name] = []
            else:
                expected[name].append(line)
    monkeypatch.chdir(job)                       # the input named as the reference's run named it
    sim = abi.Simulation("input.scan", gpu_ctx, outdir=str(out))
    st = sim.run(precision)
    sim.write_output(8)
    assert len(expected) == 41
    for fn, lines in expected.items():
        got = sorted(open(out / fn).read().splitlines())
        assert len(got) == len(lines), fn
        bad = sum(a != b for a, b in zip(got, lines))
        assert bad == 0, f"{fn}: {bad} of {len(lines)} lines differ"
    assert st["total_steps"] == sum(int(l.split()[-1]) for ls in expected.values() for l in ls)
    stats = open(out / "statistics.txt").read()
    assert f"Needed {st['total_steps']} minimization steps" in stats
    # console: the charge warning and the settings echo as the reference
    # prints them (its credits banner between the two is not reproduced)
    ref = open(os.path.join(src, "console.txt")).read()
    warn, echo = ref.split("+ - - -")[0], "+- \n+- Input settings" + ref.split("+- \n+- Input settings")[1]
    assert sim.settings() == warn + echo


@pytest.mark.parametrize("case", CASES)
def test_fast64_matches_reference(gpu_ctx, case, tmp_path):
    sim, tab, stats, shape = _run(gpu_ctx, case, tmp_path, "fast64")
    _, ref = golden_records(case)
    dpos, dfz = _errors(tab, ref)
    print(f"{case} fast64: max|dr| {dpos.max():.2e} A, max rel dFz {dfz.max():.2e}, "
          f"step mismatches {(tab[:, 3] != ref[:, 3]).sum()}/{len(ref)}")
    assert dpos.max() <= 1e-4
    assert dfz.max() <= 1e-3


@pytest.mark.parametrize("case", CASES)
def test_fast32_matches_reference(gpu_ctx, case, tmp_path):
    from mechafm_amd import abi
    sim, tab, stats, shape = _run(gpu_ctx, case, tmp_path, "fast32", options=abi.SCAN_REPRODUCIBLE)
    _, ref = golden_records(case)
    dpos, dfz = _errors(tab, ref)
    ok = (dpos <= 1e-4) & (dfz <= 1e-3)
    print(f"{case} fast32: within bar {ok.mean() * 100:.2f}%, median|dr| {np.median(dpos):.1e}, "
          f"step mismatches {(tab[:, 3] != ref[:, 3]).sum()}/{len(ref)}")
    assert ok.mean() >= 0.99


@pytest.mark.parametrize("compensation", [1, 2])
@pytest.mark.parametrize("case", ["c5_mini", "c2_direct", "c4_mini"])
def test_fast32_accumulation_modes(gpu_ctx, case, compensation, tmp_path):
    """Kahan per term (1) and plain FP32 sums (2) meet the fast32 bar too."""
    from mechafm_amd import abi
    sim, tab, stats, shape = _run(gpu_ctx, case, tmp_path, "fast32", compensation=compensation,
                                  options=abi.SCAN_REPRODUCIBLE)
    _, ref = golden_records(case)
    dpos, dfz = _errors(tab, ref)
    ok = (dpos <= 1e-4) & (dfz <= 1e-3)
    print(f"{case} fast32 compensation {compensation}: within bar {ok.mean() * 100:.2f}%")
    assert ok.mean() >= 0.99


@pytest.mark.parametrize("G", [1, 2, 4, 8, 16, 32, 64])
def test_lanes_per_pixel_invariance(gpu_ctx, G, tmp_path):
    """Splitting the sites over G lanes changes only the summation order."""
    sim, tab, stats, shape = _run(gpu_ctx, "c1_mini", tmp_path, "fast64", lanes_per_pixel=G)
    _, ref = golden_records("c1_mini")
    dpos, dfz = _errors(tab, ref)
    assert dpos.max() <= 1e-4 and dfz.max() <= 1e-3
    assert stats["lanes_per_pixel"] == G


@pytest.mark.parametrize("precision", ["plain64", "fast64", "fast32"])
def test_scan_past_force_grid_is_an_error(gpu_ctx, tmp_path, precision):
    """A lookup that needs a grid point past a non-periodic force grid stops
    the reference with an error (force_grid.cpp:236-240): the scan returns
    MAFM_ERR_OUT_OF_GRID (every gather stays clamped inside the grid), and
    the context runs the original scan again afterwards."""
    from mechafm_amd import abi
    sim, tab, stats, shape = _run(gpu_ctx, "c2_mini", tmp_path, precision,
                                  options=abi.SCAN_REPRODUCIBLE)
    d = sim.scan_desc(precision, options=abi.SCAN_REPRODUCIBLE)
    d.nx += 30                      # x up to 4 A: past the 1 A area and the grid's 1.5 A margin
    with pytest.raises(abi.MafmError) as err:
        gpu_ctx.scan(d)
    assert err.value.status == 3, err.value
    assert "past the grid" in str(err.value)
    again = sim.run(precision, options=abi.SCAN_REPRODUCIBLE)
    assert again["total_steps"] == stats["total_steps"]
    np.testing.assert_array_equal(abi.records_to_table(sim.records(), shape), tab)


def test_context_closed_before_simulation(tmp_path):
    """mafm_destroy of a context that a simulation still uses is completed by
    that simulation's mafm_sim_close: the simulation keeps working in between
    and frees the force grid it built on a live context."""
    from mechafm_amd import abi
    ctx = abi.Context(0)
    sim = abi.Simulation(materialize("c2_mini", str(tmp_path)), ctx)
    ctx.close()
    st = sim.run("fast64")
    assert st["total_steps"] > 0
    sim.close()


@pytest.mark.parametrize("case", [c for c in CASES if golden_eval(c) is not None])
def test_eval_tip_surface_matches_reference(gpu_ctx, case, tmp_path):
    from mechafm_amd import abi
    inp = materialize(case, str(tmp_path))
    sim = abi.Simulation(inp, gpu_ctx)
    gpu_ctx.set_systems([sim.system()])
    pos, fe = golden_eval(case)
    got = gpu_ctx.eval_tip_surface(0, pos)
    assert np.abs(got - fe).max() <= 1e-9 * max(1.0, np.abs(fe).max())


def test_batch_equals_single_runs(gpu_ctx, tmp_path):
    """Three molecules in one launch give the records of three single runs
    (fixed summation order and lanes, so the comparison is bit for bit)."""
    from mechafm_amd import abi, workloads as W
    wls = W.make_c3(str(tmp_path), mini=True, count=3)
    sims = [abi.Simulation(w.input_path, gpu_ctx) for w in wls]
    kw = dict(lanes_per_pixel=2, options=abi.SCAN_REPRODUCIBLE)
    abi.batch_run(sims, "fast64", **kw)
    batched = [s.records().copy() for s in sims]
    for s, b in zip(sims, batched):
        s.run("fast64", **kw)
        single = s.records()
        assert (single == b).all()


@pytest.mark.parametrize("precision", ["plain64", "fast64", "fast32"])
def test_system_without_sites(gpu_ctx, tmp_path, precision):
    """A system with no tip-surface sites, batched ahead of a real one: its
    tip only feels the tip-dummy terms and starts at their LJ minimum
    straight under the dummy, so every record has zero tip-surface force and
    energy, no lateral offset and rz = -d; the real system's records equal
    its single run bit for bit."""
    from mechafm_amd import abi
    sim = abi.Simulation(materialize("c1_mini", str(tmp_path)), gpu_ctx)
    kw = dict(lanes_per_pixel=2, options=abi.SCAN_REPRODUCIBLE)
    sim.run(precision, **kw)
    single = sim.records().copy()
    full = sim.system()
    empty = abi.SystemDesc.from_buffer_copy(full)
    empty.n_sites = 0
    gpu_ctx.set_systems([empty, full])
    st = gpu_ctx.scan(sim.scan_desc(precision, **kw))
    n = len(single)
    rec0, rec1 = gpu_ctx.records(0, n), gpu_ctx.records(n, n)
    for f in ("fx", "fy", "fz", "e", "rx", "ry"):
        assert (rec0[f] == 0).all(), f
    assert np.abs(rec0["rz"] + full.tip_dummy_d).max() < 1e-9
    assert (rec0["n_steps"] >= 1).all() and st["unconverged"] == 0
    assert (rec1 == single).all()


@pytest.mark.parametrize("precision", ["fast32", "fast64"])
def test_reproducible_runs_are_identical(gpu_ctx, tmp_path, precision):
    from mechafm_amd import abi, workloads as W
    wls = W.make_c3(str(tmp_path), mini=True, count=4)
    sims = [abi.Simulation(w.input_path, gpu_ctx) for w in wls]
    runs = []
    for _ in range(2):
        st = abi.batch_run(sims, precision, lanes_per_pixel=2, options=abi.SCAN_REPRODUCIBLE)
        assert st["wide_trips"] == 0
        runs.append(np.concatenate([s.records().copy() for s in sims]))
    assert (runs[0] == runs[1]).all()


@pytest.mark.parametrize("G", [1, 2, 8])
def test_wide_tail_matches_fixed_order(gpu_ctx, tmp_path, G):
    """Regrouping a wave's last pixel groups over 64 lanes changes only the
    summation order: the fast32 bar against the fixed-order run."""
    from mechafm_amd import abi, workloads as W
    wls = W.make_c3(str(tmp_path), mini=True, count=4)
    sims = [abi.Simulation(w.input_path, gpu_ctx) for w in wls]
    abi.batch_run(sims, "fast32", lanes_per_pixel=G, options=abi.SCAN_REPRODUCIBLE)
    fixed = np.concatenate([s.records().copy() for s in sims])
    st = abi.batch_run(sims, "fast32", lanes_per_pixel=G)
    wide = np.concatenate([s.records().copy() for s in sims])
    assert st["wide_trips"] > 0
    dpos = np.abs(np.stack([wide[c] - fixed[c] for c in ("rx", "ry", "rz")])).max(0)
    dfz = np.abs(wide["fz"] - fixed["fz"]) / np.maximum(np.abs(fixed["fz"]), 1e-2)
    ok = (dpos <= 1e-4) & (dfz <= 1e-3)
    print(f"G={G}: wide trips {st['wide_trips']}, within bar {ok.mean() * 100:.2f}%")
    assert ok.mean() >= 0.99


@pytest.mark.parametrize("W", [1, 2, 4, 6])
def test_waves_per_group_share_sites(gpu_ctx, tmp_path, W):
    """W waves of a workgroup share one LDS copy of a periodic system's sites
    (c4 mini: slab with 9 images plus an external potential)."""
    from mechafm_amd import abi
    sim, tab, stats, shape = _run(gpu_ctx, "c4_mini", tmp_path, "fast32", lanes_per_pixel=4, waves_per_group=W,
                                  options=abi.SCAN_REPRODUCIBLE)
    _, ref = golden_records("c4_mini")
    dpos, dfz = _errors(tab, ref)
    ok = (dpos <= 1e-4) & (dfz <= 1e-3)
    assert stats["block_threads"] == 64 * W
    assert ok.mean() >= 0.99


@pytest.mark.parametrize("gz", [False, True])
def test_output_files_match_reference_format(gpu_ctx, tmp_path, gz):
    """Per-height files: reference printf formats (simulation.cpp:180-195),
    one file per z named scan-%06.3f.dat[.gz] (mechafm.cpp:59-75)."""
    import gzip
    from mechafm_amd import abi
    inp = materialize("c1_minterm_e", str(tmp_path / "in"))
    text = open(inp).read()
    open(inp, "w").write(text.replace("gzip                     off", "gzip                     on") if gz else text)
    out = str(tmp_path / "out")
    sim = abi.Simulation(inp, gpu_ctx, outdir=out)
    sim.run("plain64")
    sim.write_output(4)
    _, ref = golden_records("c1_minterm_e")
    nz = int(ref[:, 2].max()) + 1
    files = sorted(f for f in os.listdir(out) if f.startswith("scan-"))
    suffix = ".dat.gz" if gz else ".dat"
    assert len(files) == nz and files[-1] == "scan-10.000" + suffix
    fmt = "%d %d %d %6.3f %6.3f %6.3f %8.4f %8.4f %8.4f %6.3f %6.3f %6.3f %6.3f %8.4f %8.4f %d"
    for k in (0, nz // 2, nz - 1):
        z = 10.0 - k * 0.1
        path = os.path.join(out, "scan-%06.3f" % z + suffix)
        got = (gzip.open(path, "rt") if gz else open(path)).read().splitlines()
        rows = ref[ref[:, 2] == k]
        want = [fmt % (int(r[2]), int(r[0]), int(r[1]), r[4], r[5], r[6], r[7], r[8], r[9], r[10], r[11], r[12],
                       r[13], r[14], r[15], int(r[3])) for r in rows]
        assert sorted(got) == sorted(want)


def _large_molecule(tmp_path, n_atoms, area, cell):
    from mechafm_amd import workloads as W
    rng = np.random.default_rng(7)
    mol = W.random_slab(n_atoms, (cell, cell), 4.0, rng)
    opts = W._base_options(coulomb="on", area=area, zhigh="10.0", zlow="9.0")
    return W._emit("large", str(tmp_path), mol, opts, True)


@pytest.mark.parametrize("precision,n_atoms,area", [("fast64", 1500, "0.3 0.3"), ("fast32", 3000, "2.0 2.0")])
def test_large_molecule_staged_in_lds(gpu_ctx, tmp_path, precision, n_atoms, area):
    """A molecule whose sites take more than 64 KiB (96 000 B here) is still
    staged whole in LDS, up to what one workgroup may take (159 KiB on
    gfx950), and scanned with several lanes per pixel."""
    from mechafm_amd import abi
    sim = abi.Simulation(_large_molecule(tmp_path, n_atoms, area, 40.0).input_path, gpu_ctx)
    sim.run("plain64")
    ref = sim.records().copy()
    st = sim.run(precision)
    got = sim.records()
    assert st["lanes_per_pixel"] > 1
    dpos = np.abs(np.stack([got[c] - ref[c] for c in ("rx", "ry", "rz")])).max(0)
    dfz = np.abs(got["fz"] - ref["fz"]) / np.maximum(np.abs(ref["fz"]), 1e-2)
    ok = (dpos <= 1e-4) & (dfz <= 1e-3)
    assert ok.mean() >= (1.0 if precision == "fast64" else 0.99)


@pytest.mark.parametrize("precision,n_atoms,area,cell", [("fast64", 2600, "0.3 0.3", 40.0),
                                                         ("fast32", 5200, "2.0 2.0", 60.0)])
def test_large_molecule_streams_sites(gpu_ctx, tmp_path, precision, n_atoms, area, cell):
    """A molecule whose sites exceed the LDS one workgroup may take (159 KiB:
    2544 FP64 / 5088 FP32 sites) is streamed with scalar loads, one lane per
    pixel.  (The FP32 case scans 441 columns so that one column whose FIRE
    steps FP32 rounding shifts stays well inside the 1 % the bar allows.)"""
    from mechafm_amd import abi, workloads as W
    rng = np.random.default_rng(7)
    mol = W.random_slab(n_atoms, (cell, cell), 4.0, rng)
    opts = W._base_options(coulomb="on", area=area, zhigh="10.0", zlow="9.0")
    wl = W._emit("large", str(tmp_path), mol, opts, True)
    sim = abi.Simulation(wl.input_path, gpu_ctx)
    sim.run("plain64")
    ref = sim.records().copy()
    st = sim.run(precision)
    got = sim.records()
    assert st["lanes_per_pixel"] == 1 and st["block_threads"] == 64
    dpos = np.abs(np.stack([got[c] - ref[c] for c in ("rx", "ry", "rz")])).max(0)
    dfz = np.abs(got["fz"] - ref["fz"]) / np.maximum(np.abs(ref["fz"]), 1e-2)
    ok = (dpos <= 1e-4) & (dfz <= 1e-3)
    assert ok.mean() >= (1.0 if precision == "fast64" else 0.99)
    with pytest.raises(abi.MafmError):
        sim.run(precision, lanes_per_pixel=2)


def test_autotune_on_device(gpu_ctx, tmp_path):
    """The package tuner measures launch shapes on the workload and caches
    the fastest; the cached shape is reused without launches."""
    from mechafm_amd import abi, tune, workloads as W
    wls = W.make_c3(str(tmp_path / "in"), mini=True, count=4)
    sims = [abi.Simulation(w.input_path, gpu_ctx) for w in wls]
    cache = str(tmp_path / "tune.json")
    best = tune.tune(sims, "fast32", cache=cache, device=gpu_ctx.device_info())
    assert best["lanes_per_pixel"] in tune.DEFAULT_LANES and best["waves_per_simd"] in (1, 2, 3)
    assert not best["cached"] and len(best["tried"]) == len(tune.DEFAULT_LANES) * len(tune.WAVES) + 1
    again = tune.tune(sims, "fast32", cache=cache, device=gpu_ctx.device_info(),
                      run=lambda *a, **k: pytest.fail("cached shape should not launch"))
    assert again["cached"] and again["lanes_per_pixel"] == best["lanes_per_pixel"]
    st = abi.batch_run(sims, "fast32", lanes_per_pixel=best["lanes_per_pixel"],
                       waves_per_simd=best["waves_per_simd"])
    assert st["unconverged"] == 0


def test_records_of_a_replaced_scan_are_refused(gpu_ctx, tmp_path):
    """Records live on the context: once another scan has run there, a
    simulation's records, files and frequency-shift images are refused with
    MAFM_ERR_STATE instead of returning the other job's data; running the
    simulation again restores them."""
    from mechafm_amd import abi
    a = abi.Simulation(materialize("c1_mini", str(tmp_path / "a")), gpu_ctx, outdir=str(tmp_path / "oa"))
    b = abi.Simulation(materialize("c5_mini", str(tmp_path / "b")), gpu_ctx)
    with pytest.raises(abi.MafmError) as err:
        a.records()
    assert err.value.status == 4
    a.run("fast64", options=abi.SCAN_REPRODUCIBLE)
    mine = a.records().copy()
    b.run("fast64")
    for call in (a.records, lambda: a.write_output(2), a.write_deltaf):
        with pytest.raises(abi.MafmError) as err:
            call()
        assert err.value.status == 4 and "another scan" in str(err.value)
    b.records()                                   # the latest scan's own records stay readable
    a.run("fast64", options=abi.SCAN_REPRODUCIBLE)
    assert (a.records() == mine).all()


@pytest.mark.parametrize("precision", ["fast64", "fast32"])
def test_decision_margins(gpu_ctx, tmp_path, precision):
    """MAFM_SCAN_MARGINS records, per position, how near the height's
    minimiser decisions came to flipping; recording them leaves the records
    unchanged.  c1_mini runs FIRE with minterm ef and etol = ftol = 0.01:
    the force test decides (its margin is below ftol everywhere), the
    energy difference stays far inside etol."""
    from mechafm_amd import abi
    sim, tab, _, shape = _run(gpu_ctx, "c1_mini", tmp_path, precision, options=abi.SCAN_REPRODUCIBLE)
    sim.run(precision, options=abi.SCAN_REPRODUCIBLE | abi.SCAN_MARGINS)
    again = abi.records_to_table(sim.records(), shape)
    np.testing.assert_array_equal(again, tab)
    n = len(tab)
    m = gpu_ctx.margins(0, n)
    assert (m[:, 3] == 0).all()
    assert np.isfinite(m[:, 0]).all() and (m[:, 0] >= 0).all() and (m[:, 0] < 0.01).all()
    assert (m[:, 2] >= 0).all()
    e = m[:, 1][np.isfinite(m[:, 1])]
    assert len(e) == n and (e > 0).all() and (e <= 0.01).all()
    plain = abi.Simulation(materialize("c1_mini", str(tmp_path / "p")), gpu_ctx)
    plain.run("plain64")
    with pytest.raises(abi.MafmError):
        gpu_ctx.margins(0, n)
