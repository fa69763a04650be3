"""Flexible on, on the CPU: the oracle and the host's interaction list,
pinned to the reference.

* The oracle's scan (oracle_flex.c over oracle_flex.py's set-up) reproduces
  the reference's records and the centre column's relaxed atoms (what
  System::makeXYZFile writes) bit for bit on every flexible golden case:
  FIRE with euler / midpoint / rk4, steepest descent, minterm e / ef,
  maxsteps, periodic images, Morse overwrites, the reference's own example.
* The oracle's whole-list evaluation reproduces the reference's forces and
  energies on every atom at perturbed geometries, bit for bit -- and a
  planted mistake in the list (a dropped dihedral, swapped dihedral ends, a
  wrong equilibrium angle, reordered terms) fails that comparison, so the
  comparison pins the list's content and order.
* The host library's list (mafm_sim_flex_system, built by the C++ job layer
  without a device) is byte-identical to the oracle's.
"""
import os
import shutil

import numpy as np
import pytest

from conftest import GOLDEN, flex_cases, flex_golden, materialize
from mechafm_amd import abi
from oracle import oracle_flex as OF

CASES = flex_cases()
FORCE_CASES = [c for c in CASES if os.path.exists(os.path.join(GOLDEN, c, "forces.npz"))]


@pytest.fixture(scope="module")
def setups(tmp_path_factory):
    out = {}
    for case in CASES:
        d = tmp_path_factory.mktemp(case)
        out[case] = (OF.flex_setup(materialize(case, str(d))), str(d))
    return out


@pytest.mark.parametrize("case", CASES)
def test_oracle_scan_is_the_reference(setups, case):
    fs, _ = setups[case]
    shape, ref, ref_state = flex_golden(case)
    rec, state, _ = OF.scan(fs)
    assert rec.shape == ref.shape
    assert np.array_equal(rec, ref)
    assert np.array_equal(state, ref_state)


@pytest.mark.parametrize("case", FORCE_CASES)
def test_oracle_forces_are_the_reference(setups, case):
    fs, _ = setups[case]
    g = np.load(os.path.join(GOLDEN, case, "forces.npz"))
    for c in range(len(g["pos"])):
        F, E = OF.evaluate(fs.terms, g["pos"][c])
        assert np.array_equal(F, g["forces"][c])
        assert np.array_equal(E, g["energies"][c])


def _mismatch(terms, g) -> bool:
    for c in range(len(g["pos"])):
        F, E = OF.evaluate(terms, g["pos"][c])
        if not (np.array_equal(F, g["forces"][c]) and np.array_equal(E, g["energies"][c])):
            return True
    return False


def test_planted_mistakes_fail_the_force_check(setups):
    fs, _ = setups["fx_mini"]
    g = np.load(os.path.join(GOLDEN, "fx_mini", "forces.npz"))
    t = fs.terms
    dih = np.where(t["kind"] == OF.FX_DIHEDRAL)[0]
    ang = np.where(t["kind"] == OF.FX_ANGLE)[0]
    bonded = np.where(np.isin(t["kind"], [OF.FX_BOND, OF.FX_ANGLE, OF.FX_DIHEDRAL]))[0]
    assert not _mismatch(t, g)
    dropped = np.delete(t, dih[3])
    assert _mismatch(dropped, g)
    swapped = t.copy()
    swapped["atom"][dih[5]] = swapped["atom"][dih[5]][[3, 1, 2, 0]]
    assert _mismatch(swapped, g)
    wrong_angle = t.copy()
    wrong_angle["p"][ang[2], 1] += 1e-9
    assert _mismatch(wrong_angle, g)
    reordered = t.copy()
    reordered[bonded] = reordered[bonded[::-1]]
    assert _mismatch(reordered, g)
    # an angle's shared atom is updated after its two ends (interactions.cpp:327-329)
    shared_last = t.copy()
    shared_last["atom"][ang[0]][:3] = shared_last["atom"][ang[0]][[1, 0, 2]]
    assert _mismatch(shared_last, g)


@pytest.mark.parametrize("case", CASES)
def test_host_list_is_the_oracle_list(setups, case):
    fs, d = setups[case]
    sim = abi.Simulation(os.path.join(d, "input.scan"), None)
    fx = sim.flex_system()
    assert fx["terms"].tobytes() == fs.terms.tobytes()
    assert np.array_equal(fx["positions"], fs.base.system.pos)
    assert np.array_equal(fx["fixed"], fs.fixed)
    assert fx["tip_dummy_d"] == fs.base.system.tip_dummy_d
    shape = fs.base.scan_shape
    assert fx["state_column"] == shape[0] * shape[1] // 2
    assert fx["tip_pbc"] == int(fs.base.system.tip_pbc)
    info = sim.info()
    assert info["flexible"] == 1 and info["n_terms"] == len(fs.terms) and info["n_sites"] == fs.n_tip_surface
    with pytest.raises(abi.MafmError, match="term list"):
        sim.system()
    assert "flexible:                 on" in sim.settings()
    sim.close()


def test_rigid_jobs_have_no_term_list(tmp_path):
    sim = abi.Simulation(materialize("c1_mini", str(tmp_path)), None)
    with pytest.raises(abi.MafmError, match="site table"):
        sim.flex_system()
    assert sim.info()["flexible"] == 0


@pytest.mark.parametrize("line,message", [
    ("bond", "No harmonic bond parameters found in parameter file!"),
    ("angle", "No harmonic angle parameters found in parameter file!"),
    ("dihedral", "No harmonic dihedral parameters found in parameter file!"),
    ("substrate", "No harmonic substrate support parameters found in parameter file!"),
])
def test_missing_flexible_parameters_are_the_reference_errors(tmp_path, line, message):
    inp = materialize("fx_mini", str(tmp_path))
    p = os.path.join(str(tmp_path), "parameters.dat")
    kept = [l for l in open(p).read().splitlines() if not l.startswith(line + " ")]
    open(p, "w").write("\n".join(kept) + "\n")
    with pytest.raises(abi.MafmError, match=message):
        abi.Simulation(inp, None)


def test_duplicate_flexible_parameters_warn(tmp_path):
    inp = materialize("fx_mini", str(tmp_path))
    with open(os.path.join(str(tmp_path), "parameters.dat"), "a") as f:
        f.write("bond 30.0\ntopobond H C 1.1\n")
    sim = abi.Simulation(inp, None)
    text = sim.settings()
    assert "+- WARNING: Parameters for harmonic bond defined multiple times!" in text
    assert "+- WARNING: The topobond for H and C is defined at least twice!" in text
    # the last bond constant and the last matching topobond win (parse.cpp:723, simulation.cpp:566-570)
    t = sim.flex_system()["terms"]
    bonds = t[t["kind"] == abi.FX_BOND]
    assert np.all(bonds["p"][:, 0] == 30.0)


def test_fixed_column_is_ignored_when_rigid(tmp_path):
    inp = materialize("fx_mini", str(tmp_path))
    text = open(inp).read().replace("flexible                 on", "flexible                 off")
    open(inp, "w").write(text)
    sim = abi.Simulation(inp, None)
    assert sim.info()["flexible"] == 0
    assert sim.info()["n_sites"] == 40
