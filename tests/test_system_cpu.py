"""The System-level operators (FIREMinimisation / SDMinimisation from a
given state, eulerStep / midpointStep / rk4Step, Interaction::eval over the
list) restated by the oracle, pinned to the reference on the CPU.

* Replaying pixel column 0 of a flexible golden height by height through
  the oracle's minimise operator -- the System copy of simulation.cpp:93
  with the dummy and tip placed, velocities and the tip energy carried from
  height to height, the tip lowered by dz in between -- gives the
  reference's iteration count and relaxed tip offset at every height, bit
  for bit: the operator is the reference's minimiser with its integrator.
* The eval operator from zero equals the whole-list evaluation that is
  pinned to the reference's forces goldens, and adding onto given forces
  keeps the list's order (a planted reordering of the list changes it).
* The host library exports the three entry points (tests/test_abi.py) and
  refuses bad lists before touching a device; the C++ API of
  include/mafm_system.hpp (System, FIREMinimisation, midpointStep, ...)
  compiles and links into tests/bin/system_api_check.
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, flex_cases, flex_golden, materialize
from oracle import oracle_flex as OF

REPLAY = [c for c in flex_cases() if c not in ("fx_pbc", "fx_graphene")]


def _column_start(fs, x=0.0, y=0.0):
    s = fs.base.system
    pos = np.array(s.pos, dtype=np.float64, copy=True)
    pos[0] = (x, y, fs.base.opts.zhigh)
    pos[1] = (x, y, fs.base.opts.zhigh - s.tip_dummy_d)
    return pos


@pytest.mark.parametrize("case", REPLAY)
def test_minimise_operator_replays_a_reference_column(tmp_path, case):
    shape, ref, _ = flex_golden(case)
    fs = OF.flex_setup(materialize(case, str(tmp_path)))
    so = OF._fx_opts(fs, shape)
    nz = shape[2]
    n = len(fs.base.system.pos)
    pos = _column_start(fs)
    vel, frc, ene = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros(n)
    for k in range(nz):
        steps = OF.system_op(fs.terms, fs.fixed, "minimise", pos, vel, frc, ene, integrator=so.integrator,
                             dt=so.dt, minimiser=so.minimiser, minterm=so.minterm, maxsteps=so.maxsteps,
                             etol=so.etol, ftol=so.ftol)
        row = ref[k]                                     # column 0 is the first nz records
        assert (int(row[0]), int(row[1]), int(row[2])) == (0, 0, k)
        assert steps == int(row[3]), f"height {k}: {steps} steps, the reference {int(row[3])}"
        np.testing.assert_array_equal(pos[1] - pos[0], row[10:13])
        pos[0, 2] -= fs.base.opts.dz
        pos[1, 2] -= fs.base.opts.dz


def test_eval_operator_is_the_pinned_list_evaluation(tmp_path):
    fs = OF.flex_setup(materialize("fx_mini", str(tmp_path)))
    rng = np.random.default_rng(5)
    pos = _column_start(fs, 0.3, 0.2) + rng.normal(0, 0.02, (len(fs.fixed), 3))
    F, E = OF.evaluate(fs.terms, pos)
    f, e = np.zeros_like(F), np.zeros_like(E)
    OF.system_op(fs.terms, fs.fixed, "eval", pos.copy(), None, f, e)
    np.testing.assert_array_equal(f, F)
    np.testing.assert_array_equal(e, E)
    # onto given forces: every term adds in list order after what is there
    f0 = rng.normal(0, 1, F.shape)
    e0 = rng.normal(0, 1, E.shape)
    f1, e1 = f0.copy(), e0.copy()
    OF.system_op(fs.terms, fs.fixed, "eval", pos.copy(), None, f1, e1)
    assert np.abs(f1 - (f0 + F)).max() < 1e-9 and np.abs(e1 - (e0 + E)).max() < 1e-9
    # a planted mistake: the list reversed adds the same terms in another order
    f2, e2 = f0.copy(), e0.copy()
    OF.system_op(fs.terms[::-1].copy(), fs.fixed, "eval", pos.copy(), None, f2, e2)
    assert not (np.array_equal(f2, f1) and np.array_equal(e2, e1))


def test_bad_lists_are_refused_without_a_device():
    import ctypes

    from mechafm_amd import abi
    L = abi.lib()
    terms = np.zeros(1, dtype=abi.FLEX_TERM_DTYPE)
    terms[0]["kind"] = abi.FX_BOND
    terms[0]["atom"][0], terms[0]["atom"][1] = 0, 5        # atom 5 of a 3-atom system
    fixed = np.array([1, 0, 0], dtype=np.int32)
    lst = abi.InteractionList(3, 1, terms.ctypes.data_as(ctypes.c_void_p),
                              fixed.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    pos = np.zeros((1, 3, 3))
    out = np.zeros((1, 3, 3))
    ene = np.zeros((1, 3))
    dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))   # noqa: E731
    assert L.mafm_system_eval(None, ctypes.byref(lst), 1, dp(pos), dp(out), dp(ene)) == 1   # no context


def test_cpp_system_api_builds_and_links():
    subprocess.check_call(["make", "-B", "-C", os.path.join(ROOT, "tests")], stdout=subprocess.DEVNULL)
    exe = os.path.join(ROOT, "tests", "bin", "system_api_check")
    syms = subprocess.check_output(["nm", "-C", exe], text=True)
    for name in ("mafm_system_minimise", "mafm_system_step", "mafm_system_eval", "mafm_sim_flex_system"):
        assert f"U {name}" in syms, name
    import torch
    if not torch.cuda.is_available():
        r = subprocess.run([exe, os.path.join(ROOT, "tests", "golden", "fx_mini", "input.scan")],
                           capture_output=True, text=True)
        assert r.returncode == 3 and "no device" in r.stderr
