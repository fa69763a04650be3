"""The System-level operators on the GPU (mafm_system_eval / _step /
_minimise through the C ABI) against the oracle's restatement, which
tests/test_system_cpu.py pins to the reference.

* Many independent states at once (a workgroup each): perturbed copies of
  the flexible test system with random velocities and random starting
  forces / energies.  eval and every integrator step agree with the oracle
  to 1e-9 (the device's acos / sin / atan / exp may differ from glibc's in
  the last place; everything else is the reference's operation order);
  FIRE with each integrator and steepest descent give the oracle's
  iteration count and positions within 1e-8 on every state whose oracle
  trajectory does not change when its coordinates move by one ulp (far from
  equilibrium the flexible system is chaotic: a last-place difference, such
  as the device's libm against glibc, can send FIRE elsewhere).
* Pixel column 0 of a flexible golden replayed through the minimise
  operator, height by height, gives the reference's iteration counts and
  tip offsets.
* The C++ API (include/mafm_system.hpp, driven by tests/bin/system_api_check
  as Simulation::run drives the reference's) gives the same column.
* Bad lists are refused with a message; a state array of the wrong shape
  never reaches the device.
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, flex_golden, materialize
from mechafm_amd import abi
from oracle import oracle_flex as OF

pytestmark = pytest.mark.gpu

TOL = 1e-9


@pytest.fixture(scope="module")
def system(tmp_path_factory):
    fs = OF.flex_setup(materialize("fx_mini", str(tmp_path_factory.mktemp("fx_mini"))))
    return fs


def _states(fs, ns, seed, spread=0.03, speed=0.05):
    rng = np.random.default_rng(seed)
    s = fs.base.system
    n = len(fs.fixed)
    pos = np.repeat(np.asarray(s.pos, dtype=np.float64)[None], ns, axis=0)
    xy = rng.uniform(0, 1.5, (ns, 2))
    pos[:, 0, :2] = xy
    pos[:, 0, 2] = fs.base.opts.zhigh
    pos[:, 1, :2] = xy
    pos[:, 1, 2] = fs.base.opts.zhigh - s.tip_dummy_d
    free = fs.fixed != 1
    pos[:, free] += rng.normal(0, spread, (ns, int(free.sum()), 3))
    vel = np.where(free[None, :, None], rng.normal(0, speed, (ns, n, 3)), 0.0)
    return pos, vel, rng.normal(0, 0.5, (ns, n, 3)), rng.normal(0, 0.5, (ns, n))


def test_eval_matches_the_oracle(gpu_ctx, system):
    fs = system
    pos, _, f0, e0 = _states(fs, 40, 1)
    frc, ene = gpu_ctx.system_eval(fs.terms, fs.fixed, pos, f0, e0)
    for s in range(len(pos)):
        f, e = f0[s].copy(), e0[s].copy()
        OF.system_op(fs.terms, fs.fixed, "eval", pos[s].copy(), None, f, e)
        np.testing.assert_allclose(frc[s], f, rtol=0, atol=TOL)
        np.testing.assert_allclose(ene[s], e, rtol=0, atol=TOL)
    # from zero: the plain forces of the list
    frc0, ene0 = gpu_ctx.system_eval(fs.terms, fs.fixed, pos[:3])
    F, E = OF.evaluate(fs.terms, pos[2])
    np.testing.assert_allclose(frc0[2], F, rtol=0, atol=TOL)
    np.testing.assert_allclose(ene0[2], E, rtol=0, atol=TOL)


@pytest.mark.parametrize("integrator", [0, 1, 2])
def test_step_matches_the_oracle(gpu_ctx, system, integrator):
    fs = system
    pos, vel, f0, e0 = _states(fs, 24, 10 + integrator)
    dt = 0.02
    p, v, f, e = gpu_ctx.system_step(fs.terms, fs.fixed, integrator, dt, pos, vel, f0, e0)
    for s in range(len(pos)):
        ps, vs, fs_, es = pos[s].copy(), vel[s].copy(), f0[s].copy(), e0[s].copy()
        OF.system_op(fs.terms, fs.fixed, "step", ps, vs, fs_, es, integrator=integrator, dt=dt)
        for got, want in ((p[s], ps), (v[s], vs), (f[s], fs_), (e[s], es)):
            np.testing.assert_allclose(got, want, rtol=0, atol=TOL)


@pytest.mark.parametrize("minimiser,integrator", [(1, 1), (1, 0), (1, 2), (0, 1)])
def test_minimise_matches_the_oracle(gpu_ctx, system, minimiser, integrator):
    fs = system
    pos, vel, _, e0 = _states(fs, 24, 20 + 3 * minimiser + integrator, spread=0.01, speed=0.01)
    kw = dict(minimiser=minimiser, integrator=integrator, minterm=1, maxsteps=3000, etol=0.01, ftol=0.003,
              dt=0.002)
    p, v, f, e, steps = gpu_ctx.system_minimise(fs.terms, fs.fixed, pos, vel, e0, **kw)

    def oracle(start):
        ps, vs, fs_, es = start.copy(), vel[s].copy(), np.zeros_like(start), e0[s].copy()
        n = OF.system_op(fs.terms, fs.fixed, "minimise", ps, vs, fs_, es, **kw)
        return n, ps, vs

    checked = 0
    for s in range(len(pos)):
        n, ps, vs = oracle(pos[s])
        # a state whose own trajectory changes when its coordinates move by
        # one ulp is chaotic at the last-place level, where the device's
        # acos / sin / atan may differ from glibc's (and one that never
        # converges may run away): no exact count there
        if n >= kw["maxsteps"] or oracle(np.nextafter(pos[s], np.inf))[0] != n:
            continue
        checked += 1
        assert steps[s] == n, f"state {s}: {steps[s]} iterations, the oracle {n}"
        np.testing.assert_allclose(p[s], ps, rtol=0, atol=1e-8)
        np.testing.assert_allclose(v[s], vs, rtol=0, atol=1e-8)
    print(f"minimiser {minimiser} integrator {integrator}: iterations {steps.min()}..{steps.max()}, "
          f"{checked} of {len(pos)} states converged, insensitive to one ulp and checked")
    assert checked >= 0.6 * len(pos)


def test_a_reference_column_through_the_operators(gpu_ctx, tmp_path):
    shape, ref, _ = flex_golden("fx_mini")
    fs = OF.flex_setup(materialize("fx_mini", str(tmp_path)))
    so = OF._fx_opts(fs, shape)
    s = fs.base.system
    pos = np.asarray(s.pos, dtype=np.float64).copy()[None]
    pos[0, 0] = (0, 0, fs.base.opts.zhigh)
    pos[0, 1] = (0, 0, fs.base.opts.zhigh - s.tip_dummy_d)
    vel, ene = np.zeros_like(pos), np.zeros(pos.shape[:2])
    for k in range(shape[2]):
        pos, vel, _, ene, steps = gpu_ctx.system_minimise(
            fs.terms, fs.fixed, pos, vel, ene, minimiser=so.minimiser, integrator=so.integrator,
            minterm=so.minterm, maxsteps=so.maxsteps, etol=so.etol, ftol=so.ftol, dt=so.dt)
        assert int(steps[0]) == int(ref[k, 3])
        np.testing.assert_allclose(pos[0, 1] - pos[0, 0], ref[k, 10:13], rtol=0, atol=1e-9)
        pos[0, 0, 2] -= fs.base.opts.dz
        pos[0, 1, 2] -= fs.base.opts.dz


@pytest.mark.parametrize("case", ["fx_mini", "fx_sd", "fx_rk4"])
def test_cpp_api_replays_a_reference_column(tmp_path, case):
    shape, ref, _ = flex_golden(case)
    exe = os.path.join(ROOT, "tests", "bin", "system_api_check")
    out = subprocess.run([exe, materialize(case, str(tmp_path))], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    rows = np.array([[float(v) for v in line.split()] for line in out.stdout.splitlines()])
    assert rows.shape == (shape[2], 5)
    np.testing.assert_array_equal(rows[:, 1], ref[:shape[2], 3])
    np.testing.assert_allclose(rows[:, 2:5], ref[:shape[2], 10:13], rtol=0, atol=1e-9)


def test_bad_lists_are_refused(gpu_ctx, system):
    fs = system
    pos = _states(fs, 2, 3)[0]
    bad = fs.terms.copy()
    bad[5]["atom"][1] = len(fs.fixed) + 3
    with pytest.raises(abi.MafmError, match="outside"):
        gpu_ctx.system_eval(bad, fs.fixed, pos)
    bad = fs.terms.copy()
    bad[0]["kind"] = 42
    with pytest.raises(abi.MafmError, match="unknown kind"):
        gpu_ctx.system_eval(bad, fs.fixed, pos)
    with pytest.raises(ValueError):
        gpu_ctx.system_step(fs.terms, fs.fixed, 1, 0.01, pos, np.zeros((2, 3, 3)))
    with pytest.raises(abi.MafmError, match="integrator"):
        gpu_ctx.system_step(fs.terms, fs.fixed, 7, 0.01, pos, np.zeros_like(pos))
