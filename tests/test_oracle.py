"""CPU tests of the oracle (oracle/oracle.py + oracle/oracle_scan.c).

Pins, each chosen so that a plausible slip fails at least one:
* reference golden fixtures (tests/golden/*, written by the reference's own
  objects): every record, every FIRE step count;
* closed forms of the pair operators (LJ minimum at 2^(1/6) sigma with depth
  epsilon, Morse zero at r_e and plateau D_e, Coulomb 1/r) and F = -grad E by
  central differences (catches sign / factor slips in forces);
* trilinear lookup: exact on a trilinear field;
* the FFT electrostatics: a plane-wave potential convolved with the tip's
  normalised Gaussian is q exp(-2 pi^2 sigma^2 |k|^2) cos(2 pi k.r), with
  force = -grad of it (catches the volume factor, the k aliasing, the sign).
"""
import math
import os

import numpy as np
import pytest

from conftest import golden_cases, golden_eval, golden_records, materialize
from oracle import oracle as O


@pytest.mark.parametrize("case", golden_cases())
def test_oracle_matches_reference_golden(case, tmp_path):
    inp = materialize(case, str(tmp_path))
    st = O.setup(inp)
    shape, ref = golden_records(case)
    assert st.scan_shape == shape
    rec, stats = O.scan(st)
    assert stats["rc"] == 0
    np.testing.assert_array_equal(rec[:, 3], ref[:, 3])          # step counts: exact
    tol = 1e-12 if case.startswith("c4") else 0.0                # c4: numpy FFT vs kiss_fft rounding
    assert np.abs(rec - ref).max() <= tol * max(1.0, np.abs(ref).max())
    ev = golden_eval(case)
    if ev is not None:
        pos, fe = ev
        got = O.eval_tip_surface(st, pos)
        assert np.abs(got - fe).max() <= 1e-12 * max(1.0, np.abs(fe).max())


def _one_atom_setup(tmp_path, atom="C", overwrite=None, charge=None, coulomb=False):
    d = str(tmp_path)
    with open(os.path.join(d, "m.xyz"), "w") as f:
        if charge is None:
            f.write(f"1\n\n{atom} 0.0 0.0 0.0\n")
        else:
            f.write(f"1\n\n{atom} 0.0 0.0 0.0 {charge}\n")
    with open(os.path.join(d, "p.dat"), "w") as f:
        f.write("atom C 0.07 3.55 12.011 0.0\natom X 20.0 3.55 12.011 0.021\natom O 0.192 3.15 15.999 -0.021\n")
        f.write("harm 0.72 0.0\n")
        if overwrite:
            f.write(f"pair_ovwrt {overwrite}\n")
    with open(os.path.join(d, "in.scan"), "w") as f:
        f.write("xyzfile m.xyz\nparamfile p.dat\ntipatom O\ndummyatom X\nunits kcal/mol\nminterm f\n")
        f.write("area 1.0 1.0\ncenter 0.0 0.0\n")
        f.write(f"coulomb {'on' if coulomb else 'off'}\n")
    return O.setup(os.path.join(d, "in.scan"))


def _radial(st, rs):
    pos = np.array([[0.0, 0.0, r] for r in rs])
    return O.eval_tip_surface(st, pos)


def test_lj_closed_form(tmp_path):
    st = _one_atom_setup(tmp_path)
    eps = math.sqrt(0.07 * 0.192)
    sig = (3.55 + 3.15) / 2
    rmin = 2 ** (1 / 6) * sig
    fe = _radial(st, [rmin])[0]
    assert fe[3] == pytest.approx(-eps, rel=1e-12)
    assert abs(fe[2]) < 1e-12
    rs = np.linspace(3.0, 8.0, 11)
    fe = _radial(st, rs)
    e = 4 * eps * ((sig / rs) ** 12 - (sig / rs) ** 6)
    np.testing.assert_allclose(fe[:, 3], e, rtol=1e-12)


def test_morse_closed_form(tmp_path):
    st = _one_atom_setup(tmp_path, overwrite="O C morse 0.5 1.7 3.4")
    fe = _radial(st, [3.4, 60.0])
    assert abs(fe[0, 3]) < 1e-14 and abs(fe[0, 2]) < 1e-14
    assert fe[1, 3] == pytest.approx(0.5, rel=1e-12)


def test_coulomb_closed_form(tmp_path):
    st = _one_atom_setup(tmp_path, charge=0.3, coulomb=True, overwrite="O C lj 0.0 1.0")
    rs = np.array([3.0, 5.0, 9.0])
    fe = _radial(st, rs)
    qq = 332.06371 * (-0.021) * 0.3
    np.testing.assert_allclose(fe[:, 3], qq / rs, rtol=1e-13)
    np.testing.assert_allclose(fe[:, 2], qq / rs ** 2, rtol=1e-13)


@pytest.mark.parametrize("ow,coul", [(None, False), ("O C morse 0.5 1.7 3.4", False), (None, True)])
def test_force_is_minus_gradient(tmp_path, ow, coul):
    st = _one_atom_setup(tmp_path, overwrite=ow, charge=0.3 if coul else None, coulomb=coul)
    rng = np.random.default_rng(3)
    h = 1e-5
    for _ in range(5):
        p = np.array([rng.uniform(-1, 1), rng.uniform(-1, 1), rng.uniform(3.2, 5.0)])
        f = O.eval_tip_surface(st, p[None])[0, :3]
        for a in range(3):
            dp = np.zeros(3)
            dp[a] = h
            ep = O.eval_tip_surface(st, (p + dp)[None])[0, 3]
            em = O.eval_tip_surface(st, (p - dp)[None])[0, 3]
            assert f[a] == pytest.approx(-(ep - em) / (2 * h), rel=1e-6, abs=1e-9)


def test_total_force_includes_spring_and_tip_dummy(tmp_path):
    st = _one_atom_setup(tmp_path, overwrite="O C lj 0.0 1.0")     # surface term switched off
    dummy = np.array([[0.3, -0.2, 10.0]])
    tip = np.array([[0.5, 0.1, 6.5]])
    fe = O.eval_total(st, dummy, tip)[0]
    k = 0.72
    rxy = dummy[0, :2] - tip[0, :2]
    np.testing.assert_allclose(fe[:2], 2 * k * rxy - _td_lj_force(dummy[0], tip[0])[:2], rtol=1e-12)
    h = 1e-6
    for a in range(3):
        d = np.zeros(3)
        d[a] = h
        ep = O.eval_total(st, dummy, tip + d)[0, 3]
        em = O.eval_total(st, dummy, tip - d)[0, 3]
        assert fe[a] == pytest.approx(-(ep - em) / (2 * h), rel=1e-6, abs=1e-9)


def _td_lj_force(dummy, tip):
    eps = math.sqrt(20.0 * 0.192)
    sig = (3.55 + 3.15) / 2
    r = dummy - tip
    r2 = r @ r
    a = 4 * eps * sig ** 12 / r2 ** 6
    b = 4 * eps * sig ** 6 / r2 ** 3
    return (12 * a - 6 * b) / r2 * r


def test_trilinear_exact_on_trilinear_field(tmp_path):
    st = _one_atom_setup(tmp_path)
    n = (5, 6, 7)
    h = np.array([0.3, 0.25, 0.2])
    off = np.array([-0.5, -0.4, 1.0])
    I, J, K = np.meshgrid(*(np.arange(v) for v in n), indexing="ij")
    X, Y, Z = off[0] + I * h[0], off[1] + J * h[1], off[2] + K * h[2]
    fe = np.stack([1 + 2 * X - Y + 0.5 * X * Y * Z, 3 * Z - X * Z, Y * Z + 0.25, 2 - X + Y - Z + X * Y], -1)
    st.grids = [{"n": n, "basis": np.diag(h), "offset": off, "periodic": False, "orthogonal": True,
                 "fe": fe.reshape(-1, 4)}]
    st.inter = [{"kind": O.OR_GRID, "grid": 0}]
    rng = np.random.default_rng(7)
    p = np.stack([rng.uniform(off[a], off[a] + (n[a] - 1) * h[a] - 1e-9, 50) for a in range(3)], 1)
    got = O.eval_tip_surface(st, p)
    x, y, z = p.T
    want = np.stack([1 + 2 * x - y + 0.5 * x * y * z, 3 * z - x * z, y * z + 0.25, 2 - x + y - z + x * y], -1)
    np.testing.assert_allclose(got, want, atol=1e-12)


def test_fft_electrostatics_plane_wave():
    n = (48, 40, 32)
    h = np.array([0.1, 0.125, 0.15])
    L = h * np.array(n)
    kvec = np.array([2 / L[0], 1 / L[1], 3 / L[2]])           # cycles / A on the reciprocal lattice
    I, J, K = np.meshgrid(*(np.arange(v) for v in n), indexing="ij")
    X, Y, Z = I * h[0], J * h[1], K * h[2]
    phase = 2 * np.pi * (kvec[0] * X + kvec[1] * Y + kvec[2] * Z)
    pot = np.cos(phase)
    q, sigma = -0.3, 0.5
    g = O.electrostatic_force_grid(pot, np.diag(h), np.zeros(3), q, sigma)
    damp = q * math.exp(-2 * math.pi ** 2 * sigma ** 2 * float(kvec @ kvec))
    e = g["fe"][:, 3].reshape(n)
    np.testing.assert_allclose(e, damp * np.cos(phase), atol=1e-9 * abs(damp))
    for c in range(3):
        f = g["fe"][:, c].reshape(n)
        np.testing.assert_allclose(f, damp * 2 * np.pi * kvec[c] * np.sin(phase), atol=1e-8 * abs(damp))


def test_scale_inputs_regenerate(tmp_path):
    """The scale cases' inputs come back from the seeded generators byte for
    byte (SHA-256 of every file the reference read)."""
    from conftest import materialize_scale
    for case in ("c3_batch", "c4_full"):
        materialize_scale(case, str(tmp_path / case))


def test_oracle_matches_c3_batch(tmp_path):
    """The oracle on all 64 molecules of the batched workload (2 x 2 pixels,
    40 heights each) equals the reference's records bit for bit."""
    from conftest import materialize_scale, scale_records
    paths = materialize_scale("c3_batch", str(tmp_path))
    shapes, ref = scale_records("c3_batch")
    got = np.concatenate([O.scan(O.setup(p))[0] for p in paths])
    np.testing.assert_array_equal(got[:, 3], ref[:, 3])
    np.testing.assert_array_equal(got, ref)
