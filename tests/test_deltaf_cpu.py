"""Frequency-shift images: the oracle pinned by hand-derived windows and
closed forms, and libmafm's host-side window (mafm_deltaf_weights) against
the oracle bit for bit (no GPU needed)."""
import math

import numpy as np
import pytest

from oracle import deltaf as D


def test_window_dz_0p1_lacks_minus_dz():
    """A = 0.5, dz = 0.1: zosc = 0 .. 0.5 mirrored as zosc[-1:1:-1], so the
    window is -0.4 -0.3 -0.2 0 0.1 0.2 0.3 0.4 (no -0.1) after the ends are
    dropped; no half-step adjustment (0.5 // 0.1 == 4.0, 6.0 % 5.0 == 1)."""
    zpos = D.file_heights(10.0, 0.1, 41)
    w, fac, dz, amp = D.weights(zpos)
    assert amp == 0.5 and abs(dz - 0.1) < 1e-15
    z = [-0.4, -0.3, -0.2, 0.0, 0.1, 0.2, 0.3, 0.4]
    want = [v / math.sqrt(abs(0.25 - v * v)) for v in z]
    np.testing.assert_allclose(w, want, rtol=1e-12, atol=0)
    # a symmetric window would sum to 0; this one keeps the +0.1 term
    assert abs(w.sum() - 0.1 / math.sqrt(0.24)) < 1e-12


def test_window_dz_0p25_adjusts_amplitude():
    """dz = 0.25: 0.5 / 0.25 == 0.5 // 0.25 == 2, (1 + 2) % (1 + 2) == 0 < 0.01,
    so A becomes 0.625 and zosc = 0, .25, .5, .75 -> window -.5, 0, .25, .5."""
    zpos = D.file_heights(10.0, 0.25, 20)
    w, fac, dz, amp = D.weights(zpos)
    assert amp == 0.625
    a2 = 0.625 ** 2
    want = [-0.5 / math.sqrt(a2 - 0.25), 0.0, 0.25 / math.sqrt(a2 - 0.0625), 0.5 / math.sqrt(a2 - 0.25)]
    np.testing.assert_allclose(w, want, rtol=1e-12, atol=0)


def test_prefactor_closed_form():
    """f0 / (2k) * 2 / (pi A^2), k = 1800 N/m in kcal/mol/A^2 (1.438978)."""
    _, fac, _, _ = D.weights(D.file_heights(10.0, 0.1, 41))
    k = 1800 * 1.438978
    assert fac == pytest.approx(25e3 / (2 * k) * 2 / (math.pi * 0.25), rel=1e-14)
    # 25e3 / (2 * 2590.16) * 2.5465 = 12.289 Hz per kcal/mol/A
    assert fac == pytest.approx(12.2891, rel=1e-4)


def test_images_of_linear_force():
    """Fz = a + b z: image pl = -fac dz (a sum W + b sum W z_(pl+t)), exact
    up to rounding; the topmost height never enters an image."""
    nz, zhigh, dz0 = 30, 9.0, 0.1
    zpos = D.file_heights(zhigh, dz0, nz)
    w, fac, dz, _ = D.weights(zpos)
    a, b = 0.7, -2.5
    fz = np.tile(a + b * np.array(zpos), (3, 2, 1))
    img, zlow = D.images(fz, zpos)
    assert img.shape == (nz - len(w), 3, 2)
    for pl in range(len(img)):
        want = -fac * dz * sum(w[t] * (a + b * zpos[pl + t]) for t in range(len(w)))
        np.testing.assert_allclose(img[pl], want, rtol=1e-12)
    np.testing.assert_array_equal(zlow, zpos[:len(img)])


def test_records_order():
    fz = np.arange(2 * 3 * 4, dtype=float)        # index (i*ny + j)*nz + k
    asc = D.fz_from_records(fz, 2, 3, 4)
    assert asc[1, 2, 0] == fz[(1 * 3 + 2) * 4 + 3] and asc[1, 2, 3] == fz[(1 * 3 + 2) * 4 + 0]


CASES = [(10.0, 0.1, 41, 0.5), (10.0, 0.25, 20, 0.5), (14.4, 0.1, 40, 0.5), (8.0, 0.05, 60, 0.5),
         (10.0, 0.125, 33, 0.5), (10.0, 0.2, 25, 0.5), (9.95, 0.07, 50, 0.3), (10.0, 0.1, 41, 0.75),
         (12.0, 0.04, 80, 0.5), (10.0, 0.3, 15, 0.9), (7.0, 0.15, 21, 0.45)]


@pytest.mark.parametrize("zhigh,dz,nz,amp", CASES)
def test_library_window_matches_oracle(zhigh, dz, nz, amp):
    """libmafm's C++ window (Python float % and //, numpy pairwise mean and
    arange restated) equals the numpy oracle bit for bit."""
    from mechafm_amd import abi
    got = abi.deltaf_weights(nz, zhigh, dz, amplitude=amp)
    w, fac, dzm, a = D.weights(D.file_heights(zhigh, dz, nz), amplitude=amp)
    assert got["dz"] == dzm and got["amplitude"] == a and got["prefactor"] == fac
    assert got["n_weights"] == len(w) and got["n_images"] == max(nz - len(w), 0)
    np.testing.assert_array_equal(got["weights"], w)


def test_library_rejects_short_scans():
    from mechafm_amd import abi
    with pytest.raises(abi.MafmError):
        abi.deltaf_weights(1, 10.0, 0.1)
    with pytest.raises(abi.MafmError):
        abi.deltaf_weights(40, 10.0, 0.001)      # window of 501 heights
