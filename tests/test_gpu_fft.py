"""GPU tests of the FFT path: mafm_fft3d against numpy.fft (same sign and
normalisation as kiss_fft, kiss_fft.c:359-362 / fft.cpp:74-84) and
mafm_potential_grid against the oracle's restatement of
ElectrostaticPotentialInteraction (interactions.cpp:53-234)."""
import os

import numpy as np
import pytest

from conftest import GOLDEN, materialize

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", [(16, 20, 30), (24, 24, 64), (7, 11, 13), (1, 5, 9), (160, 160, 80)])
def test_fft3d_matches_numpy(gpu_ctx, shape):
    rng = np.random.default_rng(sum(shape))
    x = rng.normal(size=shape) + 1j * rng.normal(size=shape)
    got = gpu_ctx.fft3d(x, inverse=False)
    want = np.fft.fftn(x)
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()
    back = gpu_ctx.fft3d(got, inverse=True)
    assert np.abs(back - x).max() <= 1e-12 * np.abs(x).max()


def test_fft3d_rejects_unsupported_length(gpu_ctx):
    from mechafm_amd import abi
    with pytest.raises(abi.MafmError):
        gpu_ctx.fft3d(np.zeros((17, 4, 4), dtype=complex))


def _cube_potential(tmp_path):
    from oracle import oracle as O
    inp = materialize("c4_mini", str(tmp_path))
    cube = O.read_cube(os.path.join(str(tmp_path), "hartree.cube"))
    pot = cube.values * (-627.50961)
    return pot, cube.vox.T.copy(), cube.origin


def test_potential_grid_matches_oracle(gpu_ctx, tmp_path):
    from oracle import oracle as O
    pot, basis, origin = _cube_potential(tmp_path)
    h = gpu_ctx.potential_grid(pot, basis, origin, -0.021, 0.5)
    got = gpu_ctx.grid_download(h, pot.shape)
    want = O.electrostatic_force_grid(pot, basis, origin, -0.021, 0.5)["fe"]
    assert np.abs(got - want).max() <= 1e-10 * np.abs(want).max()


def test_potential_grid_imaginary_check(gpu_ctx):
    """A 0.4 A voxel leaves the tip Gaussian's Nyquist content above the
    reference's 1e-8 threshold: the oracle raises, the library reports
    MAFM_ERR_FFT_IMAG."""
    from mechafm_amd import abi
    from oracle import oracle as O
    n = (20, 20, 40)
    basis = np.diag([0.4, 0.4, 0.4])
    rng = np.random.default_rng(5)
    pot = rng.normal(size=n)
    with pytest.raises(RuntimeError):
        O.electrostatic_force_grid(pot, basis, np.zeros(3), -0.021, 0.5)
    with pytest.raises(abi.MafmError) as e:
        gpu_ctx.potential_grid(pot, basis, np.zeros(3), -0.021, 0.5)
    assert e.value.status == 5
