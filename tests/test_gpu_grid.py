"""ForceGrid::interpolate on the device (mafm_grid_interpolate) against the
oracle's lookup, which follows the reference's operation order
(force_grid.cpp:88-275) and is pinned by the reference's 512 grid-lookup
goldens (tests/test_oracle.py), and against a closed form: a trilinear
field is reproduced exactly.  Periodic grids wrap (pbcGridPoint), skewed
bases go through the inverse basis, and a position past a non-periodic grid
returns MAFM_ERR_OUT_OF_GRID where the reference stops."""
import numpy as np
import pytest

from mechafm_amd import abi
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _setup(tmp_path):
    d = str(tmp_path)
    with open(f"{d}/m.xyz", "w") as f:
        f.write("1\n\nC 0.0 0.0 0.0\n")
    with open(f"{d}/p.dat", "w") as f:
        f.write("atom C 0.07 3.55 12.011 0.0\natom X 20.0 3.55 12.011 0.021\natom O 0.192 3.15 15.999 -0.021\n")
        f.write("harm 0.72 0.0\n")
    with open(f"{d}/in.scan", "w") as f:
        f.write("xyzfile m.xyz\nparamfile p.dat\ntipatom O\ndummyatom X\nunits kcal/mol\nminterm f\n")
        f.write("area 1.0 1.0\ncenter 0.0 0.0\ncoulomb off\n")
    return O.setup(f"{d}/in.scan")


def _oracle(tmp_path, n, basis, off, fe, periodic, p):
    st = _setup(tmp_path)
    st.grids = [{"n": n, "basis": basis, "offset": off, "periodic": periodic,
                 "orthogonal": bool(np.count_nonzero(basis - np.diag(np.diag(basis))) == 0), "fe": fe}]
    st.inter = [{"kind": O.OR_GRID, "grid": 0}]
    return O.eval_tip_surface(st, p)


def test_trilinear_field_and_the_oracle(gpu_ctx, tmp_path):
    n = (5, 6, 7)
    h = np.array([0.3, 0.25, 0.2])
    off = np.array([-0.5, -0.4, 1.0])
    I, J, K = np.meshgrid(*(np.arange(v) for v in n), indexing="ij")
    X, Y, Z = off[0] + I * h[0], off[1] + J * h[1], off[2] + K * h[2]
    fe = np.stack([1 + 2 * X - Y + 0.5 * X * Y * Z, 3 * Z - X * Z, Y * Z + 0.25, 2 - X + Y - Z + X * Y],
                  -1).reshape(-1, 4)
    g = gpu_ctx.grid_upload(n, np.diag(h), off, fe, periodic=False)
    rng = np.random.default_rng(11)
    p = np.stack([rng.uniform(off[a], off[a] + (n[a] - 1) * h[a] - 1e-9, 400) for a in range(3)], 1)
    got = gpu_ctx.grid_interpolate(g, p)
    np.testing.assert_array_equal(got, _oracle(tmp_path, n, np.diag(h), off, fe, False, p))
    x, y, z = p.T
    want = np.stack([1 + 2 * x - y + 0.5 * x * y * z, 3 * z - x * z, y * z + 0.25, 2 - x + y - z + x * y], -1)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-12)
    # past the last grid point: the reference's error, as a status
    with pytest.raises(abi.MafmError, match="past the grid"):
        gpu_ctx.grid_interpolate(g, np.array([[0.0, 0.0, off[2] + (n[2] - 1) * h[2] + 0.05]]))
    gpu_ctx.grid_free(g)


def test_periodic_skewed_grid_matches_the_oracle(gpu_ctx, tmp_path):
    n = (8, 9, 10)
    basis = np.array([[0.4, 0.1, 0.0], [0.0, 0.35, 0.0], [0.05, 0.02, 0.3]])
    off = np.array([0.1, -0.2, 0.3])
    rng = np.random.default_rng(12)
    fe = rng.normal(0, 1, (int(np.prod(n)), 4))
    g = gpu_ctx.grid_upload(n, basis, off, fe, periodic=True)
    p = rng.uniform(-6, 6, (500, 3))                 # well outside the cell: every lookup wraps
    got = gpu_ctx.grid_interpolate(g, p)
    want = _oracle(tmp_path, n, basis, off, fe, True, p)
    np.testing.assert_allclose(got, want, rtol=0, atol=1e-13)
    gpu_ctx.grid_free(g)
