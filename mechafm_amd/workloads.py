"""Seeded synthetic inputs in MechAFM's own file formats.

Writes ``input.scan`` (keyword file, parse.cpp:83-474), the parameter file
(``atom``/``harm``/``pair_ovwrt`` lines, parse.cpp:588-697), the xyz file
(4/5/6 columns, parse.cpp:477-585) and, for the external-potential workload, a
Gaussian cube file (cube_io.cpp:89-139) holding a periodic Hartree potential.

Workloads (all sizes are the reference's own scan-grid formulas,
mechafm.cpp:47-49):

* ``c1``  C96 hexagonal graphene flake, LJ only, 101 x 101 px x 41 z, FP64.
* ``c2``  40-atom planar C/N/O/H molecule, Morse tip-surface + Coulomb,
          ``rigidgrid on`` (grid spacing 0.05 A), 128 x 128 px x 40 z.
* ``c3``  64 planar molecules with N ~ U[20, 400], LJ + Coulomb,
          64 x 64 px x 40 z each (batched in one launch).
* ``c4``  2000-atom periodic slab (40 x 40 x 20 A cell), ``vdw_pbc on``,
          Hartree potential on a 160 x 160 x 80 grid from Gaussian charges,
          256 x 256 px x 40 z.
* ``c5``  150-atom molecule, LJ + Coulomb, 128 x 128 px x 30 z (adjoint).
* ``fx``  flexible on: a hydrogen-terminated coronene (C24H12, edge H
          tethered in xy, two carbons fixed) beside a separate ammonia
          molecule, LJ + Coulomb with xyz charges, so the interaction list
          holds bonds, angles, proper and improper dihedrals, non-bonded
          pairs between the two molecules, the substrate wall and xy tethers.
* ``fx-example``  the reference's example graphene job with flexible on.

Every workload has a ``mini`` variant that keeps the physics and shrinks the
pixel grid so a CPU FP64 run finishes in seconds.
"""
from __future__ import annotations

import math
import os
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

# Atom parameters (eps kcal/mol, sigma A, mass amu, charge e).  C, H, O, X are
# the reference's examples/parameters.dat:4-8 values; N is a CHARMM-like
# nitrogen chosen for the synthetic organic molecules.
BASE_ATOMS: Dict[str, Tuple[float, float, float, float]] = {
    "C": (0.07000, 3.55000, 12.01100, 0.00000),
    "H": (0.03000, 2.42000, 1.00800, 0.00000),
    "N": (0.20000, 3.30000, 14.00700, 0.00000),
    "X": (20.00000, 3.55000, 12.01100, 0.02100),
    "O": (0.19200, 3.15000, 15.99900, -0.02100),
}
HARM = (0.72000, 0.00)          # examples/parameters.dat:20
BOHR = 0.52917721092            # cube_io.hpp:17
ATOMIC_NUMBER = {"H": 1, "C": 6, "N": 7, "O": 8, "X": 0}


@dataclass
class Molecule:
    types: List[str]
    pos: np.ndarray                       # (n, 3) float64
    charges: Optional[np.ndarray] = None  # (n,) float64 or None
    fixed: Optional[np.ndarray] = None    # (n,) int: xyz sixth column (flexible on)


@dataclass
class Workload:
    """One scan job: its files on disk and the options written into them."""
    name: str
    input_path: str
    options: Dict[str, str]
    molecule: Molecule
    extra: Dict[str, object] = field(default_factory=dict)


# --------------------------------------------------------------------------
# geometry generators
# --------------------------------------------------------------------------
def hex_flake(n_rings: int, cc: float = 1.42) -> Molecule:
    """Hexagonal graphene flake of the coronene series: 6 n^2 carbons.

    The union of the six vertices of every hexagon whose centre lies within
    hex distance n-1 of the origin (n=2 coronene, n=4 gives C96)."""
    a1 = np.array([math.sqrt(3.0) * cc, 0.0])
    a2 = np.array([math.sqrt(3.0) * cc / 2.0, 1.5 * cc])
    verts = []
    for u in range(-n_rings, n_rings + 1):
        for v in range(-n_rings, n_rings + 1):
            w = -u - v
            if max(abs(u), abs(v), abs(w)) > n_rings - 1:
                continue
            c = u * a1 + v * a2
            for k in range(6):
                ang = math.radians(30.0 + 60.0 * k)
                verts.append((round(c[0] + cc * math.cos(ang), 6), round(c[1] + cc * math.sin(ang), 6)))
    pts = np.array(sorted(set(verts)), dtype=np.float64)
    pos = np.zeros((len(pts), 3))
    pos[:, :2] = pts
    return Molecule(["C"] * len(pts), pos)


def planar_molecule(n_atoms: int, rng: np.random.Generator,
                    fractions=(("C", 0.50), ("N", 0.15), ("O", 0.10), ("H", 0.25)),
                    bond=(1.0, 1.5), min_dist=1.0, charge_sigma: float = 0.2) -> Molecule:
    """Seeded planar graph growth: heavy atoms first (valence <= 3), then H on
    unsaturated heavy atoms.  Bonds ~ U[1.0, 1.5] A, no two atoms closer than
    ``min_dist``.  Charges ~ N(0, charge_sigma) shifted to zero net charge."""
    counts = {}
    left = n_atoms
    for i, (t, f) in enumerate(fractions):
        c = left if i == len(fractions) - 1 else int(round(f * n_atoms))
        c = min(c, left)
        counts[t] = c
        left -= c
    n_h = counts.get("H", 0)
    heavy_types = [t for t, _ in fractions if t != "H" for _ in range(counts[t])]
    rng.shuffle(heavy_types)
    n_heavy = len(heavy_types)
    buf = np.zeros((n_atoms + 1, 2))
    count = [1]
    deg = [0]

    class _Pos:
        def __getitem__(self, k):
            return buf[k]

        def __len__(self):
            return count[0]

        def append(self, p):
            buf[count[0]] = p
            count[0] += 1

    pos = _Pos()

    def try_place(parent: int) -> Optional[np.ndarray]:
        live = buf[: count[0]]
        for _ in range(40):
            ang = rng.uniform(0, 2 * math.pi)
            b = rng.uniform(*bond)
            p = buf[parent] + b * np.array([math.cos(ang), math.sin(ang)])
            d2 = ((live - p) ** 2).sum(axis=1)
            d2[parent] = 1e18
            dmin = math.sqrt(d2.min())
            if dmin >= max(min_dist, b * 0.999) or dmin >= 1.3:
                return p
        return None

    guard = 0
    while len(pos) < n_heavy:
        guard += 1
        if guard > 200000:
            raise RuntimeError("molecule growth stalled")
        cands = [i for i in range(len(pos)) if deg[i] < 3]
        parent = int(rng.choice(cands))
        p = try_place(parent)
        if p is None:
            deg[parent] = 3
            continue
        pos.append(p)
        deg.append(1)
        deg[parent] += 1
    types = list(heavy_types)
    n_added = 0
    guard = 0
    while n_added < n_h:
        guard += 1
        if guard > 200000:
            raise RuntimeError("hydrogen placement stalled")
        cands = [i for i in range(n_heavy) if deg[i] < 3] or list(range(n_heavy))
        parent = int(rng.choice(cands))
        p = try_place(parent)
        if p is None:
            deg[parent] = 4
            continue
        pos.append(p)
        deg.append(1)
        deg[parent] += 1
        types.append("H")
        n_added += 1
    xyz = np.zeros((len(pos), 3))
    xyz[:, :2] = buf[: len(pos)]
    q = rng.normal(0.0, charge_sigma, len(pos))
    q -= q.mean()
    return Molecule(types, xyz, q)


def random_slab(n_atoms: int, cell: Tuple[float, float], thickness: float,
                rng: np.random.Generator, min_dist: float = 1.0,
                charge_sigma: float = 0.2) -> Molecule:
    """Random sequential addition of C/N/O/H atoms into a periodic slab
    [0,Lx) x [0,Ly) x [0,thickness] with a minimum image distance."""
    lx, ly = cell
    cell_sz = min_dist
    nbx, nby = int(lx // cell_sz), int(ly // cell_sz)
    nbz = int(thickness // cell_sz) + 1
    buckets: Dict[Tuple[int, int, int], List[int]] = {}
    pts: List[np.ndarray] = []
    tries = 0
    while len(pts) < n_atoms:
        tries += 1
        if tries > 200 * n_atoms:
            raise RuntimeError("slab packing stalled")
        p = np.array([rng.uniform(0, lx), rng.uniform(0, ly), rng.uniform(0, thickness)])
        bx, by, bz = int(p[0] / lx * nbx) % nbx, int(p[1] / ly * nby) % nby, min(int(p[2] / cell_sz), nbz - 1)
        ok = True
        for ix in (-1, 0, 1):
            for iy in (-1, 0, 1):
                for iz in (-1, 0, 1):
                    for j in buckets.get(((bx + ix) % nbx, (by + iy) % nby, bz + iz), ()):
                        d = pts[j] - p
                        d[0] -= lx * round(d[0] / lx)
                        d[1] -= ly * round(d[1] / ly)
                        if d @ d < min_dist * min_dist:
                            ok = False
                            break
                    if not ok:
                        break
                if not ok:
                    break
            if not ok:
                break
        if ok:
            buckets.setdefault((bx, by, bz), []).append(len(pts))
            pts.append(p)
    xyz = np.array(pts)
    types = list(rng.choice(["C", "N", "O", "H"], size=n_atoms, p=[0.5, 0.15, 0.10, 0.25]))
    q = rng.normal(0.0, charge_sigma, n_atoms)
    q -= q.mean()
    return Molecule(types, xyz, q)


# --------------------------------------------------------------------------
# periodic Hartree potential from Gaussian charges (FFT Poisson)
# --------------------------------------------------------------------------
def gaussian_hartree_potential(mol: Molecule, cell: Sequence[float], n_grid: Sequence[int],
                               width: float = 0.5, cutoff_sigmas: float = 5.0) -> np.ndarray:
    """Electrostatic potential of Gaussian charges on a periodic orthorhombic
    grid, returned as the *Hartree potential of the electrons* in Hartree
    (i.e. minus the atomic-unit potential), the sign/units convention that the
    reference undoes with its -627.50961 factor (simulation.cpp:348-355).

    rho is deposited in real space (cut at ``cutoff_sigmas``), Poisson is
    solved spectrally: phi_k = 4 pi rho_k / |k|^2, phi_{k=0} = 0."""
    nx, ny, nz = (int(v) for v in n_grid)
    lx, ly, lz = (float(v) for v in cell)
    h = np.array([lx / nx, ly / ny, lz / nz])
    rho = np.zeros((nx, ny, nz))
    norm = 1.0 / (width * math.sqrt(2.0 * math.pi)) ** 3
    rc = cutoff_sigmas * width
    span = [int(math.ceil(rc / h[d])) for d in range(3)]
    offs = [np.arange(-span[d], span[d] + 1) for d in range(3)]
    q = mol.charges if mol.charges is not None else np.zeros(len(mol.types))
    for p, qi in zip(mol.pos, q):
        base = np.floor(p / h).astype(int)
        gx, gy, gz = (base[d] + offs[d] for d in range(3))
        dx = gx * h[0] - p[0]
        dy = gy * h[1] - p[1]
        dz = gz * h[2] - p[2]
        r2 = dx[:, None, None] ** 2 + dy[None, :, None] ** 2 + dz[None, None, :] ** 2
        val = qi * norm * np.exp(-0.5 * r2 / width ** 2)
        rho[np.ix_(gx % nx, gy % ny, gz % nz)] += val
    rho_k = np.fft.fftn(rho)
    kx = 2 * np.pi * np.fft.fftfreq(nx, d=h[0])
    ky = 2 * np.pi * np.fft.fftfreq(ny, d=h[1])
    kz = 2 * np.pi * np.fft.fftfreq(nz, d=h[2])
    k2 = kx[:, None, None] ** 2 + ky[None, :, None] ** 2 + kz[None, None, :] ** 2
    k2[0, 0, 0] = 1.0
    phi_k = 4.0 * np.pi * rho_k / k2
    phi_k[0, 0, 0] = 0.0
    phi_ang = np.real(np.fft.ifftn(phi_k))   # e / A
    phi_au = phi_ang * BOHR                  # Hartree / e
    return -phi_au


# --------------------------------------------------------------------------
# file writers
# --------------------------------------------------------------------------
def write_xyz(path: str, mol: Molecule, with_charges: bool) -> None:
    with open(path, "w") as f:
        f.write(f"{len(mol.types)}\n\n")
        for i, t in enumerate(mol.types):
            x, y, z = mol.pos[i]
            if mol.fixed is not None:
                q = mol.charges[i] if (with_charges and mol.charges is not None) else 0.0
                f.write(f"{t:<4s} {x:.8f} {y:.8f} {z:.8f} {q:.8f} {int(mol.fixed[i])}\n")
            elif with_charges:
                f.write(f"{t:<4s} {x:.8f} {y:.8f} {z:.8f} {mol.charges[i]:.8f}\n")
            else:
                f.write(f"{t:<4s} {x:.8f} {y:.8f} {z:.8f}\n")


def write_params(path: str, atoms: Dict[str, Tuple[float, float, float, float]],
                 harm=HARM, overwrites: Sequence[str] = (), extra: Sequence[str] = ()) -> None:
    with open(path, "w") as f:
        f.write("# name | epsilon | sigma | mass | charge\n")
        for t, (e, s, m, q) in atoms.items():
            f.write(f"atom {t} {e:.5f} {s:.5f} {m:.5f} {q:.5f}\n")
        f.write(f"harm {harm[0]:.5f} {harm[1]:.2f}\n")
        for line in overwrites:
            f.write(f"pair_ovwrt {line}\n")
        for line in extra:
            f.write(line + "\n")


def write_input(path: str, opts: Dict[str, str]) -> None:
    with open(path, "w") as f:
        for k, v in opts.items():
            f.write(f"{k:<24s} {v}\n")


def write_cube(path: str, values: np.ndarray, cell: Sequence[float], mol: Molecule,
               origin=(0.0, 0.0, 0.0), voxels: Optional[np.ndarray] = None, bohr: bool = False) -> None:
    """Cube file, x-major / z-fastest values, 6 per line.  Voxel vectors are
    ``voxels`` (rows) or the orthorhombic cell / counts.  Angstrom units are
    flagged by negative voxel counts on all three axes; ``bohr`` writes
    positive counts and every length in Bohr (cube_io.cpp:107-118)."""
    nx, ny, nz = values.shape
    vox = np.diag([cell[0] / nx, cell[1] / ny, cell[2] / nz]) if voxels is None else np.asarray(voxels, float)
    scale = 1.0 / BOHR if bohr else 1.0
    sign = 1 if bohr else -1
    o = np.asarray(origin, float) * scale
    lines = ["synthetic Hartree potential",
             "units: " + ("Bohr" if bohr else "Angstrom") + ", Hartree"]
    lines.append(f"{len(mol.types):5d} {o[0]:14.8f} {o[1]:14.8f} {o[2]:14.8f}")
    for a, n in enumerate((nx, ny, nz)):
        v = vox[a] * scale
        lines.append(f"{sign * n:5d} {v[0]:14.8f} {v[1]:14.8f} {v[2]:14.8f}")
    for t, p in zip(mol.types, mol.pos):
        q = np.asarray(p) * scale
        lines.append(f"{ATOMIC_NUMBER.get(t, 0):5d} {0.0:12.6f} {q[0]:14.8f} {q[1]:14.8f} {q[2]:14.8f}")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
        flat = values.reshape(nx * ny, nz)
        for row in flat:
            for s in range(0, nz, 6):
                f.write(" ".join(f"{v:13.5e}" for v in row[s:s + 6]) + "\n")


# --------------------------------------------------------------------------
# workloads
# --------------------------------------------------------------------------
def _base_options(**kw) -> Dict[str, str]:
    o = {
        "xyzfile": "molecule.xyz", "paramfile": "parameters.dat",
        "tipatom": "O", "dummyatom": "X", "units": "kcal/mol",
        "minterm": "ef", "etol": "0.01", "ftol": "0.01", "dt": "0.001", "maxsteps": "5000",
        "minimiser": "FIRE", "integrator": "midpoint", "coulomb": "off",
        "rigidgrid": "off", "flexible": "off",
        "area": "10.0 10.0", "zhigh": "10.0", "zlow": "6.0",
        "dx": "0.1", "dy": "0.1", "dz": "0.1", "bufsize": "1000", "gzip": "off",
        "statistics": "on",
    }
    o.update({k: str(v) for k, v in kw.items()})
    return o


def morse_overwrites(tip: str, types: Sequence[str], atoms=BASE_ATOMS, a: float = 1.8) -> List[str]:
    """Morse tip-surface pairs: De = mixed epsilon, re = mixed sigma * 2^(1/6)."""
    out = []
    for t in types:
        e = math.sqrt(atoms[t][0] * atoms[tip][0])
        s = 0.5 * (atoms[t][1] + atoms[tip][1])
        out.append(f"{tip} {t} morse {e:.6f} {a:.4f} {s * 2 ** (1 / 6):.6f}")
    return out


def _emit(name: str, outdir: str, mol: Molecule, opts: Dict[str, str], charges_in_xyz: bool,
          overwrites: Sequence[str] = (), atoms=BASE_ATOMS, extra=None, param_lines: Sequence[str] = ()) -> Workload:
    os.makedirs(outdir, exist_ok=True)
    write_xyz(os.path.join(outdir, opts["xyzfile"]), mol, charges_in_xyz)
    write_params(os.path.join(outdir, opts["paramfile"]), atoms, HARM, overwrites, param_lines)
    path = os.path.join(outdir, "input.scan")
    write_input(path, opts)
    return Workload(name, path, dict(opts), mol, dict(extra or {}))


def make_c1(outdir: str, mini: bool = False, **over) -> Workload:
    mol = hex_flake(4)
    opts = _base_options()
    if mini:
        opts.update(area="1.0 1.0")
    opts.update({k: str(v) for k, v in over.items()})
    return _emit("c1-mini" if mini else "c1", outdir, mol, opts, False)


def make_c2(outdir: str, mini: bool = False, seed: int = 1236, **over) -> Workload:
    rng = np.random.default_rng(seed)
    mol = planar_molecule(40, rng)
    opts = _base_options(coulomb="on", rigidgrid="on", area="12.75 12.75", zlow="6.1")
    if mini:
        opts.update(area="1.0 1.0")
    opts.update({k: str(v) for k, v in over.items()})
    ow = morse_overwrites("O", ["C", "N", "O", "H"])
    return _emit("c2-mini" if mini else "c2", outdir, mol, opts, True, ow)


def c3_molecules(seed: int = 1237, count: int = 64, nmin: int = 20, nmax: int = 400) -> List[Molecule]:
    rng = np.random.default_rng(seed)
    sizes = rng.integers(nmin, nmax + 1, size=count)
    return [planar_molecule(int(n), rng) for n in sizes]


def make_c3(outdir: str, mini: bool = False, seed: int = 1237, count: int = 64, jitter: float = 0.0,
            **over) -> List[Workload]:
    """The 64-molecule batch.  jitter > 0 moves every atom by a seeded
    uniform offset in [-jitter, jitter] A per coordinate (the xyz files keep
    8 decimals, so 1e-8 is the smallest change they can carry)."""
    mols = c3_molecules(seed, count)
    if jitter > 0:
        rng = np.random.default_rng(seed + 7919)
        for mol in mols:
            mol.pos = mol.pos + rng.uniform(-jitter, jitter, size=mol.pos.shape)
    out = []
    for m, mol in enumerate(mols):
        opts = _base_options(coulomb="on", area="6.35 6.35", zlow="6.1")
        if mini:
            opts.update(area="0.5 0.5")
        opts.update({k: str(v) for k, v in over.items()})
        out.append(_emit(f"c3-{m:02d}", os.path.join(outdir, f"mol{m:02d}"), mol, opts, True))
    return out


def make_c4(outdir: str, mini: bool = False, seed: int = 1238, **over) -> Workload:
    rng = np.random.default_rng(seed)
    if mini:
        cell, n_atoms, n_grid, thick = (6.0, 6.0, 16.0), 45, (24, 24, 64), 4.0
        area, dxy = "0.8 0.8", "0.2"
    else:
        cell, n_atoms, n_grid, thick = (40.0, 40.0, 20.0), 2000, (160, 160, 80), 4.0
        area, dxy = "39.84375 39.84375", "0.15625"
    mol = random_slab(n_atoms, cell[:2], thick, rng)
    pot = gaussian_hartree_potential(mol, cell, n_grid)
    os.makedirs(outdir, exist_ok=True)
    write_cube(os.path.join(outdir, "hartree.cube"), pot, cell, mol)
    opts = _base_options(area=area, dx=dxy, dy=dxy, zhigh="14.4", zlow="10.5",
                         vdw_pbc="on", use_external_potential="on",
                         e_potential_file="hartree.cube")
    opts.update({k: str(v) for k, v in over.items()})
    return _emit("c4-mini" if mini else "c4", outdir, mol, opts, False,
                 extra={"cell": cell, "n_grid": n_grid})


def make_c5(outdir: str, mini: bool = False, seed: int = 1239, **over) -> Workload:
    rng = np.random.default_rng(seed)
    mol = planar_molecule(150, rng)
    opts = _base_options(coulomb="on", area="12.75 12.75", zhigh="10.0", zlow="7.05")
    if mini:
        opts.update(area="0.3 0.3")
    opts.update({k: str(v) for k, v in over.items()})
    return _emit("c5-mini" if mini else "c5", outdir, mol, opts, True)


# File frame of a system whose surface normal is x or y: the reference rotates
# positions to (y, z, x) for x and (z, x, y) for y (System::rotateCoordAxes,
# system.cpp:68-89), so a system built with its normal along z is written with
# these column orders to come back unchanged.
FILE_COLUMNS = {"x": [2, 0, 1], "y": [1, 2, 0]}


def rotated_molecule(mol: Molecule, normal: str) -> Molecule:
    return Molecule(list(mol.types), mol.pos[:, FILE_COLUMNS[normal]].copy(),
                    None if mol.charges is None else mol.charges.copy())


def make_c1_normal(outdir: str, normal: str = "x", **over) -> Workload:
    """The c1 flake written with its normal along x or y (surface_normal)."""
    mol = rotated_molecule(hex_flake(4), normal)
    opts = _base_options(area="0.5 0.5", surface_normal=normal)
    opts.update({k: str(v) for k, v in over.items()})
    return _emit(f"c1-normal-{normal}", outdir, mol, opts, False)


def periodic_test_potential(n_grid: Sequence[int], amp: float = 0.004) -> np.ndarray:
    """A smooth periodic potential in fractional coordinates (a few low
    Fourier modes, Hartree), usable on any cell shape."""
    u, v, w = (np.arange(n) / n for n in n_grid)
    U, V, Wz = np.meshgrid(u, v, w, indexing="ij")
    tp = 2.0 * math.pi
    return amp * (np.cos(tp * U) + 0.7 * np.sin(tp * (V + Wz)) + 0.5 * np.cos(tp * (U - V) + 0.3)
                  + 0.4 * np.cos(tp * 2 * Wz + 1.1))


def make_c4_variant(outdir: str, variant: str, seed: int = 1238, **over) -> Workload:
    """c4-mini variants that reach the remaining branches of the cube and
    external-potential path:

    * ``bohr``     -- the Hartree cube written in Bohr (positive voxel counts);
    * ``skew``     -- a non-orthogonal in-plane cell (basis^-1 branches of the
                      Gaussian placement, the k-vectors and the grid lookup);
    * ``normal_y`` -- slab and cube written with the surface normal along y
                      (the ZXY rotation of positions and potential grid)."""
    rng = np.random.default_rng(seed)
    cell, n_atoms, n_grid, thick = (6.0, 6.0, 16.0), 45, (24, 24, 64), 4.0
    mol = random_slab(n_atoms, cell[:2], thick, rng)
    opts = _base_options(area="0.8 0.8", dx="0.2", dy="0.2", zhigh="14.4", zlow="10.5",
                         vdw_pbc="on", use_external_potential="on", e_potential_file="hartree.cube")
    os.makedirs(outdir, exist_ok=True)
    path = os.path.join(outdir, "hartree.cube")
    if variant == "bohr":
        pot = gaussian_hartree_potential(mol, cell, n_grid)
        write_cube(path, pot, cell, mol, bohr=True)
    elif variant == "skew":
        vox = np.array([[6.0, 0.0, 0.0], [1.5, 5.8, 0.0], [0.0, 0.0, 16.0]]) / np.array(n_grid)[:, None]
        pot = periodic_test_potential(n_grid)
        write_cube(path, pot, cell, mol, voxels=vox)
    elif variant == "normal_y":
        pot = gaussian_hartree_potential(mol, cell, n_grid)
        cols = FILE_COLUMNS["y"]
        # file axis a holds system axis cols[a]: values_f[i, j, k] = pot[k, i, j]
        pot_f = np.transpose(pot, (1, 2, 0)).copy()
        cell_f = [cell[c] for c in cols]
        mol = rotated_molecule(mol, "y")
        write_cube(path, pot_f, cell_f, mol)
        opts["surface_normal"] = "y"
    else:
        raise ValueError(variant)
    opts.update({k: str(v) for k, v in over.items()})
    return _emit(f"c4-{variant}", outdir, mol, opts, False)


MAKERS = {"c1": make_c1, "c2": make_c2, "c3": make_c3, "c4": make_c4, "c5": make_c5}


def scan_shape(opts: Dict[str, str]) -> Tuple[int, int, int]:
    """Pixel grid of a scan: mechafm.cpp:47-49 (floor of double division + 1)."""
    ax, ay = (float(v) for v in opts["area"].split()[:2])
    dx, dy, dz = float(opts["dx"]), float(opts["dy"]), float(opts["dz"])
    zh, zl = float(opts["zhigh"]), float(opts["zlow"])
    return (math.floor(ax / dx) + 1, math.floor(ay / dy) + 1, math.floor((zh - zl) / dz) + 1)


# --------------------------------------------------------------------------
# flexible on
# --------------------------------------------------------------------------
# The flexible-topology lines of the reference's examples/parameters.dat:26-44
# (possible bonds, harmonic constants, 10-4 substrate wall), plus N-H / C-N.
FLEX_PARAM_LINES = (
    "topobond C C 1.430", "topobond C H 1.095", "topobond N H 1.050", "topobond C N 1.470",
    "bond 25.000", "angle 0.2500", "dihedral 0.2500",
    "substrate 0.100 3.0 3.0 7.5 0.01",
)


def terminated_flake(n_rings: int = 2, ch: float = 1.09) -> Molecule:
    """hex_flake(n_rings) with one H on every edge carbon (two C neighbours),
    pointing away from the two neighbours' midpoint."""
    c = hex_flake(n_rings)
    xy = c.pos[:, :2]
    types, pos = list(c.types), [p.copy() for p in c.pos]
    for i in range(len(xy)):
        d = np.linalg.norm(xy - xy[i], axis=1)
        nb = np.where((d > 0.1) & (d < 1.6))[0]
        if len(nb) == 2:
            u = xy[i] - xy[nb].mean(axis=0)
            u /= np.linalg.norm(u)
            types.append("H")
            pos.append(np.array([xy[i][0] + ch * u[0], xy[i][1] + ch * u[1], 0.0]))
    return Molecule(types, np.array(pos))


def flex_molecule(seed: int = 1240) -> Molecule:
    """Coronene C24H12 with a slight seeded buckle, edge H tethered (fixed 2),
    two inner carbons fixed (1), and a pyramidal ammonia 6.9 A from the
    centre at z = 0.6 A, 2.3 A from the nearest flake atom (its first H
    tethered, the rest free), with seeded zero-sum xyz charges."""
    rng = np.random.default_rng(seed)
    a = terminated_flake(2)
    pos = a.pos.copy()
    pos[:, 2] += rng.normal(0.0, 0.02, len(pos))
    types = list(a.types)
    fixed = [2 if t == "H" else 0 for t in types]
    inner = np.argsort(np.linalg.norm(pos[:, :2], axis=1))[:2]
    for i in inner:
        fixed[i] = 1
    n0 = np.array([6.8, 1.2, 0.6])
    types.append("N")
    pos = np.vstack([pos, n0])
    fixed.append(0)
    for k in range(3):
        ang = math.radians(120.0 * k + 15.0)
        h = n0 + np.array([0.95 * math.cos(ang), 0.95 * math.sin(ang), 0.34])
        types.append("H")
        pos = np.vstack([pos, h])
        fixed.append(0 if k else 2)
    q = rng.normal(0.0, 0.08, len(types))
    q -= q.mean()
    return Molecule(types, pos, q, np.array(fixed))


def make_fx(outdir: str, mini: bool = True, seed: int = 1240, **over) -> Workload:
    mol = flex_molecule(seed)
    opts = _base_options(flexible="on", coulomb="on", tip_dummy_coulomb="on", area="0.5 0.5",
                         zhigh="9.0", zlow="8.0", dx="0.25", dy="0.25", center="0.1 -0.05")
    if not mini:
        opts.update(area="8.0 8.0", zlow="6.0", dx="0.2", dy="0.2")
    opts.update({k: str(v) for k, v in over.items()})
    return _emit("fx-mini" if mini else "fx", outdir, mol, opts, True, param_lines=FLEX_PARAM_LINES)


EXAMPLE_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tests", "golden",
                           "example_graphene")


def make_fx_example(outdir: str, example_dir: str = EXAMPLE_DIR, **over) -> Workload:
    """The reference's own example job (examples/input.scan, graphene.xyz,
    parameters.dat: 72 atoms, every one tethered in xy) with flexible on and
    any option overridden; the files are the copies kept with the tests."""
    os.makedirs(outdir, exist_ok=True)
    opts: Dict[str, str] = {}
    for line in open(os.path.join(example_dir, "input.scan")):
        t = line.split()
        if t and not line.startswith("#"):
            opts[t[0]] = " ".join(t[1:])
    opts["flexible"] = "on"
    opts.update({k: str(v) for k, v in over.items()})
    for fn in (opts["xyzfile"], opts["paramfile"]):
        with open(os.path.join(example_dir, fn)) as fi, open(os.path.join(outdir, fn), "w") as fo:
            fo.write(fi.read())
    path = os.path.join(outdir, "input.scan")
    write_input(path, opts)
    types, pos = [], []
    for line in open(os.path.join(outdir, opts["xyzfile"])).read().splitlines()[2:]:
        t = line.split()
        if t:
            types.append(t[0])
            pos.append([float(v) for v in t[1:4]])
    return Workload("fx-example", path, opts, Molecule(types, np.array(pos)))
