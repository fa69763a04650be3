"""TEST INFRASTRUCTURE ONLY -- the oracle.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline
leg may import this module, and only as the checker.  No product path imports
it.

A CPU FP64 restatement of MechAFM's rigid scan, in two halves:

* this module restates the set-up (the input.scan / parameter / xyz parsers of
  parse.cpp, ``Simulation::initialize`` and ``buildInteractions`` of
  simulation.cpp, the cube reader of cube_io.cpp and the
  ``ElectrostaticPotentialInteraction`` FFT pipeline of interactions.cpp with
  numpy.fft in place of kiss_fft: both compute the same unnormalised DFT, so
  the two agree to rounding);
* ``oracle_scan.c`` (liboracle.so) restates the hot loops: operators,
  integrators, minimisers, grid lookup and the pixel / z sweep.

Each function cites the reference lines it follows.
"""
from __future__ import annotations

import ctypes
import math
import os
import subprocess
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

SIXTHRT2 = 1.12246204830937298142      # globals.hpp:14
PI = 3.14159265358979323846            # globals.hpp:13
FORCE_GRID_MARGIN = 1.5                # globals.hpp:16
GAUSS_CUTOFF = 1.0e-10                 # globals.hpp:17
TIP_GAUSS_WIDTH = 0.5                  # globals.hpp:18
HARTREE = {"kcal": 627.50961, "kj": 2625.5002, "ev": 27.211386}   # globals.hpp:21-23
BOHR = 0.52917721092                   # cube_io.hpp:17
QBASE = {"kcal": 332.06371, "kj": 1389.354563, "ev": 14.39964901}  # parse.cpp:278-298

OR_LJ, OR_MORSE, OR_COUL, OR_HARM, OR_GRID = range(5)


class OrInter(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("i1", ctypes.c_int), ("i2", ctypes.c_int), ("grid", ctypes.c_int),
                ("p", ctypes.c_double * 3), ("shift", ctypes.c_double * 3)]


class OrGrid(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int * 3), ("periodic", ctypes.c_int), ("orthogonal", ctypes.c_int),
                ("basis", ctypes.c_double * 9), ("offset", ctypes.c_double * 3),
                ("fe", ctypes.POINTER(ctypes.c_double))]


class OrScanOpts(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz", ctypes.c_double),
                ("zhigh", ctypes.c_double), ("tip_dummy_d", ctypes.c_double),
                ("minimiser", ctypes.c_int), ("integrator", ctypes.c_int), ("minterm", ctypes.c_int),
                ("etol", ctypes.c_double), ("ftol", ctypes.c_double), ("dt", ctypes.c_double),
                ("maxsteps", ctypes.c_int), ("tip_pbc", ctypes.c_int),
                ("cell", ctypes.c_double * 9), ("offset", ctypes.c_double * 3)]


_lib = None


def lib():
    """Load (building if needed) the C half of the oracle."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            subprocess.check_call(["make", "-C", HERE, "liboracle.so"], stdout=subprocess.DEVNULL)
        _lib = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        i64p = ctypes.POINTER(ctypes.c_int64)
        _lib.oracle_scan.argtypes = [ctypes.POINTER(OrInter), ctypes.c_int, ctypes.POINTER(OrGrid), dp,
                                     ctypes.POINTER(OrScanOpts), dp, i64p, i64p]
        _lib.oracle_build_grid.argtypes = [ctypes.POINTER(OrInter), ctypes.c_int, ctypes.POINTER(OrGrid), dp,
                                           ctypes.POINTER(OrScanOpts), ctypes.POINTER(ctypes.c_int), dp, dp, dp]
        _lib.oracle_eval_tip_surface.argtypes = [ctypes.POINTER(OrInter), ctypes.c_int, ctypes.POINTER(OrGrid),
                                                 dp, ctypes.c_long, dp, dp]
        _lib.oracle_eval_total.argtypes = [ctypes.POINTER(OrInter), ctypes.c_int, ctypes.POINTER(OrGrid),
                                           dp, ctypes.c_long, dp, dp, dp]
    return _lib


def _dp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


# --------------------------------------------------------------------------
# parsing (parse.cpp)
# --------------------------------------------------------------------------
def _is_comment(line: str) -> bool:
    """checkForComments (parse.cpp:17-27)."""
    return line == "" or line[0] in "#%\n"


@dataclass
class Options:
    inputfolder: str = ""
    xyzfile: str = ""
    paramfile: str = ""
    e_potential_file: str = ""
    tipatom: str = ""
    dummyatom: str = ""
    units: str = "kcal"
    units_given: bool = False
    coulomb: bool = False
    tip_dummy_coulomb: bool = False
    use_external_potential: bool = False
    area: Tuple[float, float] = (10.0, 10.0)
    center: Tuple[float, float] = (-1.0, -1.0)
    dx: float = 0.1
    dy: float = 0.1
    dz: float = 0.1
    zlow: float = 6.0
    zhigh: float = 10.0
    vdw_pbc: bool = False
    cell: List[Tuple[float, float, float]] = field(default_factory=lambda: [(0.0, 0.0, 0.0)] * 3)
    normal: str = "z"
    etol: float = 0.01
    ftol: float = 0.01
    dt: float = 0.001
    maxsteps: int = 5000
    minterm: Optional[str] = None
    rigidgrid: bool = False
    flexible: bool = False
    minimiser: str = "FIRE"
    integrator: str = "midpoint"
    xyz_charges: bool = False


def read_input(path: str) -> Options:
    """readInputFile (parse.cpp:83-405): defaults :96-134, keywords :143-320,
    mandatory keys :322-337, area/2 default centre :342-344, sanity :386-405."""
    o = Options()
    sl = path.rfind("/")
    o.inputfolder = path[: sl + 1] if sl >= 0 else ""
    onoff = {"on": True, "off": False}
    with open(path) as f:
        for line in f:
            if _is_comment(line):
                continue
            tok = line.split()
            if not tok:
                continue
            key = tok[0].lower()
            val = tok[1] if len(tok) > 1 else ""
            if key in ("xyzfile", "paramfile", "e_potential_file"):
                setattr(o, key, o.inputfolder + val)
            elif key in ("tipatom", "dummyatom"):
                setattr(o, key, val)
            elif key in ("area", "center"):
                setattr(o, key, (float(tok[1]), float(tok[2])))
            elif key in ("zhigh", "zlow", "dx", "dy", "dz", "etol", "ftol", "dt"):
                setattr(o, key, float(val))
            elif key in ("maxsteps",):
                o.maxsteps = int(val)
            elif key in ("bufsize",):
                pass
            elif key in ("cell_a", "cell_b", "cell_c"):
                o.cell["abc".index(key[-1])] = (float(tok[1]), float(tok[2]), float(tok[3]))
            elif key in ("gzip", "statistics", "flexible", "rigidgrid", "coulomb", "tip_dummy_coulomb",
                         "use_external_potential", "vdw_pbc"):
                if val not in onoff:
                    raise ValueError(f"Option {key} must be either on or off!")
                if key not in ("gzip", "statistics"):
                    setattr(o, key, onoff[val])
            elif key == "surface_normal":
                o.normal = val.lower()
            elif key == "minterm":
                if val not in ("e", "f", "ef"):
                    raise ValueError("minterm must be e, f or ef")
                o.minterm = val
            elif key == "units":
                m = {"kcal/mol": "kcal", "kcal": "kcal", "kJ/mol": "kj", "kJ": "kj", "eV": "ev"}
                if val not in m:
                    raise ValueError("unknown units")
                o.units = m[val]
                o.units_given = True
            elif key == "minimiser":
                if val not in ("SD", "FIRE"):
                    raise ValueError("Unrecognised minimiser type!")
                o.minimiser = val
            elif key == "integrator":
                if val not in ("euler", "midpoint", "rk4"):
                    raise ValueError("Unrecognised integrator type!")
                o.integrator = val
            else:
                raise ValueError(f"Unknown option {key}!")
    for k in ("xyzfile", "paramfile", "tipatom", "dummyatom"):
        if not getattr(o, k):
            raise ValueError(f"missing {k}")
    if o.minterm is None:
        raise ValueError("missing minterm")
    if o.center == (-1.0, -1.0):
        o.center = (o.area[0] / 2, o.area[1] / 2)
    if o.rigidgrid and o.flexible:
        raise ValueError("Cannot use a flexible molecule with a static force grid!")
    if o.coulomb and o.use_external_potential:
        raise ValueError("Coulomb and external potential are exclusive")
    if o.vdw_pbc and o.coulomb:
        raise ValueError("vdw_pbc does not support Coulomb")
    if o.vdw_pbc and not o.use_external_potential and any(c == (0.0, 0.0, 0.0) for c in o.cell):
        raise ValueError("The unit cell vectors must be given if periodic vdW is used.")
    return o


@dataclass
class System:
    types: List[str]
    pos: np.ndarray          # (n_atoms + 2, 3); rows 0/1 dummy/tip
    charges: np.ndarray      # (n_atoms + 2,)
    offset: np.ndarray = field(default_factory=lambda: np.zeros(3))
    tip_pbc: bool = False
    cell: np.ndarray = field(default_factory=lambda: np.zeros((3, 3)))
    tip_dummy_d: float = 0.0


def read_xyz(o: Options) -> System:
    """readXYZFile (parse.cpp:477-585): 4/5/6 columns, rigid => all fixed,
    xyz charges used iff sum q^2 >= 1e-10 (:571-583)."""
    with open(o.xyzfile) as f:
        lines = f.readlines()
    n_atoms, realxyz = 0, False
    first = True
    for line in lines:
        if _is_comment(line):
            continue
        if first:
            t = line.split()[0]
            if all(ch.isdigit() or ch in "+-" for ch in t):
                n_atoms, realxyz = int(t), True
                break
            first = False
        n_atoms += 1
    types = [o.dummyatom, o.tipatom]
    pos = [(0.0, 0.0, 0.0), (0.0, 0.0, 0.0)]
    q = [0.0, 0.0]
    ncols = None
    for n, line in enumerate(lines):
        if realxyz and n < 2:
            continue
        if _is_comment(line):
            continue
        tok = line.split()
        if ncols is None:
            ncols = len(tok)
        if ncols not in (4, 5, 6):
            raise ValueError("Invalid number of columns in the xyz file.")
        types.append(tok[0])
        pos.append((float(tok[1]), float(tok[2]), float(tok[3])))
        q.append(float(tok[4]) if ncols >= 5 else 0.0)
    if len(types) - 2 != n_atoms:
        raise ValueError("atom count mismatch")
    qa = np.array(q)
    s = 0.0
    for v in qa:
        s += v ** 2
    o.xyz_charges = not (abs(s) < 1e-10)
    return System(types, np.array(pos, dtype=np.float64), qa)


@dataclass
class Params:
    atoms: Dict[str, Tuple[float, float, float, float]] = field(default_factory=dict)
    harm: Optional[Tuple[float, float]] = None
    overwrites: List[dict] = field(default_factory=list)


def read_params(o: Options, s: System) -> Params:
    """readParameterFile (parse.cpp:588-697): atom / harm / pair_ovwrt; the
    dummy and tip charges come from their atom lines (:626-631)."""
    p = Params()
    with open(o.paramfile) as f:
        for line in f:
            if _is_comment(line):
                continue
            tok = line.split()
            if not tok:
                continue
            if tok[0] == "atom":
                name = tok[1]
                eps, sig, mass, q = (float(v) for v in tok[2:6])
                p.atoms[name] = (eps, sig, q, mass)
                if o.dummyatom == name:
                    s.charges[0] = q
                if o.tipatom == name:
                    s.charges[1] = q
            elif tok[0] == "harm":
                if p.harm is not None:
                    raise ValueError("harm specified twice")
                p.harm = (float(tok[1]), float(tok[2]))
            elif tok[0] == "pair_ovwrt":
                style = tok[3]
                if style == "lj":
                    if len(tok) != 6:
                        raise ValueError("LJ overwrite needs eps, sig")
                    p.overwrites.append({"atoms": sorted([tok[1], tok[2]]), "morse": False,
                                         "eps": float(tok[4]), "sig": float(tok[5])})
                elif style == "morse":
                    if len(tok) != 7:
                        raise ValueError("Morse overwrite needs De, a, re")
                    p.overwrites.append({"atoms": sorted([tok[1], tok[2]]), "morse": True,
                                         "de": float(tok[4]), "a": float(tok[5]), "re": float(tok[6])})
                else:
                    raise ValueError(f"Unknown pair style overwrite '{style}'")
    if p.harm is None:
        raise ValueError("No harmonic spring parameters found in parameter file!")
    return p


# --------------------------------------------------------------------------
# cube file + external potential (cube_io.cpp, interactions.cpp:53-234)
# --------------------------------------------------------------------------
@dataclass
class Cube:
    n: Tuple[int, int, int]
    vox: np.ndarray          # (3, 3): row i = voxel vector i, in A
    origin: np.ndarray
    values: np.ndarray       # (nx, ny, nz)


def read_cube(path: str, values: bool = True) -> Cube:
    """CubeReader::readMetadata / readVolumetricData (cube_io.cpp:40-54, 89-139):
    negative voxel count = Angstrom, else Bohr (x 0.52917721092)."""
    with open(path) as f:
        f.readline(); f.readline()
        rest = f.read().split()
    it = iter(rest)
    n_atoms = int(next(it))
    origin = np.array([float(next(it)) for _ in range(3)])
    n = []
    vox = np.zeros((3, 3))
    for a in range(3):
        n.append(int(next(it)))
        vox[a] = [float(next(it)) for _ in range(3)]
    if n[0] < 0:
        n = [-v for v in n]
    else:
        vox = vox * BOHR
        origin = origin * BOHR
    for _ in range(n_atoms):
        for _ in range(5):
            next(it)
    vals = None
    if values:
        data = np.array([float(v) for v in it][: n[0] * n[1] * n[2]])
        vals = data.reshape(n)
    return Cube(tuple(n), vox, origin, vals)


def _det3(m: np.ndarray) -> float:
    """Matrix3::determinant (matrices.hpp:135-143)."""
    m = [float(v) for v in m.ravel()]
    det = m[0] * (m[4] * m[8] - m[5] * m[7])
    det += m[1] * (m[5] * m[6] - m[3] * m[8])
    det += m[2] * (m[3] * m[7] - m[4] * m[6])
    return det


def _inv3(mm: np.ndarray) -> np.ndarray:
    """Matrix3::inverse (matrices.hpp:89-114)."""
    inv_det = 1 / _det3(mm)
    a, b, c, d, e, f, g, h, i = (float(v) for v in mm.ravel())
    return np.array([[inv_det * (e * i - f * h), -inv_det * (b * i - c * h), inv_det * (b * f - c * e)],
                     [-inv_det * (d * i - f * g), inv_det * (a * i - c * g), -inv_det * (a * f - c * d)],
                     [inv_det * (d * h - e * g), -inv_det * (a * h - b * g), inv_det * (a * e - b * d)]])


def _is_diag(m: np.ndarray) -> bool:
    """Matrix3::isDiagonal (matrices.hpp:72-86)."""
    diag = sum(abs(m[i, i]) for i in range(3))
    nondiag = sum(abs(m[i, j]) for i in range(3) for j in range(3) if i != j)
    return nondiag < 1.0e-10 * diag


def electrostatic_force_grid(pot: np.ndarray, basis: np.ndarray, origin: np.ndarray, tip_charge: float,
                             width: float = TIP_GAUSS_WIDTH) -> dict:
    """ElectrostaticPotentialInteraction::ElectrostaticPotentialInteraction
    (interactions.cpp:53-234) with ``basis`` columns = voxel vectors.

    Returns a periodic grid dict {n, basis, offset, fe (N,4)}."""
    nx, ny, nz = pot.shape
    ntot = nx * ny * nz
    w2 = width ** 2
    norm = 1.0 / (width * math.sqrt(2.0 * PI)) ** 3
    n_min = min(nx, ny, nz)
    cut = []
    for a in range(3):                                   # :71-95
        col = basis[:, a]
        c = n_min
        for i in range(n_min - 1):
            v = (i * col[0], i * col[1], i * col[2])
            r2 = (v[0] ** 2 + v[1] ** 2) + v[2] ** 2
            if math.exp(-0.5 * r2 / w2) < GAUSS_CUTOFF:
                c = i
                break
        cut.append(c)
    rho = np.zeros(ntot)
    ortho = _is_diag(basis)
    ix = np.arange(-cut[0], cut[0] + 1)
    iy = np.arange(-cut[1], cut[1] + 1)
    iz = np.arange(-cut[2], cut[2])                      # :99  (iz < z_cutoff)
    IX, IY, IZ = np.meshgrid(ix, iy, iz, indexing="ij")
    IX, IY, IZ = IX.ravel(), IY.ravel(), IZ.ravel()
    if ortho:                                            # DataGrid::positionAt (data_grid.cpp:131-142)
        px, py, pz = 0.0 + IX * basis[0, 0], 0.0 + IY * basis[1, 1], 0.0 + IZ * basis[2, 2]
    else:
        px = 0.0 + (basis[0, 0] * IX + basis[0, 1] * IY + basis[0, 2] * IZ)
        py = 0.0 + (basis[1, 0] * IX + basis[1, 1] * IY + basis[1, 2] * IZ)
        pz = 0.0 + (basis[2, 0] * IX + basis[2, 1] * IY + basis[2, 2] * IZ)
    r2 = (px ** 2 + py ** 2) + pz ** 2
    val = tip_charge * norm * np.exp(-0.5 * r2 / w2)
    wx, wy, wz = IX % nx, IY % ny, IZ % nz               # DataGrid::atPBC (data_grid.cpp:105-127)
    np.add.at(rho, (wx * ny + wy) * nz + wz, val)
    rho = rho.reshape(nx, ny, nz)

    pot_k = np.fft.fftn(pot)                             # :126-127
    rho_k = np.fft.fftn(rho)
    vol = _det3(basis)                                   # :139-146
    e_k = pot_k * (vol * rho_k)

    def ifft_real(a):                                    # fft.cpp:51-94
        out = np.fft.ifftn(a)
        if np.any(out.imag > 1.0e-8):
            raise RuntimeError("Imaginary part of inverse FFT exceeded threshold 1.0e-8!")
        return out.real

    energy = ifft_real(e_k)                              # :162
    nmat = np.diag([float(nx), float(ny), float(nz)])
    kb = _inv3(basis @ nmat).T                           # fft.cpp:41-47
    kvec = []
    for a, n in enumerate((nx, ny, nz)):                 # :165-183
        col = kb[:, a]
        idx = np.arange(n, dtype=np.float64)
        k = idx[:, None] * col[None, :]
        hi = np.arange(n) >= n // 2
        k[hi] = k[hi] - n * col[None, :]
        kvec.append(k)
    fe = np.zeros((ntot, 4))
    fe[:, 3] = energy.ravel()
    for c in range(3):                                   # :185-225
        t = np.zeros_like(e_k)
        for a in range(3):
            kk = kvec[a][:, c]
            b = (-2.0 * PI) * kk
            shape = [1, 1, 1]
            shape[a] = -1
            b = b.reshape(shape)
            t = t + (-(b * e_k.imag) + 1j * (b * e_k.real))
        fe[:, c] = ifft_real(t).ravel()
    return {"n": (nx, ny, nz), "basis": basis.copy(), "offset": np.array(origin, dtype=np.float64),
            "periodic": True, "orthogonal": ortho, "fe": fe}


# --------------------------------------------------------------------------
# Simulation::initialize + buildInteractions (simulation.cpp:31-65, 201-515)
# --------------------------------------------------------------------------
@dataclass
class Setup:
    opts: Options
    system: System
    params: Params
    inter: List[dict]
    grids: List[dict]
    scan_shape: Tuple[int, int, int]


def _mix(a, b):
    """mixeps / mixsig (interactions.hpp:58-64)."""
    return math.sqrt(a[0] * b[0]), (a[1] + b[1]) / 2


def _vdw(s: System, p: Params, i1: int, i2: int, shift=(0.0, 0.0, 0.0)) -> dict:
    """addVDWInteraction (simulation.cpp:238-267)."""
    key = sorted([s.types[i1], s.types[i2]])
    for ow in p.overwrites:
        if ow["atoms"] == key:
            if ow["morse"]:
                return {"kind": OR_MORSE, "i1": i1, "i2": i2, "p": (ow["de"], ow["a"], ow["re"]), "shift": shift}
            es6 = 4 * ow["eps"] * ow["sig"] ** 6
            es12 = 4 * ow["eps"] * ow["sig"] ** 12
            return {"kind": OR_LJ, "i1": i1, "i2": i2, "p": (es6, es12, 0.0), "shift": shift}
    for t in (s.types[i1], s.types[i2]):
        if t not in p.atoms:
            raise ValueError(f"Parameters for atom type {t} not found in parameter file!")
    eps, sig = _mix(p.atoms[s.types[i1]], p.atoms[s.types[i2]])
    es6 = 4 * eps * sig ** 6
    es12 = 4 * eps * sig ** 12
    return {"kind": OR_LJ, "i1": i1, "i2": i2, "p": (es6, es12, 0.0), "shift": shift}


def _coul(o: Options, s: System, p: Params, i1: int, i2: int) -> Optional[dict]:
    """addCoulombInteraction (simulation.cpp:269-295); qbase needs `units`."""
    if o.xyz_charges:
        q1, q2 = s.charges[i1], s.charges[i2]
    else:
        q1, q2 = p.atoms[s.types[i1]][2], p.atoms[s.types[i2]][2]
    if q1 != 0 and q2 != 0:
        if not o.units_given:
            raise ValueError("qbase is undefined without a `units` line")
        return {"kind": OR_COUL, "i1": i1, "i2": i2, "p": (QBASE[o.units] * q1 * q2, 0.0, 0.0),
                "shift": (0.0, 0.0, 0.0)}
    return None


def setup(input_path: str) -> Setup:
    o = read_input(input_path)
    s = read_xyz(o)
    p = read_params(o, s)
    nx = math.floor(o.area[0] / o.dx) + 1                # mechafm.cpp:47-49
    ny = math.floor(o.area[1] / o.dy) + 1
    nz = math.floor((o.zhigh - o.zlow) / o.dz) + 1
    # Simulation::initialize (simulation.cpp:31-65)
    if o.normal == "x":
        s.pos[:, :] = s.pos[:, [1, 2, 0]]                # System::rotateCoordAxes("YZX")
    elif o.normal == "y":
        s.pos[:, :] = s.pos[:, [2, 0, 1]]                # "ZXY"
    cube = None
    if o.vdw_pbc:
        if o.use_external_potential:
            cube = read_cube(o.e_potential_file, values=False)
            cellv = [cube.n[a] * cube.vox[a] for a in range(3)]
        else:
            cellv = [np.array(c, dtype=np.float64) for c in o.cell]
        s.cell = np.array(cellv).T                       # System::setUnitCell (system.cpp:170-177)
        s.tip_pbc = True
    # centerMolecule (system.cpp:119-136)
    avgx = 0.0
    avgy = 0.0
    for i in range(2, len(s.types)):
        avgx += s.pos[i, 0]
        avgy += s.pos[i, 1]
    avgx /= len(s.types) - 2
    avgy /= len(s.types) - 2
    ddx, ddy = o.center[0] - avgx, o.center[1] - avgy
    s.pos[2:, 0] += ddx
    s.pos[2:, 1] += ddy
    s.offset[0], s.offset[1] = ddx, ddy
    # setMoleculeZ (system.cpp:103-116)
    min_z = 10e10
    for i in range(2, len(s.types)):
        if s.pos[i, 2] < min_z and s.types[i] != "H":
            min_z = s.pos[i, 2]
    s.pos[2:, 2] -= min_z
    s.offset[2] = -min_z
    # calculateTipDummyDistance (simulation.cpp:218-225)
    s.tip_dummy_d = (p.atoms[s.types[0]][1] + p.atoms[s.types[1]][1]) / 2 * SIXTHRT2

    inter: List[dict] = []
    grids: List[dict] = []

    def tip_surface() -> List[dict]:                      # simulation.cpp:297-418
        lst = []
        if o.vdw_pbc:
            for a in (-1, 0, 1):
                for b in (-1, 0, 1):
                    c0, c1 = s.cell[:, 0], s.cell[:, 1]
                    shift = tuple(float(v) for v in (a * c0 + b * c1))
                    for i in range(2, len(s.types)):
                        lst.append(_vdw(s, p, 1, i, shift))
        else:
            for i in range(2, len(s.types)):
                lst.append(_vdw(s, p, 1, i))
                if o.coulomb:
                    c = _coul(o, s, p, 1, i)
                    if c is not None:
                        lst.append(c)
        if o.use_external_potential:
            cb = read_cube(o.e_potential_file)
            pot = cb.values * (-HARTREE[o.units])          # :348-355
            vox, corner = cb.vox, cb.origin
            if o.normal == "x":                            # :358-359: "ZYX" throws (data_grid.cpp:93-95)
                raise ValueError("rotateCoordAxes(ZYX) is not implemented for data grids")
            if o.normal == "y":                            # DataGrid::rotateCoordAxes("ZXY") (data_grid.cpp:49-70)
                pot = np.transpose(pot, (2, 0, 1)).copy()
                vox = vox[[2, 0, 1]][:, [2, 0, 1]]
                corner = corner[[2, 0, 1]]
            origin = corner + s.offset                     # :365
            g = electrostatic_force_grid(pot, vox.T.copy(), origin, float(s.charges[1]))
            grids.append(g)
            lst.append({"kind": OR_GRID, "grid": len(grids) - 1})
        return lst

    if o.rigidgrid:                                       # simulation.cpp:420-501
        ts = tip_surface()
        sp = np.array([o.dx, o.dy, o.dz]) / 2
        border = [math.ceil(FORCE_GRID_MARGIN / v) for v in sp]
        n = [math.floor(o.area[0] / sp[0]) + 2 * border[0] + 1,
             math.floor(o.area[1] / sp[1]) + 2 * border[1] + 1,
             math.floor((o.zhigh - o.zlow) / sp[2]) + 2 * border[2] + 1]
        off = [-border[0] * sp[0], -border[1] * sp[1], o.zlow - s.tip_dummy_d - border[2] * sp[2]]
        fe = build_grid(ts, grids, s, o, n, sp, off)
        grids.append({"n": tuple(n), "basis": np.diag(sp), "offset": np.array(off), "periodic": False,
                      "orthogonal": True, "fe": fe})
        inter = [{"kind": OR_GRID, "grid": len(grids) - 1}]
    else:
        inter = tip_surface()
    # buildTipDummyInteractions (simulation.cpp:503-515)
    inter.append(_vdw(s, p, 0, 1))
    if o.tip_dummy_coulomb:
        c = _coul(o, s, p, 0, 1)
        if c is not None:
            inter.append(c)
    inter.append({"kind": OR_HARM, "i1": 0, "i2": 1, "p": (p.harm[0], p.harm[1], 0.0), "shift": (0, 0, 0)})
    return Setup(o, s, p, inter, grids, (nx, ny, nz))


# --------------------------------------------------------------------------
# ctypes marshalling
# --------------------------------------------------------------------------
def _pack(inter: List[dict], grids: List[dict]):
    arr = (OrInter * max(1, len(inter)))()
    for k, d in enumerate(inter):
        arr[k].kind = d["kind"]
        arr[k].i1 = d.get("i1", 1)
        arr[k].i2 = d.get("i2", 1)
        arr[k].grid = d.get("grid", 0)
        for a, v in enumerate(d.get("p", (0, 0, 0))):
            arr[k].p[a] = v
        for a, v in enumerate(d.get("shift", (0, 0, 0))):
            arr[k].shift[a] = v
    garr = (OrGrid * max(1, len(grids)))()
    keep = []
    for k, g in enumerate(grids):
        fe = np.ascontiguousarray(g["fe"], dtype=np.float64)
        keep.append(fe)
        for a in range(3):
            garr[k].n[a] = g["n"][a]
            garr[k].offset[a] = g["offset"][a]
        garr[k].periodic = int(g["periodic"])
        garr[k].orthogonal = int(g["orthogonal"])
        for a, v in enumerate(np.asarray(g["basis"], dtype=np.float64).ravel()):
            garr[k].basis[a] = v
        garr[k].fe = _dp(fe)
    return arr, garr, keep


def _opts(st_o: Options, s: System, shape, **over) -> OrScanOpts:
    so = OrScanOpts()
    so.nx, so.ny, so.nz = shape
    so.dx, so.dy, so.dz, so.zhigh = st_o.dx, st_o.dy, st_o.dz, st_o.zhigh
    so.tip_dummy_d = s.tip_dummy_d
    so.minimiser = 0 if st_o.minimiser == "SD" else 1
    so.integrator = {"euler": 0, "midpoint": 1, "rk4": 2}[st_o.integrator]
    so.minterm = {"e": 0, "f": 1, "ef": 2}[st_o.minterm]
    so.etol, so.ftol, so.dt, so.maxsteps = st_o.etol, st_o.ftol, st_o.dt, st_o.maxsteps
    so.tip_pbc = int(s.tip_pbc)
    for a, v in enumerate(s.cell.ravel()):
        so.cell[a] = v
    for a in range(3):
        so.offset[a] = s.offset[a]
    for k, v in over.items():
        setattr(so, k, v)
    return so


def build_grid(ts, grids, s: System, o: Options, n, spacing, offset) -> np.ndarray:
    arr, garr, keep = _pack(ts, grids)
    atoms = np.ascontiguousarray(s.pos, dtype=np.float64)
    so = _opts(o, s, (1, 1, 1))
    nn = (ctypes.c_int * 3)(*n)
    sp = np.array(spacing, dtype=np.float64)
    off = np.array(offset, dtype=np.float64)
    fe = np.zeros((n[0] * n[1] * n[2], 4))
    rc = lib().oracle_build_grid(arr, len(ts), garr, _dp(atoms), ctypes.byref(so), nn, _dp(sp), _dp(off), _dp(fe))
    if rc != 0:
        raise RuntimeError("grid lookup left the grid during grid build")
    return fe


RECORD_FIELDS = ("i", "j", "k", "n", "x", "y", "z", "fx", "fy", "fz", "rx", "ry", "rz", "r", "angle", "e")


def scan(st: Setup, shape=None, **over) -> Tuple[np.ndarray, dict]:
    """Run the restated scan; returns (records (nx*ny*nz, 16), stats)."""
    shape = shape or st.scan_shape
    arr, garr, keep = _pack(st.inter, st.grids)
    atoms = np.ascontiguousarray(st.system.pos, dtype=np.float64)
    so = _opts(st.opts, st.system, shape, **over)
    rec = np.zeros((shape[0] * shape[1] * shape[2], 16))
    warn = ctypes.c_int64(0)
    steps = ctypes.c_int64(0)
    rc = lib().oracle_scan(arr, len(st.inter), garr, _dp(atoms), ctypes.byref(so), _dp(rec),
                           ctypes.byref(warn), ctypes.byref(steps))
    return rec, {"rc": rc, "warnings": warn.value, "total_steps": steps.value}


def eval_tip_surface(st: Setup, pos: np.ndarray) -> np.ndarray:
    arr, garr, keep = _pack(st.inter, st.grids)
    atoms = np.ascontiguousarray(st.system.pos, dtype=np.float64)
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    out = np.zeros((len(pos), 4))
    lib().oracle_eval_tip_surface(arr, len(st.inter), garr, _dp(atoms), len(pos), _dp(pos), _dp(out))
    return out


def eval_total(st: Setup, dummy: np.ndarray, tip: np.ndarray) -> np.ndarray:
    arr, garr, keep = _pack(st.inter, st.grids)
    atoms = np.ascontiguousarray(st.system.pos, dtype=np.float64)
    dummy = np.ascontiguousarray(dummy, dtype=np.float64)
    tip = np.ascontiguousarray(tip, dtype=np.float64)
    out = np.zeros((len(tip), 4))
    lib().oracle_eval_total(arr, len(st.inter), garr, _dp(atoms), len(tip), _dp(dummy), _dp(tip), _dp(out))
    return out


def load_harness_records(path: str) -> np.ndarray:
    """Read ref_harness scan output: 4 int32 header + n x 16 doubles."""
    with open(path, "rb") as f:
        raw = f.read()
    hdr = np.frombuffer(raw[:16], dtype=np.int32)
    return np.frombuffer(raw[16:], dtype=np.float64).reshape(-1, int(hdr[3])).copy()
