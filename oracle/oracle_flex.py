"""TEST INFRASTRUCTURE ONLY -- the oracle for flexible on.

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline
leg may import this module, and only as the checker.  No product path imports
it.

Restates the flexible-surface set-up of MechAFM in Python -- the xyz file's
sixth (fixed) column (parse.cpp:548-563), readFlexibleParameters
(parse.cpp:700-782), buildSurfaceSurfaceInteractions with getConnectedAtoms
(simulation.cpp:533-693) and buildSubstrateInteractions (:517-531) on top of
the rigid set-up of ``oracle.setup`` -- as a term table in the layout of
mafm_flex_term, and drives ``oracle_flex.c`` (liboracle.so) for the
whole-list evaluation and the scan.  Python floats are IEEE doubles and
``math`` calls glibc, so each constant is the reference's to the bit.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass
from typing import List, Optional, Tuple

import numpy as np

from . import oracle as O

FX_LJ, FX_MORSE, FX_COULOMB, FX_TIP_HARMONIC, FX_BOND, FX_ANGLE, FX_DIHEDRAL, FX_SUBSTRATE, FX_XY = range(9)

TERM_DTYPE = np.dtype([("kind", "<i4"), ("atom", "<i4", (4,)), ("reserved", "<i4"), ("p", "<f8", (6,)),
                       ("shift", "<f8", (3,))])
assert TERM_DTYPE.itemsize == 96

_OR2FX = {O.OR_LJ: FX_LJ, O.OR_MORSE: FX_MORSE, O.OR_COUL: FX_COULOMB, O.OR_HARM: FX_TIP_HARMONIC}


class FxOpts(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("dx", ctypes.c_double), ("dy", ctypes.c_double), ("dz", ctypes.c_double),
                ("zhigh", ctypes.c_double), ("tip_dummy_d", ctypes.c_double),
                ("minimiser", ctypes.c_int), ("integrator", ctypes.c_int), ("minterm", ctypes.c_int),
                ("etol", ctypes.c_double), ("ftol", ctypes.c_double), ("dt", ctypes.c_double),
                ("maxsteps", ctypes.c_int), ("tip_pbc", ctypes.c_int),
                ("cell", ctypes.c_double * 9), ("offset", ctypes.c_double * 3)]


@dataclass
class FlexParams:
    bond_k: float = 0.0
    angle_k: float = 0.0
    dihedral_k: float = 0.0
    substrate: Tuple[float, float, float, float, float] = (0.0, 0.0, 0.0, 0.0, 0.0)  # eps sig lambda rc k
    topobonds: Optional[List[Tuple[str, str, float]]] = None


@dataclass
class FlexSetup:
    base: "O.Setup"
    fixed: np.ndarray          # (n_atoms,) int32
    params: FlexParams
    terms: np.ndarray          # TERM_DTYPE, the interaction list in order
    n_tip_surface: int


def read_fixed(o: "O.Options", n_atoms: int) -> np.ndarray:
    """The fixed flags of readXYZFile (parse.cpp:544-563): the sixth column
    when the first atom line has six, else 0; every surface atom 1 when not
    flexible; the dummy fixed, the tip free (system.cpp:23-24)."""
    with open(o.xyzfile) as f:
        lines = f.readlines()
    realxyz = False
    for line in lines:
        if O._is_comment(line):
            continue
        t = line.split()[0]
        realxyz = all(ch.isdigit() or ch in "+-" for ch in t)
        break
    fixed = [1, 0]
    ncols = None
    for n, line in enumerate(lines):
        if realxyz and n < 2:
            continue
        if O._is_comment(line):
            continue
        tok = line.split()
        if ncols is None:
            ncols = len(tok)
        f6 = int(tok[5]) if ncols == 6 else 0
        fixed.append(f6 if o.flexible else 1)
    assert len(fixed) == n_atoms
    return np.array(fixed, dtype=np.int32)


def read_flex_params(path: str) -> FlexParams:
    """readFlexibleParameters (parse.cpp:700-782)."""
    p = FlexParams(topobonds=[])
    seen = set()
    with open(path) as f:
        for line in f:
            if O._is_comment(line):
                continue
            tok = line.split()
            if not tok:
                continue
            if tok[0] == "bond":
                p.bond_k = float(tok[1]); seen.add("b")
            elif tok[0] == "angle":
                p.angle_k = float(tok[1]); seen.add("a")
            elif tok[0] == "dihedral":
                p.dihedral_k = float(tok[1]); seen.add("d")
            elif tok[0] == "substrate":
                p.substrate = tuple(float(v) for v in tok[1:6]); seen.add("s")
            elif tok[0] == "topobond":
                p.topobonds.append((tok[1], tok[2], float(tok[3])))
    for k, what in (("b", "bond"), ("a", "angle"), ("d", "dihedral"), ("s", "substrate support")):
        if k not in seen:
            raise ValueError(f"No harmonic {what} parameters found in parameter file!")
    return p


def _sub(a, b):
    return (a[0] - b[0], a[1] - b[1], a[2] - b[2])


def _len(a):                     # vectors.hpp:103-105
    return math.sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2])


def _dot(a, b):                  # vectors.hpp:116-118
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]


def _cross(a, b):                # vectors.hpp:120-122
    return (a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0])


def _normalized(a):              # vectors.hpp:107-114
    l = _len(a)
    return (a[0] / l, a[1] / l, a[2] / l) if l > 1e-10 else (0.0, 0.0, 0.0)


def _term(kind, atoms, p, shift=(0.0, 0.0, 0.0)):
    t = np.zeros((), dtype=TERM_DTYPE)
    t["kind"] = kind
    t["atom"][: len(atoms)] = atoms
    t["p"][: len(p)] = p
    t["shift"][:] = shift
    return t


def flex_setup(input_path: str) -> FlexSetup:
    """Simulation::initialize with flexible on: the rigid set-up of
    ``oracle.setup`` (tip-surface and tip-dummy terms, centring), then the
    surface-surface and substrate terms."""
    st = O.setup(input_path)
    o, s, prm = st.opts, st.system, st.params
    if not o.flexible:
        raise ValueError("flexible is off")
    n = len(s.types)
    fixed = read_fixed(o, n)
    fp = read_flex_params(o.paramfile)
    terms: List[np.ndarray] = []
    n_ts = 0
    for d in st.inter:
        if d["kind"] == O.OR_GRID:
            raise ValueError("grids cannot be flexible")
        kind = _OR2FX[d["kind"]]
        if d["kind"] != O.OR_HARM and d["i1"] == 1 and len(terms) == n_ts:
            n_ts += 1
        terms.append(_term(kind, (d["i1"], d["i2"]), tuple(d["p"][: 3 if kind == FX_MORSE else 2]),
                           tuple(d.get("shift", (0.0, 0.0, 0.0)))))
    P = [tuple(float(v) for v in row) for row in s.pos]
    types = s.types
    # buildSurfaceSurfaceInteractions (simulation.cpp:557-693)
    adjacent = [set() for _ in range(n)]
    bonds = []
    for i in range(2, n):
        for j in range(i + 1, n):
            r0 = 0.0
            for a, b, r in fp.topobonds:
                if sorted((a, b)) == sorted((types[i], types[j])):
                    r0 = r
            d = _len(_sub(P[i], P[j]))
            if d < 1.1 * r0:
                bonds.append((i, j))
                adjacent[i].add(j)
                adjacent[j].add(i)
                terms.append(_term(FX_BOND, (i, j), (fp.bond_k, d)))
    # getConnectedAtoms (simulation.cpp:533-555): breadth-first from each atom's neighbours
    connected = []
    for i in range(n):
        seen = set()
        todo = [(a, 1) for a in adjacent[i]]
        while todo:
            a, dist = todo.pop(0)
            seen.add(a)
            if dist < n:
                todo.extend((b, dist + 1) for b in adjacent[a] if b not in seen)
        connected.append(seen)
    angles = []
    for i in range(len(bonds)):
        for j in range(i + 1, len(bonds)):
            b1, b2 = bonds[i], bonds[j]
            shared = -1
            if b1[0] == b2[0]:
                shared, a1, a2 = b1[0], b1[1], b2[1]
            elif b1[0] == b2[1]:
                shared, a1, a2 = b1[0], b1[1], b2[0]
            elif b1[1] == b2[0]:
                shared, a1, a2 = b1[1], b1[0], b2[1]
            elif b1[1] == b2[1]:
                shared, a1, a2 = b1[1], b1[0], b2[0]
            if shared != -1:
                angles.append((shared, a1, a2))
                cos_t = _dot(_normalized(_sub(P[a1], P[shared])), _normalized(_sub(P[a2], P[shared])))
                terms.append(_term(FX_ANGLE, (shared, a1, a2), (fp.angle_k, math.acos(cos_t))))
    improper = set()
    for i in range(len(angles)):
        for j in range(i + 1, len(angles)):
            A, B = angles[i], angles[j]
            q = None
            if A[0] == B[0] and A[0] not in improper:
                improper.add(A[0])
                q = (A[0], A[1], A[2], B[2] if (B[1] == A[1] or B[1] == A[2]) else B[1])
            elif A[1] == B[0] and B[1] == A[0]:
                q = (A[2], A[0], B[0], B[2])
            elif A[2] == B[0] and B[1] == A[0]:
                q = (A[1], A[0], B[0], B[2])
            elif A[1] == B[0] and B[2] == A[0]:
                q = (A[2], A[0], B[0], B[1])
            elif A[2] == B[0] and B[2] == A[0]:
                q = (A[1], A[0], B[0], B[1])
            if q is not None:
                d12, d23, d43 = _sub(P[q[0]], P[q[1]]), _sub(P[q[1]], P[q[2]]), _sub(P[q[3]], P[q[2]])
                m, nn = _cross(d12, d23), _cross(d43, d23)
                tan_s = _dot(nn, d12) * _len(d23) / _dot(m, nn)
                terms.append(_term(FX_DIHEDRAL, q, (fp.dihedral_k, math.atan(tan_s))))
    for i in range(2, n):
        for j in range(i + 1, n):
            if j in connected[i]:
                continue
            d = O._vdw(s, prm, i, j)
            kind = _OR2FX[d["kind"]]
            terms.append(_term(kind, (i, j), tuple(d["p"][: 3 if kind == FX_MORSE else 2])))
            if o.coulomb:
                c = O._coul(o, s, prm, i, j)
                if c is not None:
                    terms.append(_term(FX_COULOMB, (i, j), (c["p"][0],)))
    # buildSubstrateInteractions (simulation.cpp:517-531), constants of interactions.hpp:222-229
    eps, sig, lam, rc, k = fp.substrate
    mult = 2 * O.PI * eps * ((sig / lam) * (sig / lam))   # pow(x, 2) is x * x
    ulj = (O.PI / (lam * lam)) * 4 * eps * ((sig / rc) ** 12 - (sig / rc) ** 6)
    ushift = mult * ((12. / 5) * (sig / rc) ** 10 - 3 * (sig / rc) ** 4)
    for i in range(2, n):
        terms.append(_term(FX_SUBSTRATE, (i,), (mult, ulj, ushift, -sig, rc, sig)))
        if fixed[i] == 2:
            terms.append(_term(FX_XY, (i,), (k, P[i][0], P[i][1])))
    return FlexSetup(st, fixed, fp, np.array(terms, dtype=TERM_DTYPE), n_ts)


def _terms_ptr(terms: np.ndarray):
    terms = np.ascontiguousarray(terms, dtype=TERM_DTYPE)
    return terms, terms.ctypes.data_as(ctypes.c_void_p)


def evaluate(terms: np.ndarray, pos: np.ndarray, only_tip_surface: bool = False) -> Tuple[np.ndarray, np.ndarray]:
    """Forces (n, 3) and energies (n) of the whole list at pos (n, 3)."""
    terms, tp = _terms_ptr(terms)
    pos = np.ascontiguousarray(pos, dtype=np.float64)
    n = len(pos)
    F = np.zeros((n, 3))
    E = np.zeros(n)
    O.lib().oracle_flex_eval(n, tp, len(terms), O._dp(pos), O._dp(F), O._dp(E), int(only_tip_surface))
    return F, E


def system_op(terms: np.ndarray, fixed: np.ndarray, op: str, pos: np.ndarray, vel: Optional[np.ndarray],
              forces: np.ndarray, energies: np.ndarray, integrator: int = 1, dt: float = 0.0,
              minimiser: int = 1, minterm: int = 1, maxsteps: int = 5000, etol: float = 0.01,
              ftol: float = 0.01) -> int:
    """One System-level operator on one state, in place (oracle_flex.c:
    oracle_flex_system): "eval" adds the list onto forces / energies, "step"
    is one eulerStep / midpointStep / rk4Step (integrator 0 / 1 / 2),
    "minimise" SDMinimisation (minimiser 0) or FIREMinimisation (1) with the
    given options.  Returns the minimiser's step count ("minimise")."""
    terms, tp = _terms_ptr(terms)
    fixed = np.ascontiguousarray(fixed, dtype=np.int32)
    for a in (pos, forces, energies) + ((vel,) if vel is not None else ()):
        assert a.dtype == np.float64 and a.flags.c_contiguous
    so = FxOpts()
    so.minimiser, so.integrator, so.minterm, so.maxsteps = minimiser, integrator, minterm, maxsteps
    so.etol, so.ftol, so.dt = etol, ftol, dt
    code = {"eval": 0, "step": 1, "minimise": 2}[op]
    return O.lib().oracle_flex_system(len(pos), tp, len(terms), fixed.ctypes.data_as(ctypes.c_void_p), code,
                                      ctypes.byref(so), integrator, ctypes.c_double(dt), O._dp(pos),
                                      O._dp(vel) if vel is not None else None, O._dp(forces), O._dp(energies))


def _fx_opts(fs: FlexSetup, shape) -> FxOpts:
    o, s = fs.base.opts, fs.base.system
    so = FxOpts()
    so.nx, so.ny, so.nz = shape
    so.dx, so.dy, so.dz, so.zhigh = o.dx, o.dy, o.dz, o.zhigh
    so.tip_dummy_d = s.tip_dummy_d
    so.minimiser = 0 if o.minimiser == "SD" else 1
    so.integrator = {"euler": 0, "midpoint": 1, "rk4": 2}[o.integrator]
    so.minterm = {"e": 0, "f": 1, "ef": 2}[o.minterm]
    so.etol, so.ftol, so.dt, so.maxsteps = o.etol, o.ftol, o.dt, o.maxsteps
    so.tip_pbc = int(s.tip_pbc)
    for a, v in enumerate(s.cell.ravel()):
        so.cell[a] = v
    for a in range(3):
        so.offset[a] = s.offset[a]
    return so


def scan(fs: FlexSetup, cols: Optional[Tuple[int, int]] = None, terms: Optional[np.ndarray] = None,
         pos: Optional[np.ndarray] = None, fixed: Optional[np.ndarray] = None):
    """The restated flexible scan over columns [first, first + count) (all by
    default).  Returns (records (count*nz, 16) in ref_harness's layout, the
    centre column's states (nz, n, 3) or None, total steps)."""
    shape = fs.base.scan_shape
    nx, ny, nz = shape
    first, count = cols if cols is not None else (0, nx * ny)
    terms, tp = _terms_ptr(fs.terms if terms is None else terms)
    pos0 = np.ascontiguousarray(fs.base.system.pos if pos is None else pos, dtype=np.float64)
    fixed = np.ascontiguousarray(fs.fixed if fixed is None else fixed, dtype=np.int32)
    n = len(pos0)
    rec = np.zeros((count * nz, 16))
    state_col = (nx * ny) // 2
    state = np.zeros((nz, n, 3))
    steps = ctypes.c_int64(0)
    so = _fx_opts(fs, shape)
    O.lib().oracle_flex_scan(n, tp, len(terms), O._dp(pos0), fixed.ctypes.data_as(ctypes.c_void_p),
                             ctypes.byref(so), first, count, O._dp(rec), state_col, O._dp(state),
                             ctypes.byref(steps))
    has_state = first <= state_col < first + count
    return rec, (state if has_state else None), steps.value
