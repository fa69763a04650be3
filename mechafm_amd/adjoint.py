"""Differentiable AFM image: torch.autograd.Function over libmafm's scan and
adjoint kernels.

``AFMImage.apply(es6, es12, qq, pos, model)`` returns the Fz image (one value
per scan position, tip-surface z force at the relaxed tip, as the reference
reports it, system.cpp:28-51) of a rigid system whose tip-surface sites carry
the given LJ constants (es6 = 4 eps sigma^6, es12 = 4 eps sigma^12,
simulation.cpp:238-267), Coulomb prefactors (qq = qbase q_tip q_atom,
simulation.cpp:269-295) and positions (scan frame).  Its backward runs
``mafm_adjoint``: an implicit-function-theorem solve at every relaxed tip.

``ScanModel`` reads a MechAFM job, exposes per-type epsilon / sigma,
per-atom charges and positions as leaf tensors and builds the site constants
from them with torch ops (Lorentz-Berthelot mixing, interactions.hpp:58-64),
so autograd chains the site gradients to the physical parameters.

Scans and adjoints run on the GPU through the C ABI; torch holds the small
parameter vectors on the CPU.
"""
from __future__ import annotations

import ctypes
import math
from typing import Dict, List, Optional

import numpy as np
import torch

from . import abi


def _read_job(input_path: str):
    """Atom types / xyz charges, parameter-file atoms and the tip type of a
    MechAFM job (the subset the harness needs; the library's own parser is
    authoritative and the model checks its result against the lowered sites)."""
    folder = input_path[: input_path.rfind("/") + 1]
    opts = {}
    for line in open(input_path):
        t = line.split()
        if t and line[0] not in "#%":
            opts[t[0].lower()] = t[1:]
    lines = [l for l in open(folder + opts["xyzfile"][0]) if l.strip() and l[0] not in "#%"]
    if lines[0].split()[0].lstrip("+-").isdigit():
        raw = open(folder + opts["xyzfile"][0]).read().splitlines()[2:]
        lines = [l for l in raw if l.strip() and l[0] not in "#%"]
    types = [l.split()[0] for l in lines]
    q = np.array([float(l.split()[4]) if len(l.split()) >= 5 else 0.0 for l in lines])
    atoms = {}
    for line in open(folder + opts["paramfile"][0]):
        t = line.split()
        if t and t[0] == "atom":
            atoms[t[1]] = (float(t[2]), float(t[3]), float(t[5]))
    units = opts.get("units", ["kcal/mol"])[0]
    qbase = {"kcal/mol": 332.06371, "kcal": 332.06371, "kJ/mol": 1389.354563, "kJ": 1389.354563,
             "eV": 14.39964901}[units]
    return types, q, atoms, opts["tipatom"][0], qbase


def _inv3(m):
    """Inverse of a row-major 3x3 by cofactors, operation for operation as
    the library computes the cell inverse (Matrix3::inverse,
    matrices.hpp:89-115)."""
    det = m[0] * (m[4] * m[8] - m[5] * m[7])
    det += m[1] * (m[5] * m[6] - m[3] * m[8])
    det += m[2] * (m[3] * m[7] - m[4] * m[6])
    inv_det = 1 / det
    a, b, c, d, e, f, g, h, i = m
    return [inv_det * (e * i - f * h), -inv_det * (b * i - c * h), inv_det * (b * f - c * e),
            -inv_det * (d * i - f * g), inv_det * (a * i - c * g), -inv_det * (a * f - c * d),
            inv_det * (d * h - e * g), -inv_det * (a * h - b * g), inv_det * (a * e - b * d)]


def wrap_dummy_xy(cell, offset, x: np.ndarray, y: np.ndarray):
    """System::setDummyXY's wrap of pixels (x, y) into the unit cell
    (system.cpp:139-162): the dummy position the scan actually used when the
    job has a cell.  Every product and sum is rounded on its own, in the
    order the scan kernels use, so pixels on a cell face wrap the same way."""
    m = [float(v) for v in cell]
    inv = _inv3(m)
    o = [float(v) for v in offset]
    tx, ty, tz = x - o[0], y - o[1], 0.0 - o[2]
    qx = inv[0] * tx + inv[1] * ty + inv[2] * tz
    qy = inv[3] * tx + inv[4] * ty + inv[5] * tz
    qz = inv[6] * tx + inv[7] * ty + inv[8] * tz
    qx = qx - np.floor(qx)
    qy = qy - np.floor(qy)
    ox = m[0] * qx + m[1] * qy + m[2] * qz
    oy = m[3] * qx + m[4] * qy + m[5] * qz
    return ox + o[0], oy + o[1]


class ScanModel:
    """One rigid LJ (+ Coulomb) job whose image is differentiable."""

    def __init__(self, input_path: str, ctx: abi.Context, precision: str = "fast64",
                 adjoint_precision: Optional[str] = None):
        self.ctx = ctx
        self.sim = abi.Simulation(input_path, ctx)
        self.precision = precision
        self.adjoint_precision = adjoint_precision or precision
        self.desc = self.sim.system()
        sites = self.sim.sites()
        if any(s.vdw != abi.VDW_LJ for s in sites) or self.desc.force_grid >= 0 or self.desc.potential_grid >= 0:
            raise ValueError("the differentiable image covers LJ (+ Coulomb) sites without grids")
        types, q, atoms, tip, qbase = _read_job(input_path)
        if len(types) != len(sites):
            raise ValueError("site table does not map one-to-one onto atoms (pbc images?)")
        self.types = types
        self.type_names = sorted(set(types))
        self.type_index = torch.tensor([self.type_names.index(t) for t in types])
        self.eps = torch.tensor([atoms[t][0] for t in self.type_names], dtype=torch.float64)
        self.sig = torch.tensor([atoms[t][1] for t in self.type_names], dtype=torch.float64)
        xyz_charges = float((q ** 2).sum()) >= 1e-10
        self.q = torch.tensor(q if xyz_charges else [atoms[t][2] for t in types], dtype=torch.float64)
        self.pos = torch.tensor([[s.x, s.y, s.z] for s in sites], dtype=torch.float64)
        self.eps_tip, self.sig_tip, self.q_tip = atoms[tip]
        self.qfac = qbase * self.q_tip
        self.has_coulomb = np.array([bool(s.has_coulomb) for s in sites])
        es6, es12, qq = self.site_constants(self.eps, self.sig, self.q)
        ref = np.array([[s.p0, s.p1, s.qq if s.has_coulomb else 0.0] for s in sites])
        got = np.stack([es6.numpy(), es12.numpy(), qq.numpy()], 1)
        if not np.allclose(got, ref, rtol=1e-12, atol=0):
            raise ValueError("harness parameters do not reproduce the library's lowered sites")
        self.scan_desc = self.sim.scan_desc(precision)
        self.shape = (self.scan_desc.nx, self.scan_desc.ny, self.scan_desc.nz)
        self.last_records = None

    def site_constants(self, eps: torch.Tensor, sig: torch.Tensor, q: torch.Tensor):
        e = torch.sqrt(self.eps_tip * eps[self.type_index])
        s = (self.sig_tip + sig[self.type_index]) / 2
        es6 = 4 * e * s ** 6
        es12 = 4 * e * s ** 12
        mask = torch.tensor(self.has_coulomb, dtype=torch.float64)
        qq = self.qfac * q * mask
        return es6, es12, qq

    def image(self, eps=None, sig=None, q=None, pos=None) -> torch.Tensor:
        eps = self.eps if eps is None else eps
        sig = self.sig if sig is None else sig
        q = self.q if q is None else q
        pos = self.pos if pos is None else pos
        es6, es12, qq = self.site_constants(eps, sig, q)
        return AFMImage.apply(es6, es12, qq, pos, self)

    # -- device calls ------------------------------------------------------
    def _system(self, es6, es12, qq, pos) -> abi.SystemDesc:
        n = len(es6)
        arr = (abi.Site * n)()
        for i in range(n):
            arr[i].x, arr[i].y, arr[i].z = (float(v) for v in pos[i])
            arr[i].vdw = abi.VDW_LJ
            arr[i].has_coulomb = int(self.has_coulomb[i])
            arr[i].p0, arr[i].p1, arr[i].p2 = float(es6[i]), float(es12[i]), 0.0
            arr[i].qq = float(qq[i])
        d = abi.SystemDesc()
        ctypes.pointer(d)[0] = self.desc
        d.n_sites = n
        d.sites = ctypes.cast(arr, ctypes.POINTER(abi.Site))
        self._keep = arr
        return d

    def forward_np(self, es6, es12, qq, pos):
        self.ctx.set_systems([self._system(es6, es12, qq, pos)])
        stats = self.ctx.scan(self.scan_desc)
        rec = self.ctx.records(0, self.shape[0] * self.shape[1] * self.shape[2])
        self.last_records, self.last_stats = rec, stats
        # records report the requested pixel (x, y); with a unit cell the
        # scan ran at its wrapped image, where the tip relaxed
        x, y = rec["x"].copy(), rec["y"].copy()
        if self.desc.tip_pbc:
            x, y = wrap_dummy_xy(self.desc.cell, self.desc.offset, x, y)
        dummy = np.stack([x, y, rec["z"]], 1)
        tip = dummy + np.stack([rec["rx"], rec["ry"], rec["rz"]], 1)
        return rec["fz"].copy(), tip, dummy

    def adjoint_np(self, es6, es12, qq, pos, tip, dummy, g):
        self.ctx.set_systems([self._system(es6, es12, qq, pos)])
        return self.ctx.adjoint(0, len(es6), tip, dummy, g, self.adjoint_precision)


class AFMImage(torch.autograd.Function):
    @staticmethod
    def forward(ctx, es6, es12, qq, pos, model: ScanModel):
        a = [t.detach().cpu().numpy().astype(np.float64) for t in (es6, es12, qq, pos)]
        fz, tip, dummy = model.forward_np(*a)
        ctx.model = model
        ctx.arrays = a
        ctx.tip, ctx.dummy = tip, dummy
        return torch.from_numpy(fz)

    @staticmethod
    def backward(ctx, g_fz):
        g = g_fz.detach().cpu().numpy().astype(np.float64)
        d_es6, d_es12, d_qq, d_pos = ctx.model.adjoint_np(*ctx.arrays, ctx.tip, ctx.dummy, g)
        return (torch.from_numpy(d_es6), torch.from_numpy(d_es12), torch.from_numpy(d_qq),
                torch.from_numpy(d_pos), None)
