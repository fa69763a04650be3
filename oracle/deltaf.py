"""Frequency-shift images of simple-vafm-image.py (MechAFM's post-processing
script), restated with numpy.

TEST INFRASTRUCTURE ONLY: the checker of libmafm's mafm_deltaf_weights /
mafm_delta_f. Nothing in the product imports it.

The script (Python 2, matplotlib) reads the per-height files, sorts them by
name, and for every window of heights computes Giessibl's weighted sum of
Fz (Beilstein J. Nanotechnol. 3:238 (2012)). Its arithmetic is restated here
with the same Python / numpy operations, so float semantics (Python's % and
//, numpy's pairwise sums and arange) are those of the script itself:

* heights: the file names' "%06.3f" values, ascending (:52-56), z of file i
  = zhigh - i*dz (mechafm.cpp:59-75);
* dz: mean of the differences of those heights (:70);
* the weight window (:79-92), including its quirks: the amplitude grows by
  dz/2 when (1 + A/dz) % (1 + A//dz) < 0.01, the negative half is
  zosc[-1:1:-1] (it lacks -dz), and both end points are dropped;
* image pl (:196-204): -fshiftfac * sum_t Fz[:, :, pl + t] * W[t] * dz over
  ascending heights pl .. pl + len(W) - 1, for pl < nz - len(W).
"""
from __future__ import annotations

from typing import Sequence, Tuple

import numpy as np

NM2KCALAA = 1.438978          # simple-vafm-image.py:80


def file_heights(zhigh: float, dz: float, nz: int) -> list:
    """Heights carried by the scan-%06.3f.dat names, in the order the
    script's sorted glob returns them (ascending for 0 <= z < 100)."""
    names = ["scan-%06.3f.dat" % (zhigh - i * dz) for i in range(nz)]
    return [float(".".join(n.split("-")[1].split(".")[:2])) for n in sorted(names)]


def weights(zpos: Sequence[float], amplitude: float = 0.50, k_cantilever: float = 1800.0,
            frequency0: float = 25.000e3) -> Tuple[np.ndarray, float, float, float]:
    """(W, fshiftfac, dz, adjusted amplitude) of simple-vafm-image.py:70-92."""
    dz = (np.array(zpos)[1:] - np.array(zpos)[:-1]).mean()
    k = k_cantilever * NM2KCALAA
    fshiftfac = (frequency0 / (2 * k)) * (2 / (np.pi * amplitude * amplitude))
    ampcut = (1 + (amplitude / dz)) % (1 + (amplitude // dz))
    if ampcut < 0.01:
        amplitude += dz / 2
    zosc = np.arange(0, amplitude + dz, dz)
    zosc = np.concatenate((-zosc[-1:1:-1], zosc))
    with np.errstate(divide="ignore"):
        w = zosc / np.sqrt(abs((amplitude ** 2) - (zosc ** 2)))
    return w[1:-1], float(fshiftfac), float(dz), float(amplitude)


def images(fz_ascending: np.ndarray, zpos: Sequence[float], **kw) -> Tuple[np.ndarray, np.ndarray]:
    """Delta f images [n_images, nx, ny] (Hz) and each image's lowest height
    from Fz[nx, ny, nz] ordered by ascending height."""
    w, fshiftfac, dz, _ = weights(zpos, **kw)
    nx, ny, nz = fz_ascending.shape
    nosc = len(w)
    nshifts = nz - nosc
    wxy = np.tile(w, [nx, ny, 1])
    out = np.zeros((max(nshifts, 0), nx, ny))
    for pl in range(nshifts):
        fz = fz_ascending[:, :, pl:pl + nosc]
        out[pl] = -fshiftfac * (fz * wxy).sum(axis=2) * dz
    return out, np.array(zpos[:max(nshifts, 0)])


def fz_text(fz: np.ndarray) -> np.ndarray:
    """Fz as the script reads it (simple-vafm-image.py:147, 183): the value
    printed "%8.4f" by the scan's writer (simulation.cpp:180-195) and parsed
    back with float()."""
    return np.array([float("%8.4f" % v) for v in np.ravel(fz)]).reshape(np.shape(fz))


def fz_from_files(files: dict, nx: int, ny: int) -> np.ndarray:
    """[nx, ny, nz] Fz in ascending height from per-height files given as
    {name "scan-%06.3f.dat": lines}, as the script fills FZ (column 8 of a
    line "k i j x y z fx fy fz ...")."""
    names = sorted(files, key=lambda n: float(n[5:11]))
    out = np.zeros((nx, ny, len(names)))
    for z, name in enumerate(names):
        for line in files[name]:
            t = line.split()
            out[int(t[1]), int(t[2]), z] = float(t[8])
    return out


def fz_from_records(fz_records: np.ndarray, nx: int, ny: int, nz: int) -> np.ndarray:
    """Fz of the records (index (i*ny + j)*nz + k, k = 0 at zhigh) as
    [nx, ny, nz] with ascending height."""
    return np.ascontiguousarray(fz_records.reshape(nx, ny, nz)[:, :, ::-1])
