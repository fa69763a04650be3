"""Writes the golden fixtures under tests/golden/<case>/ from the reference.

Run in a container that holds /root/reference (after ``make -C oracle ref``):

    python tests/golden/make_golden.py

For every case it generates the seeded inputs (mechafm_amd.workloads), runs
``oracle/_ref/ref_harness`` -- the reference's own objects, built from its
sources -- and stores

* the input files the reference read (cube files gzipped),
* ``records.npz``: the full-precision records of every scan position
  (i j k n_steps x y z fx fy fz rx ry rz r angle e),
* for grid / external-potential cases, ``eval.npz``: tip positions and the
  reference's System::evalTipSurfaceForces at them.

Scale cases (tests/golden/scale/<case>/) hold the reference's records for
workloads at their own size, whose inputs are too large to commit (the
2000-atom slab's 160 x 160 x 80 cube): ``inputs.json`` names the seeded
generator call and the SHA-256 of every file it writes, so a test
regenerates the inputs and checks they are the ones the reference read.

Only data is committed; no reference source or binary enters the repository.
"""
from __future__ import annotations

import gzip
import os
import shutil
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from mechafm_amd import workloads as W  # noqa: E402

HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
OUT = os.path.dirname(os.path.abspath(__file__))

# name -> (maker, kwargs, eval samples)
CASES = {
    "c1_mini": (W.make_c1, dict(mini=True), 0),
    "c1_rk4": (W.make_c1, dict(mini=True, area="0.5 0.5", integrator="rk4"), 0),
    "c1_euler": (W.make_c1, dict(mini=True, area="0.5 0.5", integrator="euler"), 0),
    "c1_sd": (W.make_c1, dict(mini=True, area="0.2 0.2", minimiser="SD", zlow="9.0"), 0),
    "c1_minterm_e": (W.make_c1, dict(mini=True, area="0.5 0.5", minterm="e"), 0),
    "c1_minterm_f": (W.make_c1, dict(mini=True, area="0.5 0.5", minterm="f", ftol="0.001"), 0),
    "c1_maxsteps": (W.make_c1, dict(mini=True, area="0.5 0.5", maxsteps="20"), 0),
    "c1_pbc": (W.make_c1, dict(mini=True, area="0.5 0.5", vdw_pbc="on", cell_a="14.0 0.0 0.0",
                               cell_b="0.0 14.0 0.0", cell_c="0.0 0.0 30.0", center="7.0 7.0"), 0),
    "c1_pbc_tilt": (W.make_c1, dict(mini=True, area="0.4 0.4", vdw_pbc="on", cell_a="14.0 0.0 0.5",
                                    cell_b="0.0 14.0 -0.3", cell_c="0.0 0.0 30.0", center="7.0 7.0"), 0),
    "c2_mini": (W.make_c2, dict(mini=True), 512),
    "c2_direct": (W.make_c2, dict(mini=True, rigidgrid="off", area="0.6 0.6"), 512),
    "c2_morse": (W.make_c2, dict(mini=True, rigidgrid="off", coulomb="off", area="0.6 0.6"), 512),
    "c2_morse_pbc": (W.make_c2, dict(mini=True, rigidgrid="off", coulomb="off", area="0.4 0.4", vdw_pbc="on",
                                     cell_a="16.0 0.0 0.0", cell_b="0.0 16.0 0.0", cell_c="0.0 0.0 30.0",
                                     center="0.2 0.2"), 0),
    "c3_mol": (lambda d, **k: W.make_c3(d, count=3, **k)[2], dict(mini=True), 0),
    "c4_mini": (W.make_c4, dict(mini=True), 512),
    "c5_mini": (W.make_c5, dict(mini=True), 0),
    "c5_param_ev": (W.make_c5, dict(mini=True, units="eV", tip_dummy_coulomb="on", coulomb="on"), 0),
    "c5_kj": (W.make_c5, dict(mini=True, units="kJ/mol", area="0.6 0.6"), 0),
    "c1_pixel": (W.make_c1, dict(mini=True, area="0.0 0.0", center="0.3 0.2"), 0),
    "c1_height": (W.make_c1, dict(mini=True, area="0.5 0.5", zlow="10.0"), 0),
    "c1_maxsteps1": (W.make_c1, dict(mini=True, area="0.3 0.3", maxsteps="1"), 0),
    "c1_normal_x": (lambda d, **k: W.make_c1_normal(d, "x", **k), {}, 0),
    "c1_normal_y": (lambda d, **k: W.make_c1_normal(d, "y", **k), {}, 0),
    "c4_bohr": (lambda d, **k: W.make_c4_variant(d, "bohr", **k), {}, 512),
    "c4_skew": (lambda d, **k: W.make_c4_variant(d, "skew", **k), {}, 512),
    "c4_normal_y": (lambda d, **k: W.make_c4_variant(d, "normal_y", **k), {}, 512),
}


def harness(*args: str) -> None:
    subprocess.check_call([HARNESS, *args], stdout=subprocess.DEVNULL,
                          env=dict(os.environ, OMP_NUM_THREADS=str(os.cpu_count() or 8)))


def param_charge_variant(work: str) -> None:
    """c5_param_ev: drop the xyz charge column, give charges in the parameter
    file and add an LJ pair overwrite (exercises parse.cpp:571-583, 639-663)."""
    xyz = os.path.join(work, "molecule.xyz")
    lines = open(xyz).read().splitlines()
    out = lines[:2] + [" ".join(l.split()[:4]) for l in lines[2:] if l.strip()]
    open(xyz, "w").write("\n".join(out) + "\n")
    atoms = dict(W.BASE_ATOMS)
    atoms["C"] = atoms["C"][:3] + (-0.05,)
    atoms["N"] = atoms["N"][:3] + (-0.15,)
    atoms["H"] = atoms["H"][:3] + (0.10,)
    W.write_params(os.path.join(work, "parameters.dat"), atoms, W.HARM, ["O H lj 0.05000 2.90000"])


def make_case(name: str) -> None:
    maker, kw, n_eval = CASES[name]
    dst = os.path.join(OUT, name)
    if os.path.isdir(dst):
        shutil.rmtree(dst)
    with tempfile.TemporaryDirectory() as work:
        wl = maker(work, **kw)
        wdir = os.path.dirname(wl.input_path)
        if name == "c5_param_ev":
            param_charge_variant(wdir)
        rec_bin = os.path.join(wdir, "records.bin")
        harness("scan", wl.input_path, rec_bin)
        raw = open(rec_bin, "rb").read()
        hdr = np.frombuffer(raw[:16], dtype=np.int32)
        rec = np.frombuffer(raw[16:], dtype=np.float64).reshape(-1, int(hdr[3]))
        os.makedirs(dst)
        for fn in sorted(os.listdir(wdir)):
            src = os.path.join(wdir, fn)
            if fn.endswith(".bin"):
                continue
            if fn.endswith(".cube"):
                with open(src, "rb") as fi, gzip.open(os.path.join(dst, fn + ".gz"), "wb", 9) as fo:
                    fo.write(fi.read())
            else:
                shutil.copy(src, os.path.join(dst, fn))
        np.savez_compressed(os.path.join(dst, "records.npz"), shape=hdr[:3], records=rec)
        if n_eval:
            rng = np.random.default_rng(99)
            ax, ay = (float(v) for v in wl.options["area"].split())
            zl, zh = float(wl.options["zlow"]), float(wl.options["zhigh"])
            pos = np.stack([rng.uniform(0, ax, n_eval), rng.uniform(0, ay, n_eval),
                            rng.uniform(zl - 3.76, zh - 3.76, n_eval)], axis=1)
            pin = os.path.join(wdir, "pos.bin")
            pout = os.path.join(wdir, "eval.bin")
            pos.astype(np.float64).tofile(pin)
            harness("eval", wl.input_path, pin, pout)
            ev = np.fromfile(pout, dtype=np.float64).reshape(-1, 4)
            np.savez_compressed(os.path.join(dst, "eval.npz"), pos=pos, fe=ev)
        print(f"{name}: {rec.shape[0]} records, steps mean {rec[:, 3].mean():.1f} max {rec[:, 3].max():.0f}")


# flexible on: name -> (maker, kwargs, perturbed configurations for the
# whole-list force check).  Each case also keeps the centre column's relaxed
# atoms at every height (state.npy: what System::makeXYZFile writes).
FLEX_CASES = {
    "fx_mini": (W.make_fx, {}, 4),
    "fx_rk4": (W.make_fx, dict(integrator="rk4"), 0),
    "fx_euler": (W.make_fx, dict(integrator="euler"), 0),
    "fx_sd": (W.make_fx, dict(minimiser="SD", maxsteps="3000", area="0.25 0.25"), 0),
    "fx_minterm_e": (W.make_fx, dict(minterm="e", etol="1e-8"), 0),
    "fx_maxsteps": (W.make_fx, dict(maxsteps="40"), 0),
    "fx_pbc": (W.make_fx, dict(coulomb="off", vdw_pbc="on", cell_a="16.0 0.0 0.0", cell_b="0.0 16.0 0.0",
                               cell_c="0.0 0.0 30.0", center="7.9 8.1", area="0.5 0.5", zhigh="8.0",
                               zlow="7.0"), 4),
    "fx_morse": (lambda d, **k: W.make_fx(d, **k), {}, 4),
    "fx_graphene": (W.make_fx_example, dict(area="1.0 1.0", zlow="9.0"), 4),
}


def morse_variant(work: str) -> None:
    """fx_morse: Morse pair overwrites for the tip (O) against C, H and N and
    for the C-N pairs between the two molecules."""
    W.write_params(os.path.join(work, "parameters.dat"), W.BASE_ATOMS, W.HARM,
                   W.morse_overwrites("O", ["C", "H", "N"]) + ["C N morse 0.118322 1.6000 3.798000"],
                   W.FLEX_PARAM_LINES)


def make_flex_full() -> None:
    """fx_example_full: the reference's example job with flexible on at its
    own size (41 x 31 columns x 41 heights, 72 free atoms): every position's
    FIRE step count, Fz and tip energy, the centre column's relaxed atoms at
    every height (the inputs are the example files kept with the tests), and
    the reference's sensitivity profile: how far its own answer moves when
    one input coordinate changes in the last printed digit."""
    dst = os.path.join(SCALE_DIR, "fx_example_full")
    if os.path.isdir(dst):
        shutil.rmtree(dst)
    os.makedirs(dst)
    with tempfile.TemporaryDirectory() as work:
        wl = W.make_fx_example(work)
        rec_bin, st_bin = os.path.join(work, "records.bin"), os.path.join(work, "state.bin")
        harness("scan", wl.input_path, rec_bin, st_bin)
        raw = open(rec_bin, "rb").read()
        hdr = np.frombuffer(raw[:16], dtype=np.int32)
        rec = np.frombuffer(raw[16:], dtype=np.float64).reshape(-1, int(hdr[3]))
        n_atoms = len(wl.molecule.types) + 2
        state = np.fromfile(st_bin, dtype=np.float64).reshape(int(hdr[2]), n_atoms, 3)
        # the reference's own sensitivity: the same job with one input
        # coordinate changed in its last printed digit (+1e-5 A); per height,
        # the fraction of positions keeping their iteration count and the
        # 50 / 90 / 99 % quantiles and maximum of |dFz|
        xyz = os.path.join(work, wl.options["xyzfile"])
        lines = open(xyz).read().splitlines()
        t = lines[6].split()
        lines[6] = " ".join([t[0], f"{float(t[1]) + 1e-5:.5f}"] + t[2:])
        open(xyz, "w").write("\n".join(lines) + "\n")
        harness("scan", wl.input_path, rec_bin)
        raw = open(rec_bin, "rb").read()
        pert = np.frombuffer(raw[16:], dtype=np.float64).reshape(-1, int(hdr[3]))
        nz = int(hdr[2])
        k = np.arange(len(rec)) % nz
        same = pert[:, 3] == rec[:, 3]
        dfz = np.abs(pert[:, 9] - rec[:, 9])
        sens = np.array([[same[k == kk].mean()] + list(np.quantile(dfz[k == kk], [0.5, 0.9, 0.99])) +
                         [dfz[k == kk].max()] for kk in range(nz)])
        np.savez_compressed(os.path.join(dst, "records.npz"), shape=hdr[:3], n_steps=rec[:, 3].astype(np.int32),
                            fz=rec[:, 9], e=rec[:, 15], state=state, sensitivity=sens)
    print(f"fx_example_full: {rec.shape[0]} records, {int(rec[:, 3].sum())} steps, "
          f"{same.mean():.3f} keep their count under a 1e-5 A input change")


def make_flex_case(name: str) -> None:
    maker, kw, n_conf = FLEX_CASES[name]
    dst = os.path.join(OUT, name)
    if os.path.isdir(dst):
        shutil.rmtree(dst)
    with tempfile.TemporaryDirectory() as work:
        wl = maker(work, **kw)
        wdir = os.path.dirname(wl.input_path)
        if name == "fx_morse":
            morse_variant(wdir)
        rec_bin, st_bin = os.path.join(wdir, "records.bin"), os.path.join(wdir, "state.bin")
        harness("scan", wl.input_path, rec_bin, st_bin)
        raw = open(rec_bin, "rb").read()
        hdr = np.frombuffer(raw[:16], dtype=np.int32)
        rec = np.frombuffer(raw[16:], dtype=np.float64).reshape(-1, int(hdr[3]))
        n_atoms = len(wl.molecule.types) + 2
        state = np.fromfile(st_bin, dtype=np.float64).reshape(int(hdr[2]), n_atoms, 3)
        os.makedirs(dst)
        for fn in sorted(os.listdir(wdir)):
            if not fn.endswith(".bin"):
                shutil.copy(os.path.join(wdir, fn), os.path.join(dst, fn))
        np.savez_compressed(os.path.join(dst, "records.npz"), shape=hdr[:3], records=rec, state=state)
        if n_conf:
            # the surface as the scan starts (centred) perturbed by 0.05 A, tip 3.4 A above the centre
            rng = np.random.default_rng(7)
            base = state[0].copy()
            cx, cy = (float(v) for v in wl.options.get("center", "0 0").split()) if "center" in wl.options \
                else (0.0, 0.0)
            confs = []
            for c in range(n_conf):
                p = base + rng.normal(0.0, 0.05, base.shape)
                p[0] = (cx, cy, 9.0 + 0.1 * c)
                p[1] = (cx + 0.1, cy - 0.05, 5.6 + 0.1 * c)
                confs.append(p)
            confs = np.array(confs)
            pin, pout = os.path.join(wdir, "pos.bin"), os.path.join(wdir, "forces.bin")
            confs.tofile(pin)
            harness("forces", wl.input_path, pin, pout)
            fe = np.fromfile(pout, dtype=np.float64).reshape(n_conf, 4 * n_atoms)
            np.savez_compressed(os.path.join(dst, "forces.npz"), pos=confs,
                                forces=fe[:, :3 * n_atoms].reshape(n_conf, n_atoms, 3), energies=fe[:, 3 * n_atoms:])
    print(f"{name}: {rec.shape[0]} records, steps mean {rec[:, 3].mean():.1f} max {rec[:, 3].max():.0f}")


# scale cases: name -> (generator name in mechafm_amd.workloads, kwargs, eval samples)
SCALE = {
    # the periodic slab at full size (2000 atoms x 9 images, 160 x 160 x 80
    # Hartree cube), 4 x 4 pixels spread over the cell, all 40 heights
    "c4_full": ("make_c4", dict(dx="13.28125", dy="13.28125"), 512),
    # all 64 molecules of the batched workload, 2 x 2 pixels each, 40 heights
    "c3_batch": ("make_c3", dict(area="0.1 0.1"), 0),
}
SCALE_DIR = os.path.join(OUT, "scale")


def file_hashes(folder: str) -> dict:
    import hashlib
    return {fn: hashlib.sha256(open(os.path.join(folder, fn), "rb").read()).hexdigest()
            for fn in sorted(os.listdir(folder)) if not fn.endswith(".bin")}


def make_scale_case(name: str) -> None:
    import json
    gen, kw, n_eval = SCALE[name]
    dst = os.path.join(SCALE_DIR, name)
    if os.path.isdir(dst):
        shutil.rmtree(dst)
    os.makedirs(dst)
    with tempfile.TemporaryDirectory() as work:
        wls = getattr(W, gen)(work, **kw)
        wls = wls if isinstance(wls, list) else [wls]
        shapes, recs, hashes = [], [], []
        for wl in wls:
            wdir = os.path.dirname(wl.input_path)
            rec_bin = os.path.join(wdir, "records.bin")
            harness("scan", wl.input_path, rec_bin)
            raw = open(rec_bin, "rb").read()
            hdr = np.frombuffer(raw[:16], dtype=np.int32)
            recs.append(np.frombuffer(raw[16:], dtype=np.float64).reshape(-1, int(hdr[3])))
            shapes.append(hdr[:3].copy())
            os.remove(rec_bin)
            hashes.append({"folder": os.path.relpath(wdir, work), "files": file_hashes(wdir)})
        np.savez_compressed(os.path.join(dst, "records.npz"), shapes=np.array(shapes), records=np.concatenate(recs))
        json.dump({"generator": gen, "kwargs": kw, "systems": hashes}, open(os.path.join(dst, "inputs.json"), "w"),
                  indent=1)
        if n_eval:
            wl = wls[0]
            wdir = os.path.dirname(wl.input_path)
            rng = np.random.default_rng(99)
            ax, ay = (float(v) for v in wl.options["area"].split())
            zl, zh = float(wl.options["zlow"]), float(wl.options["zhigh"])
            pos = np.stack([rng.uniform(0, ax, n_eval), rng.uniform(0, ay, n_eval),
                            rng.uniform(zl - 3.76, zh - 3.76, n_eval)], axis=1)
            pin, pout = os.path.join(wdir, "pos.bin"), os.path.join(wdir, "eval.bin")
            pos.astype(np.float64).tofile(pin)
            harness("eval", wl.input_path, pin, pout)
            ev = np.fromfile(pout, dtype=np.float64).reshape(-1, 4)
            np.savez_compressed(os.path.join(dst, "eval.npz"), pos=pos, fe=ev)
    n = sum(len(r) for r in recs)
    print(f"{name}: {len(wls)} systems, {n} records, steps mean {np.concatenate(recs)[:, 3].mean():.1f}")


REF_EXAMPLES = "/root/reference/examples"
REF_BINARY = os.path.join(ROOT, "oracle", "_ref", "mechafm-omp")


def make_example() -> None:
    """example_graphene: the reference's own example job (examples/input.scan,
    graphene.xyz, parameters.dat) and the per-height files the reference
    program writes for it, each file's lines sorted (the reference writes
    them in thread order), concatenated under '## <name>' headers."""
    dst = os.path.join(OUT, "example_graphene")
    if os.path.isdir(dst):
        shutil.rmtree(dst)
    os.makedirs(dst)
    for fn in ("input.scan", "graphene.xyz", "parameters.dat"):
        shutil.copy(os.path.join(REF_EXAMPLES, fn), os.path.join(dst, fn))
    with tempfile.TemporaryDirectory() as work:
        for fn in ("input.scan", "graphene.xyz", "parameters.dat"):
            shutil.copy(os.path.join(dst, fn), os.path.join(work, fn))
        out = os.path.join(work, "out")
        os.makedirs(out)
        run = subprocess.run([REF_BINARY, "input.scan", out], cwd=work, capture_output=True, text=True, check=True,
                             env=dict(os.environ, OMP_NUM_THREADS=str(os.cpu_count() or 8)))
        # the console text up to the scan (warnings on stderr, settings echo on stdout)
        echo = run.stdout.split("+- 3D data grid")[0]
        with open(os.path.join(dst, "console.txt"), "w") as f:
            f.write(run.stderr + echo)
        parts = []
        for fn in sorted(os.listdir(out)):
            if fn.startswith("scan-"):
                lines = sorted(open(os.path.join(out, fn)).read().splitlines())
                parts.append("## " + fn + "\n" + "\n".join(lines) + "\n")
        with gzip.open(os.path.join(dst, "outputs.txt.gz"), "wt", 9) as f:
            f.write("".join(parts))
    print(f"example_graphene: {len(parts)} files")


def main(names=None) -> None:
    if not os.path.exists(HARNESS):
        raise SystemExit("build the reference harness first: make -C oracle ref")
    for name in names or list(CASES) + ["example_graphene"] + list(FLEX_CASES) + list(SCALE) + ["fx_example_full"]:
        if name == "example_graphene":
            make_example()
        elif name == "fx_example_full":
            make_flex_full()
        elif name in FLEX_CASES:
            make_flex_case(name)
        elif name in SCALE:
            make_scale_case(name)
        else:
            make_case(name)


if __name__ == "__main__":
    main(sys.argv[1:] or None)
