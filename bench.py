"""Throughput of the rigid probe-particle scan on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c3]

A step is one full scan of the workload (every pixel column, every height,
FIRE + midpoint to convergence) in one launch of the work-stealing kernel,
with the molecules' tables already resident in HBM.  Default workload: the
batched one (64 seeded planar molecules, N ~ U[20, 400], LJ + Coulomb,
64 x 64 px x 40 z each).

Headline arithmetic: fast64 (FP64 pair terms with refined hardware
reciprocals, FP64 minimiser), the arithmetic that keeps every position
within 1e-4 A / 1e-3 relative Fz of the plain FP64 scan.  The FP32 scan
(fast32: FP32 pair terms summed in register blocks folded into FP64) is
timed beside it and reported with the fraction of positions it keeps within
that bar: an FP32 convergence test can come out one FIRE step away from the
FP64 one, which moves that position and the rest of its column
(``--margins`` prints the decision-margin analysis behind this).

Printed JSON line: metric/value (relaxed tip positions per second, whole
job), the roofline of the headline scan kernel (algorithmic VALU FLOPs per
launch over its HIP-event duration, against the 78.6 TFLOP/s FP64 or
157.3 TFLOP/s FP32 vector peak; lookup bytes against 8 TB/s for a
force-grid scan), the FP32 scan, the plain FP64 GPU scan on the same
workload (speed and the fraction of positions within 1e-4 A / 1e-3 relative
Fz of it), and a CPU baseline: the reference program timed on this host on
a sample of the same workload.
"""
import argparse
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_VALU_PEAK = 157.3e12      # MI355X FP32 vector (packed math), FLOP/s
FP32_UNPACKED_PEAK = 78.6e12   # one FP32 FMA per lane per 4-cycle wave issue, FLOP/s
FP64_VALU_PEAK = 78.6e12       # MI355X FP64 vector, FLOP/s
HBM_PEAK = 8.0e12              # MI355X HBM3E, B/s
LOOKUP_BYTES = 8 * 16          # one trilinear lookup: 8 corners of {fx, fy, fz, e} in FP32 (x2 in FP64)
# algorithmic FLOPs per pair term and per evaluation / iteration overheads
PAIR_FLOPS = {(0, 0): 26, (0, 1): 31, (1, 0): 28, (1, 1): 33}   # (vdw, coulomb)
TIP_DUMMY_FLOPS = 50
ITER_FLOPS = 52
EVALS_PER_ITER = {0: 1, 1: 2, 2: 4}                             # euler, midpoint, rk4


def workloads(name, tmp):
    from mechafm_amd import workloads as W
    if name == "c3":
        return W.make_c3(tmp)
    return [W.MAKERS[name](tmp)]


def pair_cost(sim):
    """Algorithmic FLOPs of one evaluation of a system's sites (every
    periodic image counts: the reference evaluates 9 x N pairs with vdw_pbc)."""
    d = sim.system()
    per_image = sum(PAIR_FLOPS[(s.vdw, int(s.has_coulomb and s.qq != 0.0))] for s in sim.sites())
    return per_image * max(1, d.n_images)


def algorithmic_flops(sims, recs, integrator=1):
    e = EVALS_PER_ITER[integrator]
    tot = 0.0
    for s, r in zip(sims, recs):
        c = pair_cost(s)
        steps = float(r["n_steps"].astype(np.int64).sum())
        tot += steps * (e * (c + TIP_DUMMY_FLOPS) + ITER_FLOPS) + len(r) * c
    return tot


def grid_lookup_bytes(sims, recs, integrator=1, fp64=False):
    """Algorithmic bytes of the force-grid lookups of a rigidgrid scan: one
    lookup per evaluation (E per minimiser iteration + the final one), FP32
    samples for fast32, FP64 samples otherwise."""
    e = EVALS_PER_ITER[integrator]
    per = LOOKUP_BYTES * (2 if fp64 else 1)
    return float(sum((r["n_steps"].astype(np.int64).sum() * e + len(r)) * per for r in recs))


def run_batch(abi, sims, prec, **kw):
    if len(sims) == 1:
        return sims[0].run(prec, **kw)
    return abi.batch_run(sims, prec, **kw)


def cpu_baseline(name, tmp, budget_s=25.0):
    """The reference program (oracle/_ref/mechafm-omp, built from the
    reference sources) on a sample of the same workload: every molecule, the
    same scan area and heights, every 8th pixel in x and y (dx = dy = 0.8)."""
    exe = os.path.join(ROOT, "oracle", "_ref", "mechafm-omp")
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    from mechafm_amd import workloads as W
    sample_dir = os.path.join(tmp, "cpu_sample")
    if name == "c3":
        wls = W.make_c3(sample_dir, dx="0.8", dy="0.8")
    else:
        wls = [W.MAKERS[name](sample_dir, dx="0.8", dy="0.8")]
    positions, elapsed = 0, 0.0
    kind = "reference" if os.path.exists(exe) else "port"
    for w in wls:
        shape = W.scan_shape(w.options)
        out = os.path.join(tmp, "cpu_out")
        t0 = time.perf_counter()
        if kind == "reference":
            subprocess.run([exe, w.input_path, out], check=True, stdout=subprocess.DEVNULL,
                           stderr=subprocess.DEVNULL, env=dict(os.environ, OMP_NUM_THREADS=str(threads)))
        else:
            from oracle import oracle as O
            O.scan(O.setup(w.input_path))
        elapsed += time.perf_counter() - t0
        positions += shape[0] * shape[1] * shape[2]
        shutil.rmtree(out, ignore_errors=True)
        if elapsed > budget_s:
            break
    return {"value": positions / elapsed, "unit": "positions/s", "cores": threads, "kind": kind,
            "sample": f"{positions} positions: {len(wls)} system(s) of the workload at every 8th pixel in "
                      f"x and y (dx = dy = 0.8 A), all heights; {elapsed:.1f} s wall on {threads} threads"}


def adjoint_bench(tmp, args):
    """Config 5: gradients of an image loss through the implicit-function
    adjoint, timed, and checked against central finite differences.

    The adjoint differentiates the relaxed (minimum) tip position, so the
    scans of this check relax the tip to |F| < 1e-8 (FP64 pair terms) instead
    of the workload's 0.01, and positions where a parameter step moves the
    tip to another minimum (the response is not smooth: the central
    differences with steps h and h/2 disagree) or that do not converge are
    left out of the loss for both the adjoint and the finite differences;
    their fraction is reported."""
    import torch
    from mechafm_amd import abi
    from mechafm_amd import workloads as W
    from mechafm_amd.adjoint import ScanModel
    ctx = abi.Context(0)
    wl = W.make_c5(tmp, minterm="f", ftol="1e-8", maxsteps="200000")
    m = ScanModel(wl.input_path, ctx, precision="fast64", adjoint_precision="fast32")
    rng = np.random.default_rng(5)

    def image(**kw):
        with torch.no_grad():
            a = dict(eps=m.eps, sig=m.sig, q=m.q, pos=m.pos)
            a.update(kw)
            img = m.image(**a).numpy().copy()
        return img, m.last_records["n_steps"].copy()

    target, _ = image(eps=m.eps * torch.from_numpy(1 + 0.05 * rng.choice([-1, 1], len(m.eps))),
                      sig=m.sig * torch.from_numpy(1 + 0.05 * rng.choice([-1, 1], len(m.sig))),
                      q=m.q * torch.from_numpy(1 + 0.05 * rng.choice([-1, 1], len(m.q))))
    t0 = time.perf_counter()
    base, steps0 = image()
    t_scan = time.perf_counter() - t0
    ok = steps0 < 200000

    # parameters checked: eps and sigma of every type, the charge and the three
    # coordinates with the largest gradient of the full loss
    leaves = [t.clone().requires_grad_(True) for t in (m.eps, m.sig, m.q, m.pos)]
    ((m.image(*leaves) - torch.from_numpy(target)) ** 2).sum().backward()
    params = []
    for k in range(len(m.eps)):
        params.append((f"eps[{m.type_names[k]}]", "eps", k, 1e-4 * float(m.eps[k])))
        params.append((f"sig[{m.type_names[k]}]", "sig", k, 1e-4 * float(m.sig[k])))
    i = int(torch.argmax(leaves[2].grad.abs()))
    params.append((f"q[{i}]", "q", i, 1e-4 * abs(float(m.q[i]))))
    j = int(torch.argmax(leaves[3].grad.abs().max(1).values))
    params += [(f"pos[{j},{c}]", "pos", (j, c), 1e-5) for c in range(3)]

    fz = {}
    for name, key, idx, h in params:
        for sgn, hh in ((1, h), (-1, h), (1, h / 2), (-1, h / 2)):
            t = getattr(m, key).clone()
            t[idx] += sgn * hh
            fz[(name, sgn, hh)], st = image(**{key: t})
            ok &= st < 200000
        d1 = (fz[(name, 1, h)] - fz[(name, -1, h)]) / (2 * h)
        d2 = (fz[(name, 1, h / 2)] - fz[(name, -1, h / 2)]) / h
        ok &= np.abs(d1 - d2) <= 1e-3 * np.abs(d1) + 1e-4 * np.abs(d1).max()
    mask = torch.from_numpy(ok.astype(np.float64))
    tgt = torch.from_numpy(target)

    def loss_of(img):
        return float((((img - target) ** 2) * ok).sum())

    # timed: forward scan + adjoint of the masked loss (FP32 adjoint pair terms)
    runs = []
    for r in range(args.warmup + max(1, args.steps)):
        lv = [t.clone().requires_grad_(True) for t in (m.eps, m.sig, m.q, m.pos)]
        t0 = time.perf_counter()
        img = m.image(*lv)
        loss = (((img - tgt) ** 2) * mask).sum()
        t1 = time.perf_counter()
        loss.backward()
        t2 = time.perf_counter()
        if r >= args.warmup:
            runs.append((t1 - t0, t2 - t1))
        g32 = [t.grad.clone() for t in lv]
    m.adjoint_precision = "fast64"
    lv = [t.clone().requires_grad_(True) for t in (m.eps, m.sig, m.q, m.pos)]
    (((m.image(*lv) - tgt) ** 2) * mask).sum().backward()
    g64 = [t.grad.clone() for t in lv]
    rel32 = max(float(torch.linalg.norm(a - b) / torch.linalg.norm(b)) for a, b in zip(g32, g64))
    grad = dict(zip(("eps", "sig", "q", "pos"), g64))
    checks = []
    for name, key, idx, h in params:
        f = (loss_of(fz[(name, 1, h)]) - loss_of(fz[(name, -1, h)])) / (2 * h)
        checks.append({"param": name, "adjoint": float(grad[key][idx]), "finite_difference": f})
    err = max(abs(c["adjoint"] - c["finite_difference"]) / max(abs(c["finite_difference"]), 1e-12) for c in checks)
    pos = m.shape[0] * m.shape[1] * m.shape[2]
    fwd = float(np.mean([r[0] for r in runs]))
    bwd = float(np.mean([r[1] for r in runs]))
    out = {"metric": "adjoint gradient of an image loss vs central finite differences (max relative error)",
           "value": err, "unit": "relative error", "n_gpus": 1, "steps": max(1, args.steps), "warmup": args.warmup,
           "higher_is_better": False, "dtype": "f32", "data": "synthetic",
           "config": {"workload": "150-atom seeded molecule, LJ + Coulomb, 128x128 px x 30 z, loss = sum (Fz - Fz*)^2 "
                                  "against a target with eps/sigma/q perturbed by 5 %, tip relaxed to |F| < 1e-8",
                      "positions": pos, "sites": len(m.types), "types": m.type_names},
           "positions_in_loss": int(ok.sum()), "positions_left_out": int(pos - ok.sum()),
           "forward_ms": fwd * 1e3, "backward_ms": bwd * 1e3, "scan_ms": t_scan * 1e3,
           "backward_positions_per_s": pos / bwd,
           "fp32_vs_fp64_gradient_rel": rel32, "checks": checks}
    print(json.dumps(out), flush=True)


# algorithmic FP64 operations of one evaluation of each flexible term kind, counted
# from the reference's formulas (interactions.cpp:15-51, 252-362); a libm call
# (exp, acos, sin, atan) counts as one
FLEX_TERM_FLOPS = {0: 24, 1: 30, 2: 16, 3: 16, 4: 19, 5: 71, 6: 137, 7: 24, 8: 12}
FLEX_TERM_ARITY = {0: 2, 1: 2, 2: 2, 3: 2, 4: 2, 5: 3, 6: 4, 7: 1, 8: 1}
FLEX_ATOM_STEP_FLOPS = 40      # integrator update, P term, FIRE velocity mix of one free atom


def flexible_bench(tmp, args):
    """The reference's example job with flexible on at its own size (72 free
    atoms, 781 terms, 41 x 31 columns x 41 heights): every surface atom
    relaxes with the tip (mafm_scan_flexible)."""
    from mechafm_amd import abi, workloads as W
    wl = W.make_fx_example(os.path.join(tmp, "fx"))
    ctx = abi.Context(0)
    sim = abi.Simulation(wl.input_path, ctx)
    info = sim.info()
    positions = info["nx"] * info["ny"] * info["nz"]
    for _ in range(args.warmup):
        sim.run("plain64")
    kms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st = sim.run("plain64")
        kms.append(st["kernel_ms"])
    wall = time.perf_counter() - t0
    fx = sim.flex_system()
    t = fx["terms"]
    free = fx["fixed"] != 1
    act = np.array([free[t["atom"][q][:FLEX_TERM_ARITY[int(t["kind"][q])]]].any() for q in range(len(t))])
    term_flops = sum(FLEX_TERM_FLOPS[int(k)] for k in t["kind"][act])
    slots = sum(FLEX_TERM_ARITY[int(k)] for k in t["kind"][act])
    evals = EVALS_PER_ITER[1]
    steps = st["total_steps"]
    flops = steps * (evals * (term_flops + 3 * slots) + int(free.sum()) * FLEX_ATOM_STEP_FLOPS)
    k = float(np.mean(kms)) / 1e3
    out = {"metric": "flexible scan: relaxed tip positions/s (reference example, flexible on)",
           "value": positions * args.steps / wall, "unit": "positions/s", "ms_per_step": wall / args.steps * 1e3,
           "kernel_ms": k * 1e3, "dtype": "f64", "data": "the reference's examples/ job",
           "config": {"workload": "examples/graphene.xyz, 72 free atoms, 781 terms, 41x31 px x 41 z, FIRE+midpoint",
                      "positions": positions, "total_steps": steps, "terms": len(t),
                      "workgroup_threads": st["block_threads"], "workgroups": st["grid_blocks"]},
           "us_per_fire_step_per_column": k * 1e6 * st["grid_blocks"] / steps,
           "roofline": {"bound": "valu", "achieved": flops / k / 1e12, "peak": FP64_VALU_PEAK / 1e12,
                        "unit": "TFLOP/s", "frac": flops / k / FP64_VALU_PEAK, "traffic": None,
                        "flops_per_launch": flops}}
    if not args.no_cpu:
        exe = os.path.join(ROOT, "oracle", "_ref", "mechafm-omp")
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        # every 4th pixel in x and y (dx = dy = 2.0 A): 11 x 8 columns, all heights
        swl = W.make_fx_example(os.path.join(tmp, "fx_cpu"), dx="2.0", dy="2.0")
        shape = W.scan_shape(swl.options)
        n = shape[0] * shape[1] * shape[2]
        t0 = time.perf_counter()
        if os.path.exists(exe):
            kind = "reference"
            subprocess.run([exe, swl.input_path, os.path.join(tmp, "fx_cpu_out")], check=True,
                           stdout=subprocess.DEVNULL, stderr=subprocess.DEVNULL,
                           env=dict(os.environ, OMP_NUM_THREADS=str(threads)))
        else:
            kind = "port"
            from oracle import oracle_flex as OF
            OF.scan(OF.flex_setup(swl.input_path))
        el = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": n / el, "unit": "positions/s", "cores": threads, "kind": kind,
                               "sample": f"{n} positions: every 4th pixel in x and y (dx = dy = 2.0 A), all "
                                         f"heights; {el:.1f} s wall on {threads} threads"}
    print(json.dumps(out), flush=True)


def job_bench(abi, ctx, args, tmp, precision, launch, scan_s):
    """The whole job: scan + the per-height files of every molecule
    (mafm_batch_run_output: one launch streaming records to pinned host
    memory; each molecule's files are written as soon as its columns are in,
    while the rest scans), gzip on and off, against the scan alone and the
    writing alone (the same files again from the records on the host, on the
    same writer pool: mafm_batch_write_output).  A
    first, untimed job on the context allocates the pinned buffers."""
    from mechafm_amd import workloads as W
    out = {}
    warm = os.path.join(tmp, "job_warm")
    wls = W.make_c3(os.path.join(warm, "in"), gzip="off")
    sims = [abi.Simulation(w.input_path, ctx, outdir=os.path.join(warm, "out", f"mol{m:02d}"))
            for m, w in enumerate(wls)]
    abi.batch_run_output(sims, precision, threads=0, **launch)
    for s_ in sims:
        s_.close()
    shutil.rmtree(warm, ignore_errors=True)
    for gz in (True, False):
        d = os.path.join(tmp, "job_gz" if gz else "job_txt")
        wls = W.make_c3(os.path.join(d, "in"), gzip="on" if gz else "off")
        sims = [abi.Simulation(w.input_path, ctx, outdir=os.path.join(d, "out", f"mol{m:02d}"))
                for m, w in enumerate(wls)]
        t0 = time.perf_counter()
        st = abi.batch_run_output(sims, precision, threads=0, **launch)
        job = time.perf_counter() - t0
        t0 = time.perf_counter()
        abi.batch_write_output(sims, threads=0)
        write = time.perf_counter() - t0
        size = sum(os.path.getsize(os.path.join(r, f)) for r, _, fs in os.walk(os.path.join(d, "out")) for f in fs)
        out["gzip" if gz else "text"] = {
            "job_s": job, "scan_s": scan_s, "write_s": write, "ratio_to_max_scan_write": job / max(scan_s, write),
            "job_scan_kernel_ms": st["kernel_ms"], "launches": 1, "files": 40 * len(sims), "bytes": size,
            "host_threads": len(os.sched_getaffinity(0))}
        for s_ in sims:
            s_.close()
        shutil.rmtree(d, ignore_errors=True)
    return out


def bar_compare(recs, prec_recs):
    """Per-position |dr| (max over x, y, z) and relative Fz error of recs
    against prec_recs, and the 1e-4 A / 1e-3 bar."""
    dr = np.concatenate([np.abs(np.stack([a["rx"] - b["rx"], a["ry"] - b["ry"], a["rz"] - b["rz"]])).max(0)
                         for a, b in zip(recs, prec_recs)])
    dfz = np.concatenate([np.abs(a["fz"] - b["fz"]) / np.maximum(np.abs(b["fz"]), 1e-2)
                          for a, b in zip(recs, prec_recs)])
    steps = sum(int((a["n_steps"] != b["n_steps"]).sum()) for a, b in zip(recs, prec_recs))
    ok = (dr <= 1e-4) & (dfz <= 1e-3)
    return {"within_1e-4A_1e-3Fz": float(ok.mean()), "max_dr": float(dr.max()), "median_dr": float(np.median(dr)),
            "step_count_mismatches": steps}


def timed_scan(abi, sims, precision, args, dev, options, positions, tag):
    """Tune (unless fixed), warm up, then time args.steps full scans; returns
    the wall rate, the kernel times and the last records."""
    G, wps = args.lanes, args.waves
    if not args.no_tune and (G == 0 or wps == 0):
        from mechafm_amd import tune
        best = tune.tune(sims, precision, compensation=args.compensation, options=options,
                         lanes=[G] if G else tune.DEFAULT_LANES, use_cache=False, device=dev,
                         log=lambda m: print(f"# tune {tag} " + m, file=sys.stderr, flush=True))
        G, wps = best["lanes_per_pixel"], best["waves_per_simd"]
    kw = dict(lanes_per_pixel=G, waves_per_simd=wps, compensation=args.compensation,
              waves_per_group=args.group, options=options)
    for _ in range(args.warmup):
        run_batch(abi, sims, precision, **kw)
    kms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        st = run_batch(abi, sims, precision, **kw)
        kms.append(st["kernel_ms"])
    wall = time.perf_counter() - t0
    recs = [s.records().copy() for s in sims]
    return {"value": positions * args.steps / wall, "wall": wall, "ms_per_step": wall / args.steps * 1e3,
            "kernel_s": float(np.mean(kms)) / 1e3, "stats": st, "recs": recs}


def roofline_of(sims, run, precision, info):
    """Roofline object of a timed scan: algorithmic VALU FLOPs (site terms,
    tip-dummy terms and FIRE updates per the iteration counts of the records)
    over the kernel's HIP-event duration against the FP32 (fast32) or FP64
    vector peak; for a force-grid scan, lookup bytes against HBM."""
    flops = algorithmic_flops(sims, run["recs"])
    k = run["kernel_s"]
    peak = FP32_VALU_PEAK if precision == "fast32" else FP64_VALU_PEAK
    r = {"bound": "valu", "achieved": flops / k / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s",
         "frac": flops / k / peak, "traffic": None, "kernel_ms": k * 1e3, "flops_per_launch": flops}
    if precision == "fast32":
        # one VALU instruction per 4 cycles and lane without packed FP32 math
        r["frac_of_unpacked_fp32"] = flops / k / FP32_UNPACKED_PEAK
    if info["rigidgrid"]:
        gb = grid_lookup_bytes(sims, run["recs"], fp64=precision != "fast32")
        r = {"bound": "hbm", "achieved": gb / k / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
             "frac": gb / k / HBM_PEAK, "traffic": None, "kernel_ms": k * 1e3, "bytes_per_launch": gb}
    return r


def margins_report(abi, sims, ctx, positions, nz, tmp=None):
    """Why FP32 pair sums cannot keep the bar on every position: the FP32
    and FP64 scans with their decision margins (MAFM_SCAN_MARGINS) against
    the plain FP64 scan.  For every column whose FIRE step counts leave the
    FP64 ones, the margins at the first height where they differ; for every
    column, its smallest margins; and the share of columns an FP64 replay
    would have to rescan to catch them all at a given band."""
    def go(prec, options):
        st = run_batch(abi, sims, prec, options=options)
        recs = np.concatenate([s.records().copy() for s in sims])
        m = ctx.margins(0, positions)[:, :3] if options & abi.SCAN_MARGINS else None
        return st, recs, m
    st64, r64, m64 = go("fast64", abi.SCAN_MARGINS)
    st32, r32, m32 = go("fast32", abi.SCAN_MARGINS)
    stp, rp, _ = go("plain64", 0)
    out = {"positions": positions, "fast64_ms": st64["kernel_ms"], "fast32_ms": st32["kernel_ms"],
           "plain64_ms": stp["kernel_ms"], "fast64_vs_plain64": bar_compare([r64], [rp]),
           "fast32_vs_plain64": bar_compare([r32], [rp])}
    ncol = positions // nz
    s32, s64 = r32["n_steps"].reshape(ncol, nz), rp["n_steps"].reshape(ncol, nz)
    diff = s32 != s64
    div = np.where(diff.any(1))[0]
    kstar = diff[div].argmax(1)
    M32, M64 = m32.reshape(ncol, nz, 3), m64.reshape(ncol, nz, 3)
    at64 = M64[div, kstar]
    kind = np.argmin(at64, 1)
    out["divergent_columns"] = int(len(div))
    out["columns"] = int(ncol)
    out["first_divergence_decision"] = {n: int((kind == i).sum()) for i, n in enumerate(("force", "energy", "power"))}
    out["fp64_margin_at_first_divergence_quantiles_50_90_99_100"] = [
        float(v) for v in np.quantile(at64.min(1), [0.5, 0.9, 0.99, 1.0])] if len(div) else []
    colmin = M32.min(1).min(1)
    out["replay_band_sweep"] = [{"band": b, "columns_flagged": float((colmin < b).mean()),
                                 "divergent_columns_missed": int((colmin[div] >= b).sum())}
                                for b in (1e-7, 1e-6, 1e-5, 3e-5, 1e-4)]
    if tmp is not None:
        # the reference's own arithmetic on inputs moved by FP32-sized noise:
        # how far plain FP64 itself keeps the bar when every atom coordinate
        # changes in its last written digit (1e-8 A) or by the rounding of an
        # FP32 coordinate near 10 A (1e-6 A)
        from mechafm_amd import workloads as W
        out["plain64_input_sensitivity"] = []
        for jitter in (1e-8, 1e-6):
            wls = W.make_c3(os.path.join(tmp, f"jitter{jitter:g}"), jitter=jitter)
            jsims = [abi.Simulation(w.input_path, ctx) for w in wls]
            run_batch(abi, jsims, "plain64")
            out["plain64_input_sensitivity"].append(
                dict({"jitter_A": jitter}, **bar_compare([np.concatenate([s.records() for s in jsims])], [rp])))
            for s in jsims:
                s.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c3", choices=["c1", "c2", "c3", "c4", "c5"])
    ap.add_argument("--precision", default="fast64", choices=["fast32", "fast64", "plain64"],
                    help="arithmetic of the headline scan (fast64: the one that keeps every position within "
                         "1e-4 A of the plain FP64 scan)")
    ap.add_argument("--lanes", type=int, default=0)
    ap.add_argument("--waves", type=int, default=0)
    ap.add_argument("--group", type=int, default=0, help="waves per workgroup (0 = automatic)")
    ap.add_argument("--reproducible", action="store_true",
                    help="fixed summation order per position (no wide evaluation of a wave's last groups)")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--no-fast32", action="store_true", help="skip the FP32 scan reported beside a fast64 headline")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="only run a device-to-device copy of this many bytes (calibrates HBM byte counters)")
    ap.add_argument("--compensation", type=int, default=0,
                    help="FP32 sums: 0 blocks of 4-6 sites folded into FP64, 1 Kahan, 2 plain")
    ap.add_argument("--adjoint", action="store_true",
                    help="the adjoint workload (150 atoms, 128x128 px x 30 z): gradient timing and finite differences")
    ap.add_argument("--no-job", action="store_true", help="skip the job-level (scan + files) measurement")
    ap.add_argument("--flexible", action="store_true",
                    help="time the flexible scan of the reference's example job at full size instead")
    ap.add_argument("--margins", action="store_true",
                    help="decision-margin analysis of the FP32 scan against FP64 (no timing line)")
    ap.add_argument("--valu-calibration", action="store_true",
                    help="only run the VALU calibration chains (FP32 FMA, FP64 FMA, FP64 rsq + add), one launch each "
                         "after a warm-up, and print their wave instructions and times")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: there is no multi-GPU path (pixel columns would shard "
              "across GPUs with no exchange during the scan). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    if args.copy_bytes:
        import torch
        n = args.copy_bytes // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": n * 4, "launches": 3}))
        return

    if args.valu_calibration:
        from mechafm_amd import abi
        ctx = abi.Context(0)
        out = {}
        for kind, name in ((0, "fp32_fma"), (1, "fp64_fma"), (2, "fp64_rsq_add")):
            ctx.valu_calibration(kind)                     # warm-up launch
            r = ctx.valu_calibration(kind)
            # wave instructions per SIMD per nanosecond over the launch (the
            # counters give the clock: SQ_BUSY_CYCLES)
            r["wave_insts_per_simd_ns"] = r["wave_instructions"] / (r["ms"] * 1e6) / (4 * ctx.device_info()["compute_units"])
            out[name] = r
        print(json.dumps({"valu_calibration": out, "launches_per_kind": 2}), flush=True)
        return

    from mechafm_amd import abi
    tmp = tempfile.mkdtemp(prefix="mafm_bench_")
    if args.adjoint:
        try:
            adjoint_bench(tmp, args)
        finally:
            shutil.rmtree(tmp, ignore_errors=True)
        return
    if args.flexible:
        try:
            flexible_bench(tmp, args)
        finally:
            shutil.rmtree(tmp, ignore_errors=True)
        return
    try:
        t0 = time.perf_counter()
        wls = workloads(args.config, tmp)
        ctx = abi.Context(0)
        dev = ctx.device_info()
        sims = [abi.Simulation(w.input_path, ctx) for w in wls]
        info = sims[0].info()
        shape = (info["nx"], info["ny"], info["nz"])
        positions = shape[0] * shape[1] * shape[2] * len(sims)
        print(f"# workload {args.config}: {len(sims)} systems, {positions} positions, set-up "
              f"{time.perf_counter() - t0:.1f} s on {dev['name']} ({dev['compute_units']} CUs)",
              file=sys.stderr, flush=True)
        if args.margins:
            print(json.dumps({"margins_" + args.config: margins_report(
                abi, sims, ctx, positions, shape[2], tmp if args.config == "c3" else None)}),
                  flush=True)
            return

        options = abi.SCAN_REPRODUCIBLE if args.reproducible else 0
        head = timed_scan(abi, sims, args.precision, args, dev, options, positions, args.precision)
        st = head["stats"]
        out = {
            "metric": "relaxed tip positions/s at 1e-4 Å error; % FP32 VALU peak vs plain FP64",
            "value": head["value"], "unit": "positions/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": head["ms_per_step"], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": {"fast32": "f32"}.get(args.precision, "f64"), "data": "synthetic",
            "config": {"workload": {"c3": "64 seeded planar molecules, N~U[20,400], LJ+Coulomb, "
                                          "64x64 px x 40 z each, FIRE+midpoint, one batched launch"}.get(
                args.config, args.config), "systems": len(sims), "positions": positions,
                "precision": args.precision, "lanes_per_pixel": st["lanes_per_pixel"],
                "waves_per_simd": st["waves_per_simd"], "workgroup_threads": st["block_threads"],
                "total_steps": st["total_steps"],
                "group_utilisation": st["active_group_trips"] / max(1, st["group_trips"]),
                "wide_trip_fraction": st["wide_trips"] * 64 / st["lanes_per_pixel"] / max(1, st["group_trips"]),
                "unconverged": st["unconverged"]},
            "roofline": roofline_of(sims, head, args.precision, info),
        }
        if args.precision == "fast32":
            out["config"]["compensation"] = {1: "kahan", 2: "none"}.get(args.compensation, "block-fp64")
        traffic = os.path.join(ROOT, "profiles", f"traffic_scan_{args.config}_{args.precision}.json")
        if os.path.exists(traffic):
            out["roofline"]["traffic"] = json.load(open(traffic)).get("hbm_bytes_per_launch")
        f32 = None
        if args.precision == "fast64" and not args.no_fast32:
            # the FP32 scan beside it: faster, but off the bar on a fraction of positions
            f32 = timed_scan(abi, sims, "fast32", args, dev, options, positions, "fast32")
            out["fast32"] = {"value": f32["value"], "ms_per_step": f32["ms_per_step"],
                             "lanes_per_pixel": f32["stats"]["lanes_per_pixel"],
                             "waves_per_simd": f32["stats"]["waves_per_simd"],
                             "roofline": roofline_of(sims, f32, "fast32", info)}
            t32 = os.path.join(ROOT, "profiles", f"traffic_scan_{args.config}_fast32.json")
            if os.path.exists(t32):
                out["fast32"]["roofline"]["traffic"] = json.load(open(t32)).get("hbm_bytes_per_launch")
        if not args.no_plain and args.precision != "plain64":
            t0 = time.perf_counter()
            pst = run_batch(abi, sims, "plain64")
            pwall = time.perf_counter() - t0
            prec = [s.records() for s in sims]
            p_s = pst["kernel_ms"] / 1e3
            out["plain_fp64"] = dict({"value": positions / p_s, "kernel_ms": pst["kernel_ms"], "wall_s": pwall,
                                      "speedup": p_s / head["kernel_s"]}, **bar_compare(head["recs"], prec))
            if f32 is not None:
                out["fast32"].update({"speedup_vs_plain_fp64": p_s / f32["kernel_s"]},
                                     **bar_compare(f32["recs"], prec))
        if not args.no_job and args.config == "c3":
            st = head["stats"]
            launch = dict(lanes_per_pixel=st["lanes_per_pixel"], waves_per_simd=st["waves_per_simd"],
                          compensation=args.compensation, waves_per_group=args.group, options=options)
            out["job"] = job_bench(abi, ctx, args, tmp, args.precision, launch, head["kernel_s"])
        if not args.no_cpu:
            out["cpu_baseline"] = cpu_baseline(args.config, tmp)
        print(json.dumps(out), flush=True)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


if __name__ == "__main__":
    main()
