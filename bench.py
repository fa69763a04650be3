"""Throughput of the rigid probe-particle scan on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--config c3]

A step is one full scan of the workload (every pixel column, every height,
FIRE + midpoint to convergence) in one launch of the work-stealing kernel,
with the molecules' tables already resident in HBM.  Default workload: the
batched one (64 seeded planar molecules, N ~ U[20, 400], LJ + Coulomb,
64 x 64 px x 40 z each), FP32 pair terms summed in register blocks folded
into FP64, FP64 minimiser state.  The FP64 workloads (c1, c2) default to the
FP64 fast path.

Printed JSON line: metric/value (relaxed tip positions per second, whole
job), the roofline of the scan kernel (algorithmic VALU FLOPs per launch
over its HIP-event duration, against the 157.3 TFLOP/s FP32 or 78.6 TFLOP/s
FP64 vector peak; lookup bytes against 8 TB/s for a force-grid scan),
the plain FP64 GPU scan on the same workload (speed and the fraction of
positions within 1e-4 A / 1e-3 relative Fz of it), and a CPU baseline: the
reference program timed on this host on a sample of the same workload.
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

FP32_VALU_PEAK = 157.3e12      # MI355X FP32 vector, FLOP/s
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="c3", choices=["c1", "c2", "c3", "c4", "c5"])
    ap.add_argument("--precision", default=None, choices=["fast32", "fast64", "plain64"],
                    help="default: fast64 for the FP64 workloads (c1, c2), fast32 for the others")
    ap.add_argument("--lanes", type=int, default=0)
    ap.add_argument("--waves", type=int, default=0)
    ap.add_argument("--group", type=int, default=0, help="waves per workgroup (0 = automatic)")
    ap.add_argument("--reproducible", action="store_true",
                    help="fixed summation order per position (no wide evaluation of a wave's last groups)")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="only run a device-to-device copy of this many bytes (calibrates HBM byte counters)")
    ap.add_argument("--compensation", type=int, default=0,
                    help="FP32 sums: 0 blocks of 4-6 sites folded into FP64, 1 Kahan, 2 plain")
    ap.add_argument("--adjoint", action="store_true",
                    help="the adjoint workload (150 atoms, 128x128 px x 30 z): gradient timing and finite differences")
    args = ap.parse_args()
    if args.precision is None:
        args.precision = "fast64" if args.config in ("c1", "c2") else "fast32"
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

    from mechafm_amd import abi
    tmp = tempfile.mkdtemp(prefix="mafm_bench_")
    if args.adjoint:
        try:
            adjoint_bench(tmp, args)
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

        G, wps = args.lanes, args.waves
        options = abi.SCAN_REPRODUCIBLE if args.reproducible else 0
        if args.precision != "plain64" and not args.no_tune and (G == 0 or wps == 0):
            from mechafm_amd import tune
            best = tune.tune(sims, args.precision, compensation=args.compensation, options=options,
                             lanes=[G] if G else tune.DEFAULT_LANES, use_cache=False, device=dev,
                             log=lambda m: print("# tune " + m, file=sys.stderr, flush=True))
            G, wps = best["lanes_per_pixel"], best["waves_per_simd"]
        kw = dict(lanes_per_pixel=G, waves_per_simd=wps, compensation=args.compensation,
                  waves_per_group=args.group, options=options)
        for _ in range(args.warmup):
            run_batch(abi, sims, args.precision, **kw)
        kms = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            st = run_batch(abi, sims, args.precision, **kw)
            kms.append(st["kernel_ms"])
        wall = time.perf_counter() - t0
        value = positions * args.steps / wall
        recs = [s.records().copy() for s in sims]
        flops = algorithmic_flops(sims, recs)
        k_avg = float(np.mean(kms)) / 1e3
        achieved = flops / k_avg
        peak = FP32_VALU_PEAK if args.precision == "fast32" else FP64_VALU_PEAK
        out = {
            "metric": "relaxed tip positions/s at 1e-4 A error; % FP32 VALU peak vs plain FP64",
            "value": value, "unit": "positions/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "f32" if args.precision == "fast32" else "f64", "data": "synthetic",
            "config": {"workload": {"c3": "64 seeded planar molecules, N~U[20,400], LJ+Coulomb, "
                                          "64x64 px x 40 z each, FIRE+midpoint, one batched launch"}.get(
                args.config, args.config), "systems": len(sims), "positions": positions,
                "precision": args.precision, "lanes_per_pixel": st["lanes_per_pixel"],
                "waves_per_simd": st["waves_per_simd"], "workgroup_threads": st["block_threads"],
                "total_steps": st["total_steps"],
                "compensation": {1: "kahan", 2: "none"}.get(args.compensation, "block-fp64"),
                "group_utilisation": st["active_group_trips"] / max(1, st["group_trips"]),
                "wide_trip_fraction": st["wide_trips"] * 64 / st["lanes_per_pixel"] / max(1, st["group_trips"]),
                "unconverged": st["unconverged"]},
            "roofline": {"bound": "valu", "achieved": achieved / 1e12, "peak": peak / 1e12,
                         "unit": "TFLOP/s", "frac": achieved / peak, "traffic": None,
                         "kernel_ms": k_avg * 1e3, "flops_per_launch": flops},
        }
        if info["rigidgrid"]:
            # force-grid scan: no site terms, the work is the trilinear gathers
            gb = grid_lookup_bytes(sims, recs, fp64=args.precision != "fast32")
            out["roofline"] = {"bound": "hbm", "achieved": gb / k_avg / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                               "frac": gb / k_avg / HBM_PEAK, "traffic": None, "kernel_ms": k_avg * 1e3,
                               "bytes_per_launch": gb}
        traffic = os.path.join(ROOT, "profiles", f"traffic_scan_{args.config}_{args.precision}.json")
        if os.path.exists(traffic):
            out["roofline"]["traffic"] = json.load(open(traffic)).get("hbm_bytes_per_launch")
        if not args.no_plain and args.precision != "plain64":
            t0 = time.perf_counter()
            pst = run_batch(abi, sims, "plain64")
            pwall = time.perf_counter() - t0
            prec = [s.records() for s in sims]
            dr = np.concatenate([np.abs(np.stack([a["rx"] - b["rx"], a["ry"] - b["ry"], a["rz"] - b["rz"]])).max(0)
                                 for a, b in zip(recs, prec)])
            dfz = np.concatenate([np.abs(a["fz"] - b["fz"]) / np.maximum(np.abs(b["fz"]), 1e-2)
                                  for a, b in zip(recs, prec)])
            ok = (dr <= 1e-4) & (dfz <= 1e-3)
            out["plain_fp64"] = {"value": positions / (pst["kernel_ms"] / 1e3), "kernel_ms": pst["kernel_ms"],
                                 "wall_s": pwall, "speedup": (pst["kernel_ms"] / 1e3) / k_avg,
                                 "within_1e-4A_1e-3Fz": float(ok.mean()), "max_dr": float(dr.max()),
                                 "median_dr": float(np.median(dr))}
        if not args.no_cpu:
            out["cpu_baseline"] = cpu_baseline(args.config, tmp)
        print(json.dumps(out), flush=True)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


if __name__ == "__main__":
    main()
