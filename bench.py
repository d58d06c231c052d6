#!/usr/bin/env python3
"""Benchmark of the 2D Laplace KIFMM on one MI355X.

Workload (the reference README's headline setup, README.md:9-33): 1,000,000
sources uniform in [0,1]^2 with N(0,1) charges, 1,000,000 targets on a
1000 x 1000 grid over [0,1]^2, kernel -log|x-y|/(2 pi), Nequiv = 60,
Ncutoff = 100, bbox = [0,1,0,1]; seeds as listed in "config".

One step = what the reference's harness times for a fresh solve
(examples/pure_kifmm_laplace.py:128-146): construct the FMM (tree build on
the device; per-width operators are cached after the first call, the
reference's warm-up), build_expansions(tau), target_evaluation(tx, ty).
Inputs are resident in HBM before the timed region.

Prints ONE JSON line.  `--gpus N` with N > 1 exits: the build targets one GPU.
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
N_SRC = 1_000_000
GRID = 1000
P = 60
NCUT = 100
SEED = 2
FP64_PEAK_TFLOPS = 78.6    # MI355X FP64 (vector = matrix)
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-targets", type=int, default=200_000)
    ap.add_argument("--only", choices=["fp64", "mixed"], default=None,
                    help="time one form only and print its phase times (for profiling runs)")
    return ap.parse_args()


def workload():
    rng = np.random.default_rng(SEED)
    x = rng.random(N_SRC)
    y = rng.random(N_SRC)
    tau = rng.standard_normal(N_SRC)
    g = np.linspace(0.0, 1.0, GRID)
    tx, ty = np.meshgrid(g, g, indexing="ij")
    return x, y, tau, tx.ravel(), ty.ravel()


def direct_sample(torch, xd, yd, taud, tx, ty, chunk=32):
    """FP64 direct sum at a few targets with torch (reference for the error)."""
    out = []
    for s in range(0, tx.numel(), chunk):
        dx = tx[s:s + chunk, None] - xd[None, :]
        dy = ty[s:s + chunk, None] - yd[None, :]
        v = -torch.log(dx * dx + dy * dy) * (1.0 / (4 * np.pi))
        out.append((v * taud[None, :]).sum(dim=1))
    return torch.cat(out)


def cpu_baseline(sample_targets):
    """Time the NumPy oracle (CPU restatement of the reference algorithm) on
    one core: full tree + expansions, evaluation on a target sample."""
    code = r"""
import json, sys, time
import numpy as np
sys.path.insert(0, %r)
from oracle import kifmm2d, kernels
rng = np.random.default_rng(%d)
n = %d
x = rng.random(n); y = rng.random(n); tau = rng.standard_normal(n)
g = np.linspace(0.0, 1.0, %d)
tx, ty = np.meshgrid(g, g, indexing="ij"); tx = tx.ravel(); ty = ty.ravel()
m = %d
for lev in range(12):
    kifmm2d.get_precomputation(kernels.laplace, %d, 2.0 ** -lev)
t0 = time.perf_counter()
f = kifmm2d.FMM(x, y, kernels.laplace, %d, %d, bbox=[0, 1, 0, 1])
t1 = time.perf_counter()
f.build_expansions(tau)
t2 = time.perf_counter()
u = f.target_evaluation(tx[:m], ty[:m])
t3 = time.perf_counter()
print(json.dumps(dict(tree=t1 - t0, expansions=t2 - t1, eval_sample=t3 - t2, sample=m, nt=tx.size)))
""" % (HERE, SEED, N_SRC, GRID, sample_targets, P, NCUT, P)
    env = dict(os.environ, OMP_NUM_THREADS="1", OPENBLAS_NUM_THREADS="1", MKL_NUM_THREADS="1")
    res = subprocess.run([sys.executable, "-c", code], env=env, stdout=subprocess.PIPE,
                         stderr=subprocess.PIPE, text=True, timeout=600)
    if res.returncode != 0:
        return dict(error=res.stderr[-500:])
    r = json.loads(res.stdout.strip().splitlines()[-1])
    total = r["tree"] + r["expansions"] + r["eval_sample"] * r["nt"] / r["sample"]
    return dict(
        value=round(r["nt"] / total / 1e6, 4), unit="Mpts/s", cores=1, kind="port",
        sample=("NumPy oracle (CPU restatement of the reference kifmm2d) on 1 core: full tree + "
                "build_expansions for the 1M sources, target_evaluation on the first %d of the "
                "1M grid targets, evaluation time scaled to 1M" % r["sample"]),
        seconds=dict(tree=round(r["tree"], 3), expansions=round(r["expansions"], 3),
                     eval_sample=round(r["eval_sample"], 3)))


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py measures one GPU only (this build targets a single MI355X); "
              "got --gpus %d" % args.gpus, file=sys.stderr)
        sys.exit(2)
    import torch
    from flexmm_amd import kernels
    from flexmm_amd.kifmm2d import FMM

    dev = torch.device("cuda", 0)
    x, y, tau, tx, ty = workload()
    xd = torch.from_numpy(x).to(dev)
    yd = torch.from_numpy(y).to(dev)
    taud = torch.from_numpy(tau).to(dev)
    txd = torch.from_numpy(tx).to(dev)
    tyd = torch.from_numpy(ty).to(dev)
    bbox = [0.0, 1.0, 0.0, 1.0]

    rng = np.random.default_rng(7)
    sel = torch.from_numpy(np.sort(rng.choice(tx.size, 2000, replace=False))).to(dev)
    ref = direct_sample(torch, xd, yd, taud, txd[sel], tyd[sel])

    def measure(precision):
        def step(profile=False):
            f = FMM(xd, yd, kernels.laplace, NCUT, P, bbox=bbox, precision=precision)
            if profile:
                f.set_profiling(True)
                f.build_tree()
            f.build_expansions(taud)
            u = f.target_evaluation(txd, tyd)
            return f, u

        for _ in range(args.warmup):
            f, u = step()
            del f
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            f, u = step()
            del f
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        ms = (t1 - t0) * 1000.0 / args.steps
        # Phase and kernel durations (HIP events on the launch stream),
        # averaged over separate profiled steps.
        acc = {}
        nprof = max(3, min(args.steps, 10))
        for _ in range(nprof):
            f, u = step(profile=True)
            for k, v in f.phase_times().items():
                acc[k] = acc.get(k, 0.0) + v / nprof
            work = f.work_counts()
            del f
        f, u = step()
        # max-norm error relative to max|u| (examples/pure_kifmm_laplace.py:149-151)
        err = float((u[0][sel] - ref).abs().max() / ref.abs().max())
        return dict(ms=ms, value=tx.size / (ms / 1000.0) / 1e6, phases=acc, work=work, err=err,
                    u=u[0])

    if args.only:
        r = measure(args.only)
        r.pop("u")
        print(json.dumps(dict(precision=args.only, ms_per_step=round(r["ms"], 4),
                              value=round(r["value"], 3), error_vs_direct=r["err"],
                              phase_ms={k: round(v, 4) for k, v in r["phases"].items()})))
        return
    r64 = measure("fp64")
    u64 = r64.pop("u")
    rmx = measure("mixed")
    umx = rmx.pop("u")
    err_mx64 = float((umx - u64).abs().max() / u64.abs().max())
    del u64, umx
    ms, value, acc, work, err = r64["ms"], r64["value"], r64["phases"], r64["work"], r64["err"]

    # Algorithmic work (flop conventions: a Laplace pair is 20 flops, a GEMM 2mnk)
    pairs = work["p2p_pairs"]
    f_p2p = 20.0 * pairs
    f_l2t = 20.0 * P * work["l2t_targets"]
    f_s2m = 20.0 * P * work["s2m_points"]
    f_m2l = 2.0 * P * P * work["m2l_pairs"]
    f_pinv = 4.0 * P * P * (work["up_rows"] + work["down_boxes"])
    f_m2m = 8.0 * P * P * (work["m2m_parents"]) + 2.0 * P * P * work["l2l_children"]
    f_total = f_p2p + f_l2t + f_s2m + f_m2l + f_pinv + f_m2m
    ek = acc.get("eval_kernel", float("nan"))
    achieved = (f_p2p + f_l2t) / (ek / 1000.0) / 1e12

    out = {
        "metric": "Mpts/s, 1M-src/1M-tgt Laplace KIFMM: FP64 vs mixed-prec at stated error",
        "value": round(value, 3),
        "unit": "Mpts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded uniform sources, N(0,1) charges, 1000x1000 grid targets)",
        "config": {"workload": "Laplace KIFMM, 1M uniform sources -> 1M grid targets, "
                               "Nequiv 60, Ncutoff 100, bbox [0,1]^2, seed %d; value = FP64 form, "
                               "'mixed' = mixed-precision form on the same workload" % SEED,
                   "n_sources": N_SRC, "n_targets": int(tx.size), "Nequiv": P, "Ncutoff": NCUT,
                   "precision": "fp64"},
        "error_vs_direct": err,
        "error_sample": "max|u-u_direct|/max|u_direct| over 2000 random grid targets",
        "mixed": {
            "value": round(rmx["value"], 3),
            "unit": "Mpts/s",
            "ms_per_step": round(rmx["ms"], 4),
            "speedup_vs_fp64": round(rmx["value"] / value, 3),
            "error_vs_direct": rmx["err"],
            "error_vs_fp64": err_mx64,
            "dtype": "f32 M2L (MFMA) + f32 leaf near field, f64 elsewhere",
            "phase_ms": {k: round(v, 4) for k, v in rmx["phases"].items()},
        },
        "phase_ms": {k: round(v, 4) for k, v in acc.items()},
        "work": work,
        "algorithmic_gflop": round(f_total / 1e9, 3),
        "end_to_end_tflops": round(f_total / (ms / 1000.0) / 1e12, 3),
        "roofline": {
            "kernel": "k_eval (fused L2T + P2P, FP64)",
            "bound": "mfma",
            "note": "FP64 VALU-bound (log per pair); MI355X FP64 vector peak = FP64 matrix peak",
            "achieved": round(achieved, 3),
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / FP64_PEAK_TFLOPS, 4),
            "traffic": None,
        },
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.cpu_sample_targets)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
