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


def _profile_lines(name):
    try:
        return open(os.path.join(HERE, "profiles", name)).read().splitlines()
    except OSError:
        return []


def measured_traffic(what):
    """HBM bytes per launch (what = "k_eval") or per tree build (what =
    "tree") from the rocprofv3 FETCH_SIZE / WRITE_SIZE runs summarised in
    profiles/ (scaled by the same counters' readings for a copy kernel of known
    size); None if absent."""
    key = "%s traffic_bytes=" % what
    for line in _profile_lines("laplace_1M_fp64_traffic.txt"):
        if key in line:
            return float(line.split(key)[1].split()[0])
    return None


def profiled_avg_us(kernel_prefix):
    """Average duration of a kernel in the committed rocprofv3 summary of this
    workload (profiles/laplace_1M_fp64_kernels.txt): over the evented steps,
    whose kernel times this script measures, when the summary gives it, else
    over every call; None if absent."""
    key = "evented_mean_us="
    for line in _profile_lines("laplace_1M_fp64_kernels.txt"):
        if key in line and line.strip().lstrip("# ").startswith(kernel_prefix):
            return float(line.split(key)[1].split()[0])
    for line in _profile_lines("laplace_1M_fp64_kernels.txt"):
        if line.startswith(kernel_prefix):
            parts = line[len(kernel_prefix):].split()
            try:   # columns: calls, total_us, avg_us, min_us, max_us
                return float(parts[-3])
            except (IndexError, ValueError):
                return None
    return None


# extra workloads of the default run: key -> (function, steps, warmup)
EXTRA_WORKLOADS = {
    "clustered": ("clustered_workload", 5, 2),
    "adjoint": ("adjoint_workload", 5, 2),
    "stokes": ("stokes_workload", 3, 1),
    "stokes_double_layer": ("stokes_dipole_workload", 3, 1),
    "bbfmm": ("bbfmm_workload", 5, 2),
}


def progress(what):
    print("bench: %s" % what, file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the extra workloads (clustered, adjoint, Stokeslet, Stokes "
                         "double layer, black-box)")
    ap.add_argument("--only", choices=["fp64", "mixed"] + sorted(EXTRA_WORKLOADS), default=None,
                    help="time one form or one extra workload only (for profiling runs)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run only a device-to-device copy of 16777216 doubles (the known-size "
                         "kernel that scales the FETCH_SIZE / WRITE_SIZE counters)")
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


# README.md:27-33 (uniform sources, grid targets; thousands of points per
# second per core on a 4-core MacBook Pro): form multipoles, target evaluation
PUBLISHED_KPTS_PER_CORE = {"form": 947.0, "target_eval": 414.0}


def host_cpu():
    """Model name, logical CPUs this process may run on, and the physical cores
    among them (from /proc/cpuinfo)."""
    model, phys, cur = None, set(), {}
    allowed = set(os.sched_getaffinity(0))
    try:
        for line in open("/proc/cpuinfo"):
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k == "processor":
                cur = {"proc": int(v)}
            elif k == "model name" and model is None:
                model = v
            elif k in ("physical id", "core id"):
                cur[k] = v
                if "physical id" in cur and "core id" in cur and cur["proc"] in allowed:
                    phys.add((cur["physical id"], cur["core id"]))
    except OSError:
        pass
    return model, len(allowed), len(phys) or len(allowed)


def cpu_baseline(min_seconds=10.0, max_reps=20):
    """Time the oracle (NumPy restatement of the reference kifmm2d) on the
    host's cores: the whole step on the full workload -- tree (serial, as the
    reference's top-level partition), build_expansions (S2M over leaves on
    forked workers, dense passes in threaded BLAS) and target_evaluation on
    all 1M targets (forked workers over target chunks) -- repeated until at
    least min_seconds have run.  Workers = the CPUs this process may use,
    capped by OMP_NUM_THREADS."""
    model, logical, physical = host_cpu()
    workers = max(1, min(logical, int(os.environ.get("OMP_NUM_THREADS", logical))))
    code = r"""
import json, sys, time
import numpy as np
sys.path.insert(0, %r)
from oracle import kernels, kifmm2d, parallel
rng = np.random.default_rng(%d)
n = %d
x = rng.random(n); y = rng.random(n); tau = rng.standard_normal(n)
g = np.linspace(0.0, 1.0, %d)
tx, ty = np.meshgrid(g, g, indexing="ij"); tx = tx.ravel(); ty = ty.ravel()
for lev in range(12):
    kifmm2d.get_precomputation(kernels.laplace, %d, 2.0 ** -lev)
acc = dict(tree=0.0, expansions=0.0, evaluation=0.0)
reps = 0
while reps < %d and (reps == 0 or sum(acc.values()) < %r):
    t0 = time.perf_counter()
    f = parallel.ParallelFMM(x, y, kernels.laplace, %d, %d, bbox=[0, 1, 0, 1], workers=%d)
    t1 = time.perf_counter()
    f.build_expansions(tau)
    t2 = time.perf_counter()
    u = f.target_evaluation(tx, ty)
    t3 = time.perf_counter()
    acc["tree"] += t1 - t0; acc["expansions"] += t2 - t1; acc["evaluation"] += t3 - t2
    reps += 1
    del f, u
print(json.dumps(dict(acc, reps=reps, nt=tx.size)))
""" % (HERE, SEED, N_SRC, GRID, P, max_reps, float(min_seconds), NCUT, P, workers)
    env = dict(os.environ, OMP_NUM_THREADS=str(workers), OPENBLAS_NUM_THREADS=str(workers),
               MKL_NUM_THREADS=str(workers))
    res = subprocess.run([sys.executable, "-c", code], env=env, stdout=subprocess.PIPE,
                         stderr=subprocess.PIPE, text=True, timeout=900)
    if res.returncode != 0:
        return dict(error=res.stderr[-500:])
    r = json.loads(res.stdout.strip().splitlines()[-1])
    reps = r["reps"]
    total = (r["tree"] + r["expansions"] + r["evaluation"]) / reps
    rate_pub = 1.0 / (1.0 / PUBLISHED_KPTS_PER_CORE["form"]
                      + 1.0 / PUBLISHED_KPTS_PER_CORE["target_eval"]) * workers / 1e3
    return dict(
        value=round(r["nt"] / total / 1e6, 4), unit="Mpts/s", cores=workers, kind="port",
        cpu_model=model, logical_cpus=logical, physical_cores=physical,
        sample=("NumPy oracle (CPU restatement of the reference kifmm2d) on %d worker "
                "processes: the whole step on the full workload (tree serial, S2M over leaves "
                "and evaluation of all 1M grid targets on forked workers, dense passes in "
                "threaded BLAS), repeated %d times" % (workers, reps)),
        seconds=dict(tree=round(r["tree"] / reps, 4), expansions=round(r["expansions"] / reps, 4),
                     evaluation=round(r["evaluation"] / reps, 4), reps=reps),
        published_reference_scaled=dict(
            value=round(rate_pub, 4), unit="Mpts/s",
            note="README.md:27-33 per-core rates (uniform sources, grid targets: form 947 and "
                 "target evaluation 414 kpts/s/core, tree excluded) x %d cores" % workers))


def _timed(torch, fn, steps, warmup):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1000.0 / steps


def adjoint_workload(torch, steps, warmup):
    """Differentiable modified Helmholtz (k = 20): 500,000 sources uniform on
    the circle of radius 0.4 about (0.5, 0.5), N(0,1) charges, 200,000 uniform
    targets, loss = ||u(t) - u_true||^2 with u_true from a second charge
    vector; forward vs forward + adjoint (d/dcharges, d/dsource positions)."""
    from flexmm_amd import kernels
    from flexmm_amd.kifmm2d import fmm_evaluate
    dev = torch.device("cuda", 0)
    rng = np.random.default_rng(4)
    n, nt = 500_000, 200_000
    th = rng.uniform(0, 2 * np.pi, n)
    T = lambda a: torch.tensor(a, dtype=torch.float64, device=dev)
    X, Y = T(0.5 + 0.4 * np.cos(th)), T(0.5 + 0.4 * np.sin(th))
    TAU, TAU2 = T(rng.standard_normal(n)), T(rng.standard_normal(n))
    tg = rng.random((nt, 2))
    TX, TY = T(tg[:, 0]), T(tg[:, 1])
    ker, _, _ = kernels.modified_helmholtz(20.0)
    box = [0.0, 1.0, 0.0, 1.0]
    with torch.no_grad():
        utrue = fmm_evaluate(X, Y, TAU2, TX, TY, ker, 100, 60, bbox=box)

    def fwd():
        with torch.no_grad():
            u = fmm_evaluate(X, Y, TAU, TX, TY, ker, 100, 60, bbox=box)
            return ((u - utrue) ** 2).sum()

    grads = {}

    def fwd_bwd():
        Xg, Yg, Tg = X.clone().requires_grad_(True), Y.clone().requires_grad_(True), TAU.clone().requires_grad_(True)
        u = fmm_evaluate(Xg, Yg, Tg, TX, TY, ker, 100, 60, bbox=box)
        loss = ((u - utrue) ** 2).sum()
        loss.backward()
        grads.update(u=u.detach(), tau=Tg.grad, x=Xg.grad)

    ms_f = _timed(torch, fwd, steps, warmup)
    ms_fb = _timed(torch, fwd_bwd, steps, warmup)
    # FP64 direct-sum check of d/dcharges on 2,000 sampled sources
    sel = torch.from_numpy(np.sort(rng.choice(n, 2000, replace=False))).to(dev)
    g = 2.0 * (grads["u"] - utrue)
    ref = []
    for s0 in range(0, sel.numel(), 8):
        ss = sel[s0:s0 + 8]
        r = torch.sqrt((X[ss, None] - TX[None, :]) ** 2 + (Y[ss, None] - TY[None, :]) ** 2)
        ref.append((torch.special.modified_bessel_k0(20.0 * r) * g[None, :]).sum(1) / (2 * np.pi))
    ref = torch.cat(ref)
    err = float((grads["tau"][sel] - ref).abs().max() / ref.abs().max())
    # d/dx_s = tau_s sum_t g_t dK/dx_s, dK/dx_s = -k K1(k r)/(2 pi) (x_s - t_x)/r
    refx = []
    for s0 in range(0, sel.numel(), 8):
        ss = sel[s0:s0 + 8]
        dx = X[ss, None] - TX[None, :]
        r = torch.sqrt(dx ** 2 + (Y[ss, None] - TY[None, :]) ** 2)
        dk = -20.0 * torch.special.modified_bessel_k1(20.0 * r) * dx / r / (2 * np.pi)
        refx.append(TAU[ss] * (dk * g[None, :]).sum(1))
    refx = torch.cat(refx)
    errx = float((grads["x"][sel] - refx).abs().max() / refx.abs().max())
    return {"workload": "modified Helmholtz K0(20 r)/(2 pi), 500K sources on a circle, 200K "
                        "uniform targets, Nequiv 60, Ncutoff 100, loss ||u - u_true||^2",
            "forward_ms": round(ms_f, 4), "forward_adjoint_ms": round(ms_fb, 4),
            "adjoint_over_forward": round(ms_fb / ms_f, 3),
            "forward_Mpts_s": round(nt / ms_f / 1e3, 3),
            "dcharges_error_vs_direct": err,
            "dpositions_error_vs_direct": errx,
            "error_sample": "d loss/d charges and d loss/d x at 2000 sampled sources vs FP64 "
                            "direct sums"}


def clustered_workload(torch, steps, warmup):
    """Adaptive tree: 1,000,000 sources in 50 isotropic Gaussian clusters
    (centres uniform in [0.1, 0.9]^2, sigma 0.005, clipped to the unit
    square), uniform(-1, 1) charges, the 1000 x 1000 grid targets, Nequiv 60,
    Ncutoff 100; FP64 and mixed forms, errors vs FP64 direct sums at 2,000
    grid targets."""
    from flexmm_amd import kernels
    from flexmm_amd.kifmm2d import FMM
    dev = torch.device("cuda", 0)
    rng = np.random.default_rng(3)
    c = rng.uniform(0.1, 0.9, (50, 2))
    pts = np.clip(c[np.repeat(np.arange(50), 20_000)]
                  + 0.005 * rng.standard_normal((1_000_000, 2)), 0.0, 1.0)
    tau = rng.uniform(-1.0, 1.0, 1_000_000)
    g = np.linspace(0.0, 1.0, GRID)
    tx, ty = np.meshgrid(g, g, indexing="ij")
    T = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device=dev)
    X, Y, TAU, TX, TY = T(pts[:, 0]), T(pts[:, 1]), T(tau), T(tx.ravel()), T(ty.ravel())
    box = [0.0, 1.0, 0.0, 1.0]
    sel = torch.from_numpy(np.sort(rng.choice(TX.numel(), 2000, replace=False))).to(dev)
    ref = direct_sample(torch, X, Y, TAU, TX[sel], TY[sel])
    res = {"workload": "Laplace KIFMM, 1M sources in 50 Gaussian clusters (sigma 0.005), "
                       "uniform(-1,1) charges, 1M grid targets, Nequiv 60, Ncutoff 100"}
    us = {}
    for prec in ("fp64", "mixed"):
        out = {}

        def step():
            f = FMM(X, Y, kernels.laplace, NCUT, P, bbox=box, precision=prec)
            f.build_expansions(TAU)
            out["u"] = f.target_evaluation(TX, TY)[0]
            out["levels"] = int(f.tree.levels)
            f.close()

        ms = _timed(torch, step, steps, warmup)
        u = out["u"]
        us[prec] = u
        res["tree_levels"] = out["levels"]
        res[prec] = {"ms_per_step": round(ms, 4),
                     "value": round(TX.numel() / ms / 1e3, 3), "unit": "Mpts/s",
                     "error_vs_direct": float((u[sel] - ref).abs().max() / ref.abs().max())}
    res["mixed"]["error_vs_fp64"] = float((us["mixed"] - us["fp64"]).abs().max()
                                          / us["fp64"].abs().max())
    res["mixed"]["speedup_vs_fp64"] = round(res["mixed"]["value"] / res["fp64"]["value"], 3)
    return res


def stokes_workload(torch, steps, warmup):
    """Stokeslet: 2,000,000 sources in 100 Gaussian clusters (sigma 0.01,
    clipped to the unit square), N(0,1) forces, self-evaluation, Nequiv 64,
    Ncutoff 100, mixed-precision M2L; forward and forward + adjoint w.r.t.
    the forces; error vs FP64 direct sums at 4,000 sampled sources."""
    from flexmm_amd.kifmm2d import stokes_evaluate
    dev = torch.device("cuda", 0)
    rng = np.random.default_rng(5)
    nc, per = 100, 20_000
    c = rng.uniform(0.1, 0.9, (nc, 2))
    pts = np.clip(c[np.repeat(np.arange(nc), per)] + 0.01 * rng.standard_normal((nc * per, 2)), 0, 1)
    n = pts.shape[0]
    T = lambda a: torch.tensor(a, dtype=torch.float64, device=dev)
    X, Y = T(pts[:, 0]), T(pts[:, 1])
    FX, FY = T(rng.standard_normal(n)), T(rng.standard_normal(n))
    box = [0.0, 1.0, 0.0, 1.0]
    out = {}

    def fwd():
        with torch.no_grad():
            out["uv"] = stokes_evaluate(X, Y, FX, FY, X, Y, 64, 100, bbox=box,
                                        precision="mixed", check=True)

    def fwd_bwd():
        fx, fy = FX.clone().requires_grad_(True), FY.clone().requires_grad_(True)
        u, v = stokes_evaluate(X, Y, fx, fy, X, Y, 64, 100, bbox=box, precision="mixed",
                               check=True)
        (u.sum() + v.sum()).backward()

    ms_f = _timed(torch, fwd, steps, warmup)
    ms_fb = _timed(torch, fwd_bwd, steps, warmup)
    u, v = out["uv"]
    sel = torch.from_numpy(np.sort(rng.choice(n, 4000, replace=False))).to(dev)
    ru, rv = [], []
    for s0 in range(0, sel.numel(), 16):
        ss = sel[s0:s0 + 16]
        dx = X[ss, None] - X[None, :]
        dy = Y[ss, None] - Y[None, :]
        d2 = dx * dx + dy * dy
        keep = d2 > 0
        id2 = torch.where(keep, 1.0 / torch.where(keep, d2, 1.0), 0.0)
        lg = torch.where(keep, -0.5 * torch.log(torch.where(keep, d2, 1.0)), 0.0)
        ru.append(((lg + dx * dx * id2) * FX[None, :] + dx * dy * id2 * FY[None, :]).sum(1))
        rv.append((dx * dy * id2 * FX[None, :] + (lg + dy * dy * id2) * FY[None, :]).sum(1))
    ru = torch.cat(ru) / (4 * np.pi)
    rv = torch.cat(rv) / (4 * np.pi)
    err = max(float((u[sel] - ru).abs().max() / ru.abs().max()),
              float((v[sel] - rv).abs().max() / rv.abs().max()))
    return {"workload": "Stokeslet, 2M sources in 100 Gaussian clusters (sigma 0.01), N(0,1) "
                        "forces, self-evaluation, Nequiv 64, Ncutoff 100, mixed-precision M2L",
            "forward_ms": round(ms_f, 4), "forward_adjoint_ms": round(ms_fb, 4),
            "forward_Mpts_s": round(n / ms_f / 1e3, 3),
            "velocity_error_vs_direct": err,
            "error_sample": "max over u, v of max|err|/max|ref| at 4000 sampled sources"}

def bbfmm_workload(torch, steps, warmup):
    """Black-box (Chebyshev) FMM on the headline shape: 1,000,000 uniform
    sources, N(0,1) charges, the 1000 x 1000 grid targets, order 8 and
    Ncutoff 50 (the reference's BB_FMM defaults, flexmm2d/bbfmm.py:201); FP64
    and mixed forms; errors vs FP64 direct sums at 2,000 grid targets."""
    from flexmm_amd import kernels
    from flexmm_amd.flexmm2d import BB_FMM
    dev = torch.device("cuda", 0)
    x, y, tau, tx, ty = workload()
    T = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device=dev)
    X, Y, TAU, TX, TY = T(x), T(y), T(tau), T(tx), T(ty)
    box = [0.0, 1.0, 0.0, 1.0]
    rng = np.random.default_rng(7)
    sel = torch.from_numpy(np.sort(rng.choice(TX.numel(), 2000, replace=False))).to(dev)
    ref = direct_sample(torch, X, Y, TAU, TX[sel], TY[sel])
    res = {"workload": "Laplace black-box FMM, 1M uniform sources, N(0,1) charges, 1M grid "
                       "targets, Chebyshev order 8, Ncutoff 50"}
    us = {}
    for prec in ("fp64", "mixed"):
        out = {}

        def step():
            f = BB_FMM(X, Y, kernels.laplace, Ncutoff=50, p=8, bbox=box, precision=prec)
            f.build_expansions(TAU)
            out["u"] = f.target_evaluation(TX, TY)
            out["levels"] = int(f.tree.levels)
            f.close()

        ms = _timed(torch, step, steps, warmup)
        us[prec] = out["u"]
        res["tree_levels"] = out["levels"]
        res[prec] = {"ms_per_step": round(ms, 4),
                     "value": round(TX.numel() / ms / 1e3, 3), "unit": "Mpts/s",
                     "error_vs_direct": float((out["u"][sel] - ref).abs().max()
                                              / ref.abs().max())}
    res["mixed"]["error_vs_fp64"] = float((us["mixed"] - us["fp64"]).abs().max()
                                          / us["fp64"].abs().max())
    return res


def stokes_dipole_workload(torch, steps, warmup):
    """Stokeslets plus stresslets: 2,000,000 sources in 100 Gaussian clusters
    (sigma 0.01), N(0,1) forces and double-layer densities, random unit
    normals, self-evaluation of velocity and pressure, Nequiv 64, Ncutoff
    100, FP64; the single-layer forward on the same points beside it."""
    from flexmm_amd.kifmm2d import StokesFMM
    dev = torch.device("cuda", 0)
    rng = np.random.default_rng(9)
    nc, per = 100, 20_000
    c = rng.uniform(0.1, 0.9, (nc, 2))
    pts = np.clip(c[np.repeat(np.arange(nc), per)] + 0.01 * rng.standard_normal((nc * per, 2)), 0, 1)
    n = pts.shape[0]
    T = lambda a: torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device=dev)
    X, Y = T(pts[:, 0]), T(pts[:, 1])
    FX, FY, GX, GY = [T(rng.standard_normal(n)) for _ in range(4)]
    th = rng.uniform(0, 2 * np.pi, n)
    NX, NY = T(np.cos(th)), T(np.sin(th))
    box = [0.0, 1.0, 0.0, 1.0]

    def single():
        f = StokesFMM(X, Y, 64, 100, bbox=box)
        f.build_expansions(FX, FY)
        f.evaluate_to_points(X, Y, True)
        f.close()

    out = {}

    def double():
        f = StokesFMM(X, Y, 64, 100, bbox=box)
        f.build_expansions_dipole(FX, FY, GX, GY, NX, NY)
        out["uvp"] = f.evaluate_to_points_dipole(X, Y, True)
        f.close()

    ms_s = _timed(torch, single, steps, warmup)
    ms_d = _timed(torch, double, steps, warmup)
    # FP64 direct sums (Stokeslet + stresslet, self-excluded) at 2,000 sources
    sel = torch.from_numpy(np.sort(rng.choice(n, 2000, replace=False))).to(dev)
    ref = [[], [], []]
    for s0 in range(0, sel.numel(), 16):
        ss = sel[s0:s0 + 16]
        dx = X[ss, None] - X[None, :]
        dy = Y[ss, None] - Y[None, :]
        d2 = dx * dx + dy * dy
        keep = d2 > 0
        id2 = torch.where(keep, 1.0 / torch.where(keep, d2, 1.0), 0.0)
        lg = torch.where(keep, -0.5 * torch.log(torch.where(keep, d2, 1.0)), 0.0)
        w = (dx * NX + dy * NY) * (dx * GX + dy * GY) * id2 * id2
        ref[0].append((((lg + dx * dx * id2) * FX + dx * dy * id2 * FY) / (4 * np.pi)
                       + w * dx / np.pi).sum(1))
        ref[1].append(((dx * dy * id2 * FX + (lg + dy * dy * id2) * FY) / (4 * np.pi)
                       + w * dy / np.pi).sum(1))
        ref[2].append(((dx * FX + dy * FY) * id2 / (2 * np.pi)
                       + (-(NX * GX + NY * GY) * id2 + 2.0 * w) / np.pi).sum(1))
    ref = [torch.cat(r) for r in ref]
    errs = [float((a[sel] - r).abs().max() / r.abs().max()) for a, r in zip(out["uvp"], ref)]
    return {"workload": "Stokeslets + stresslets, 2M sources in 100 Gaussian clusters "
                        "(sigma 0.01), self-evaluation of u, v, p, Nequiv 64, Ncutoff 100, FP64",
            "single_layer_ms": round(ms_s, 4), "single_plus_double_layer_ms": round(ms_d, 4),
            "single_plus_double_layer_Mpts_s": round(n / ms_d / 1e3, 3),
            "velocity_error_vs_direct": max(errs[:2]), "pressure_error_vs_direct": errs[2],
            "error_sample": "max|err|/max|ref| at 2000 sampled sources vs FP64 direct sums "
                            "with self-exclusion"}


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
    if args.copy_calibration:
        src = torch.ones(1 << 24, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": src.numel() * 8}))
        return
    if args.only in EXTRA_WORKLOADS:
        fn, steps, warmup = EXTRA_WORKLOADS[args.only]
        print(json.dumps({args.only: globals()[fn](torch, steps, warmup)}))
        return
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
        def step():
            f = FMM(xd, yd, kernels.laplace, NCUT, P, bbox=bbox, precision=precision)
            f.build_expansions(taud)
            u = f.target_evaluation(txd, tyd)
            return f, u

        # Every step reuses one pooled device plan (each FMM is released
        # before the next one is made).  The timed steps run without phase
        # events (an event record between two kernels idles the GPU for a
        # few microseconds); the same number of steps then runs again with
        # them, each run of a phase recording its own pair of HIP events on
        # the launch stream (torch's current stream), and phase_times
        # returns the mean over the runs since its previous call.
        import ctypes
        from flexmm_amd import _lib

        def phase_times(plan):
            out = (ctypes.c_double * len(_lib.PHASES))()
            _lib.check(plan.lib.fmm2d_phase_times(plan.h, out, len(out)))
            return {k: float(v) for k, v in zip(_lib.PHASES, out) if v >= 0}

        f, u = step()
        plan = f._plan
        f.close()
        for _ in range(args.warmup - 1):
            f, u = step()
            f.close()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            f, u = step()
            f.close()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        ms = (t1 - t0) * 1000.0 / args.steps
        _lib.check(plan.lib.fmm2d_set_profiling(plan.h, 1))
        phase_times(plan)
        for _ in range(args.steps):
            f, u = step()
            f.close()
        acc = phase_times(plan)
        _lib.check(plan.lib.fmm2d_set_profiling(plan.h, 0))
        f, u = step()
        work = f.work_counts()
        nodes = f.tree.n_nodes
        f.close()
        # max-norm error relative to max|u| (examples/pure_kifmm_laplace.py:149-151)
        err = float((u[0][sel] - ref).abs().max() / ref.abs().max())
        return dict(ms=ms, value=tx.size / (ms / 1000.0) / 1e6, phases=acc, work=work, err=err,
                    u=u[0], nodes=nodes)

    if args.only:
        r = measure(args.only)
        r.pop("u")
        r.pop("nodes")
        print(json.dumps(dict(precision=args.only, ms_per_step=round(r["ms"], 4),
                              value=round(r["value"], 3), error_vs_direct=r["err"],
                              phase_ms={k: round(v, 4) for k, v in r["phases"].items()})))
        return
    progress("fp64 form")
    r64 = measure("fp64")
    u64 = r64.pop("u")
    progress("mixed form")
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
    f_eval = f_p2p + f_l2t
    achieved = f_eval / (ek / 1000.0) / 1e12
    prof_us = profiled_avg_us("k_eval<64, 0, 1, false, false, false>")
    # tree build: algorithmic bytes = 40 B per source (read x, y; write sorted
    # x, y; keys and permutation) + one write of the node table (252 B per node)
    tree_bytes = 40.0 * N_SRC + 252.0 * r64["nodes"]
    tree_ms = acc.get("tree", float("nan"))

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
            "kernel_speedup_vs_fp64": {
                k: round(acc[k] / rmx["phases"][k], 3) for k in ("eval_kernel", "m2l")
                if k in acc and k in rmx["phases"] and rmx["phases"][k] > 0},
            "shared_fp64_ms": round(sum(rmx["phases"].get(k, 0.0) for k in ("tree", "upward")), 4),
            "shared_note": "the tree build and the upward pass (S2M, M2M, pseudo-inverse) run in "
                           "FP64 in both forms",
        },
        "phase_ms": {k: round(v, 4) for k, v in acc.items()},
        "work": work,
        "algorithmic_gflop": round(f_total / 1e9, 3),
        "end_to_end_tflops": round(f_total / (ms / 1000.0) / 1e12, 3),
        "roofline": {
            "kernel": "k_eval (fused L2T + P2P, FP64)",
            "bound": "valu",
            "note": "FP64 VALU issue bound (a table log per pair, about 18 VALU instructions "
                    "per pair; profiles/k_eval_1M_fp64_valu.txt); the peak is MI355X's FP64 "
                    "vector rate, equal to its FP64 matrix rate",
            "achieved": round(achieved, 3),
            "peak": FP64_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved / FP64_PEAK_TFLOPS, 4),
            "algorithmic_flop": f_eval,
            "kernel_ms": round(ek, 5),
            "kernel_ms_note": "mean of HIP-event pairs around each launch of the kernel over the "
                              "timed steps, on its launch stream",
            "profile_avg_us": prof_us,
            "frac_from_profile": (round(f_eval / (prof_us * 1e-6) / 1e12 / FP64_PEAK_TFLOPS, 4)
                                  if prof_us else None),
            "profile_note": "mean duration over the evented steps in "
                            "profiles/laplace_1M_fp64_kernels.txt (rocprofv3 --kernel-trace "
                            "--stats of bench.py --only fp64, the same steps as kernel_ms)",
            "traffic": measured_traffic("k_eval"),
            "traffic_note": "HBM bytes per launch of k_eval from rocprofv3 FETCH_SIZE/WRITE_SIZE "
                            "(profiles/laplace_1M_fp64_traffic.txt); algorithmic bytes ~6.9e7 "
                            "(targets, sources, local expansions, output)",
        },
        "roofline_tree": {
            "kernel": "tree build (every kernel of fmm2d_tree_build)",
            "bound": "hbm",
            "achieved": round(tree_bytes / (tree_ms * 1e-3) / 1e9, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(tree_bytes / (tree_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "algorithmic_bytes": tree_bytes,
            "phase_ms": round(tree_ms, 5),
            "traffic": measured_traffic("tree"),
            "note": "algorithmic bytes = 40 B per source + 252 B per node; traffic = FETCH_SIZE "
                    "+ WRITE_SIZE of all tree kernels per build (profiles/"
                    "laplace_1M_fp64_traffic.txt); the build is a chain of short kernels with "
                    "host read-backs, bound by launch and synchronisation latency",
        },
    }
    if not args.no_extra:
        for key, (fn, steps, warmup) in EXTRA_WORKLOADS.items():
            progress(key)
            out[key] = globals()[fn](torch, steps, warmup)
    if not args.no_cpu_baseline:
        progress("CPU baseline")
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
