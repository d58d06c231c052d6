"""FMM parity: device potentials against the oracle restatement (FP64, 1e-12
relative in max norm, the error metric of examples/pure_kifmm_laplace.py:149-151)
and against exact FP64 direct sums."""
import numpy as np
import pytest

from oracle import direct as odirect
from oracle import kernels as okern
from oracle import kifmm2d as okifmm
from tests import helpers as H

pytestmark = pytest.mark.gpu

PARITY_TOL = 1e-12


def _fmm(x, y, ncut, p, bbox=None, m2l="svd"):
    from flexmm_amd import kernels
    from flexmm_amd.kifmm2d import FMM, the_help
    the_help.clear()
    return FMM(x, y, kernels.laplace, ncut, p, bbox=bbox, m2l=m2l)


def _oracle(x, y, ncut, p, bbox=None):
    return okifmm.FMM(x, y, okern.laplace, ncut, p, bbox=bbox)


@pytest.mark.parametrize("p", [32, 60])
def test_self_eval_uniform(p):
    x, y, tau = H.uniform(30000, 1)
    print("self-eval uniform p", p, flush=True)
    f = _fmm(x, y, 64, p)
    f.build_expansions(tau)
    u = f.source_evaluation(x, y)[0]
    o = _oracle(x, y, 64, p)
    o.build_expansions(tau)
    uo = o.source_evaluation(x, y)[0]
    err = H.rel_max_err(u, uo)
    print("p", p, "err vs oracle %.3e" % err)
    assert err < PARITY_TOL


def test_target_eval_clustered_grid():
    x, y, tau = H.clustered(30000, 5, 0.01, 3)
    tx, ty = H.grid(120)
    print("clustered grid", flush=True)
    f = _fmm(x, y, 50, 60, bbox=[0, 1, 0, 1])
    f.build_expansions(tau)
    u = f.target_evaluation(tx, ty)[0]
    o = _oracle(x, y, 50, 60, bbox=[0, 1, 0, 1])
    o.build_expansions(tau)
    uo = o.target_evaluation(tx, ty)[0]
    err = H.rel_max_err(u, uo)
    print("err vs oracle %.3e" % err)
    assert err < PARITY_TOL
    ref = odirect.direct(okern.laplace, x, y, tau, tx[::7], ty[::7])[0]
    assert H.rel_max_err(u[::7], ref) < 1e-12


def test_gradient_evaluator():
    from flexmm_amd import kernels
    x, y, tau = H.uniform(20000, 8)
    tx, ty, _ = H.uniform(5000, 9)
    f = _fmm(x, y, 64, 60)
    f.register_evaluator(kernels.laplace_gradient, kernels.laplace_gradient, "gradient", 3)
    f.build_expansions(tau)
    g = f.evaluate_to_points(tx, ty, "gradient", False)
    o = _oracle(x, y, 64, 60)
    o.register_evaluator(okern.laplace_gradient, "gradient", 3)
    o.build_expansions(tau)
    go = o.evaluate_to_points(tx, ty, "gradient", False)
    for k in range(3):
        err = H.rel_max_err(g[k], go[k])
        print("output", k, "err %.3e" % err)
        assert err < PARITY_TOL
    ref = odirect.direct(okern.laplace_gradient, x, y, tau, tx, ty)
    for k in range(3):
        assert H.rel_max_err(g[k], ref[k]) < 1e-11


def test_empty_target_set_and_single_source():
    """No targets give empty outputs (torch and numpy inputs); one source is a
    root leaf evaluated directly."""
    import torch
    x, y, tau = H.uniform(2000, 21)
    f = _fmm(x, y, 30, 32)
    f.build_expansions(tau)
    e = np.zeros(0)
    assert f.target_evaluation(e, e)[0].shape == (0,)
    et = torch.zeros(0, dtype=torch.float64, device="cuda")
    assert f.target_evaluation(et, et)[0].numel() == 0
    g = _fmm(x[:1], y[:1], 30, 32, bbox=[0, 1, 0, 1])
    g.build_expansions(tau[:1])
    u = g.target_evaluation(x[1:9], y[1:9])[0]
    ref = odirect.direct(okern.laplace, x[:1], y[:1], tau[:1], x[1:9], y[1:9])[0]
    assert H.rel_max_err(u, ref) < 1e-14


@pytest.mark.parametrize("n,ncut", [(50, 64), (100, 30), (300, 40)])
def test_shallow_trees_are_direct(n, ncut):
    x, y, tau = H.uniform(n, n)
    f = _fmm(x, y, ncut, 32)
    f.build_expansions(tau)
    u = f.source_evaluation(x, y)[0]
    o = _oracle(x, y, ncut, 32)
    o.build_expansions(tau)
    uo = o.source_evaluation(x, y)[0]
    assert H.rel_max_err(u, uo) < PARITY_TOL


def test_direct_sum_p64():
    x, y, tau = H.uniform(40000, 12)
    f = _fmm(x, y, 64, 64)
    f.build_expansions(tau)
    u = f.source_evaluation(x, y)[0]
    sel = np.arange(0, x.size, 37)
    ref = odirect.direct(okern.laplace, x, y, tau, x[sel], y[sel], check=True)[0]
    err = H.rel_max_err(u[sel], ref)
    print("err vs direct %.3e" % err)
    assert err < 5e-13


def test_reference_cpu_case_100k():
    """The reference's own CPU-runnable case: 100,000 uniform sources, N(0,1)
    charges, self-evaluation with self-exclusion, 32 equivalent points, leaf
    capacity 64 -- against the oracle (parity) and FP64 direct sums (the
    truncation error of p = 32)."""
    rng = np.random.default_rng(100)
    x, y = rng.random(100_000), rng.random(100_000)
    tau = rng.standard_normal(100_000)
    print("100k self-eval p 32 Ncutoff 64", flush=True)
    f = _fmm(x, y, 64, 32)
    f.build_expansions(tau)
    u = f.source_evaluation(x, y)[0]
    o = _oracle(x, y, 64, 32)
    o.build_expansions(tau)
    uo = o.source_evaluation(x, y)[0]
    err = H.rel_max_err(u, uo)
    sel = np.arange(0, x.size, 97)
    ref = odirect.direct(okern.laplace, x, y, tau, x[sel], y[sel], check=True)[0]
    errd = H.rel_max_err(u[sel], ref)
    print("levels", f.tree.levels, "vs oracle %.3e, vs direct %.3e" % (err, errd))
    assert err < PARITY_TOL
    assert errd < 3e-8   # p = 32 truncation: 7.2e-9 measured


@pytest.mark.parametrize("m2l", ["svd", "dense"])
@pytest.mark.parametrize("case", ["uniform", "clustered"])
def test_m2l_compressed_and_dense_forms(m2l, case):
    """Both FP64 M2L forms against the oracle (which applies the reference's
    compressed form, helpers.py:17-45): slots of rank r with 2 r < p as
    M2 (M1 q), the others dense; and the two forms against each other."""
    if case == "uniform":
        x, y, tau = H.uniform(60000, 301)
    else:
        x, y, tau = H.clustered(60000, 6, 0.01, 302)
    tx, ty = H.grid(150)
    print("m2l", m2l, case, flush=True)
    f = _fmm(x, y, 60, 60, bbox=[0, 1, 0, 1], m2l=m2l)
    f.build_expansions(tau)
    u = f.target_evaluation(tx, ty)[0]
    o = _oracle(x, y, 60, 60, bbox=[0, 1, 0, 1])
    o.build_expansions(tau)
    uo = o.target_evaluation(tx, ty)[0]
    err = H.rel_max_err(u, uo)
    other = _fmm(x, y, 60, 60, bbox=[0, 1, 0, 1], m2l="dense" if m2l == "svd" else "svd")
    other.build_expansions(tau)
    e2 = H.rel_max_err(u, other.target_evaluation(tx, ty)[0])
    print(m2l, case, "levels", f.tree.levels, "vs oracle %.3e, vs the other form %.3e" % (err, e2))
    assert err < PARITY_TOL and e2 < PARITY_TOL
