"""Mixed-precision form: FP32 M2L on matrix cores (constant mode in FP64) and
FP32 near field for leaf colleagues.  Its potentials must stay within 1e-7
relative (max norm) of the FP64 form on the same tree and of exact FP64 direct
sums; the FP64 form itself is pinned to the oracle in test_gpu_fmm."""
import numpy as np
import pytest

from oracle import direct as odirect
from oracle import kernels as okern
from tests import helpers as H

pytestmark = pytest.mark.gpu

MIXED_TOL = 1e-7


def _pair(x, y, ncut, p, bbox=None, precision="mixed"):
    from flexmm_amd import kernels
    from flexmm_amd.kifmm2d import FMM, the_help
    the_help.clear()
    f64 = FMM(x, y, kernels.laplace, ncut, p, bbox=bbox)
    mix = FMM(x, y, kernels.laplace, ncut, p, bbox=bbox, precision=precision)
    return f64, mix


@pytest.mark.parametrize("p,precision", [(32, "mixed"), (60, "mixed"), (60, "mixed_m2l")])
def test_mixed_self_eval_uniform(p, precision):
    x, y, tau = H.uniform(200000, 61)
    print("self-eval uniform p", p, precision, flush=True)
    f64, mix = _pair(x, y, 100, p, precision=precision)
    f64.build_expansions(tau)
    mix.build_expansions(tau)
    u64 = f64.source_evaluation(x, y)[0]
    um = mix.source_evaluation(x, y)[0]
    err = H.rel_max_err(um, u64)
    sel = np.arange(0, x.size, 211)
    ref = odirect.direct(okern.laplace, x, y, tau, x[sel], y[sel], check=True)[0]
    errd = H.rel_max_err(um[sel], ref)
    print("p %d mixed vs fp64 %.3e, vs direct %.3e" % (p, err, errd))
    assert err < MIXED_TOL and errd < MIXED_TOL


def test_mixed_grid_targets_clustered():
    x, y, tau = H.clustered(150000, 20, 0.005, 62)
    tx, ty = H.grid(300)
    print("mixed clustered grid", flush=True)
    f64, mix = _pair(x, y, 100, 60, bbox=[0, 1, 0, 1])
    f64.build_expansions(tau)
    mix.build_expansions(tau)
    u64 = f64.target_evaluation(tx, ty)[0]
    um = mix.target_evaluation(tx, ty)[0]
    err = H.rel_max_err(um, u64)
    sel = np.arange(0, tx.size, 97)
    ref = odirect.direct(okern.laplace, x, y, tau, tx[sel], ty[sel])[0]
    errd = H.rel_max_err(um[sel], ref)
    print("levels %d mixed vs fp64 %.3e, vs direct %.3e" % (mix.tree.levels, err, errd))
    assert err < MIXED_TOL and errd < MIXED_TOL


def test_mixed_coincident_targets_are_infinite():
    """Without self-exclusion a target on a source gives log(0): +inf, as in
    the FP64 form (and the reference)."""
    x, y, tau = H.uniform(20000, 63)
    f64, mix = _pair(x, y, 64, 32)
    f64.build_expansions(tau)
    mix.build_expansions(tau)
    sel = np.arange(0, x.size, 101)
    tx = np.concatenate([x[sel], [0.123456, 0.5]])
    ty = np.concatenate([y[sel], [0.654321, 0.5]])
    um = mix.target_evaluation(tx, ty)[0]
    u64 = f64.target_evaluation(tx, ty)[0]
    assert np.all(np.isposinf(um[:-2])) and np.all(np.isposinf(u64[:-2]))
    assert np.all(np.isfinite(um[-2:]))
    assert H.rel_max_err(um[-2:], u64[-2:]) < 1e-6


def test_mixed_gradient_evaluator():
    """The gradient evaluator keeps an FP64 near field; only its local
    expansions carry the FP32 M2L error."""
    from flexmm_amd import kernels
    x, y, tau = H.uniform(50000, 64)
    tx, ty, _ = H.uniform(3000, 65)
    f64, mix = _pair(x, y, 64, 60)
    for f in (f64, mix):
        f.register_evaluator(kernels.laplace_gradient, kernels.laplace_gradient, "gradient", 3)
        f.build_expansions(tau)
    g64 = f64.evaluate_to_points(tx, ty, "gradient", False)
    gm = mix.evaluate_to_points(tx, ty, "gradient", False)
    errs = [H.rel_max_err(gm[k], g64[k]) for k in range(3)]
    print("gradient mixed vs fp64", errs)
    assert max(errs) < 1e-6


def test_mixed_reuse_after_rebuild():
    """Rebuilding expansions with new charges refreshes the FP32 source copy."""
    x, y, tau = H.uniform(30000, 66)
    tau2 = np.random.default_rng(67).standard_normal(x.size)
    f64, mix = _pair(x, y, 64, 32)
    mix.build_expansions(tau)
    mix.source_evaluation(x, y)
    mix.build_expansions(tau2)
    um = mix.source_evaluation(x, y)[0]
    f64.build_expansions(tau2)
    u64 = f64.source_evaluation(x, y)[0]
    assert H.rel_max_err(um, u64) < MIXED_TOL
