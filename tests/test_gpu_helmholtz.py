"""Modified-Helmholtz kernel K0(k r)/(2 pi), k = 20 (sources on a circle, as
in the differentiable workload): device FMM against the oracle (1e-12) and
against FP64 direct sums, potential and gradient."""
import numpy as np
import pytest

from oracle import direct as odirect
from oracle import kernels as okern
from oracle import kifmm2d as okifmm
from tests import helpers as H

pytestmark = pytest.mark.gpu

K = 20.0
BOX = [0, 1, 0, 1]   # the root box must hold the targets as well as the sources


def _circle(n, seed):
    rng = np.random.default_rng(seed)
    th = rng.uniform(0, 2 * np.pi, n)
    return 0.5 + 0.4 * np.cos(th), 0.5 + 0.4 * np.sin(th), rng.standard_normal(n)


@pytest.mark.parametrize("p", [32, 60])
def test_helmholtz_targets_vs_oracle_and_direct(p):
    from flexmm_amd import kernels
    from flexmm_amd.kifmm2d import FMM
    x, y, tau = _circle(20000, 71)
    tx, ty, _ = H.uniform(4000, 72)
    print("helmholtz p", p, flush=True)
    ker, pot, grad = kernels.modified_helmholtz(K)
    f = FMM(x, y, ker, 64, p, bbox=BOX)
    f.build_expansions(tau)
    u = f.target_evaluation(tx, ty)[0]
    mh, mhg = okern.make_modified_helmholtz(K)
    o = okifmm.FMM(x, y, mh, 64, p, bbox=BOX)
    o.build_expansions(tau)
    uo = o.target_evaluation(tx, ty)[0]
    err = H.rel_max_err(u, uo)
    ref = odirect.direct(mh, x, y, tau, tx, ty)[0]
    errd = H.rel_max_err(u, ref)
    print("p %d: vs oracle %.3e, vs direct %.3e" % (p, err, errd))
    assert err < 1e-12
    assert errd < (1e-12 if p == 60 else 1e-7)


def test_helmholtz_self_and_gradient():
    from flexmm_amd import kernels
    from flexmm_amd.kifmm2d import FMM
    x, y, tau = _circle(15000, 73)
    tx, ty, _ = H.uniform(2000, 74)
    ker, pot, grad = kernels.modified_helmholtz(K)
    f = FMM(x, y, ker, 64, 60, bbox=BOX)
    f.register_evaluator(grad, grad, "gradient", 3)
    f.build_expansions(tau)
    us = f.source_evaluation(x, y)[0]
    g = f.evaluate_to_points(tx, ty, "gradient", False)
    mh, mhg = okern.make_modified_helmholtz(K)
    sel = np.arange(0, x.size, 13)
    refs = odirect.direct(mh, x, y, tau, x[sel], y[sel], check=True)[0]
    assert H.rel_max_err(us[sel], refs) < 1e-11
    refg = odirect.direct(mhg, x, y, tau, tx, ty)
    errs = [H.rel_max_err(g[k], refg[k]) for k in range(3)]
    print("gradient vs direct", errs)
    assert max(errs) < 1e-10


def test_helmholtz_mixed_m2l():
    """precision='mixed' keeps the near field in FP64 for this kernel."""
    from flexmm_amd import kernels
    from flexmm_amd.kifmm2d import FMM
    x, y, tau = _circle(20000, 75)
    tx, ty, _ = H.uniform(3000, 76)
    ker, _, _ = kernels.modified_helmholtz(K)
    f64 = FMM(x, y, ker, 64, 60, bbox=BOX)
    mix = FMM(x, y, ker, 64, 60, bbox=BOX, precision="mixed")
    f64.build_expansions(tau)
    mix.build_expansions(tau)
    err = H.rel_max_err(mix.target_evaluation(tx, ty)[0], f64.target_evaluation(tx, ty)[0])
    print("helmholtz mixed vs fp64 %.3e" % err)
    assert err < 1e-7


@pytest.mark.parametrize("n,ncut", [(300, 400), (3000, 40)])
def test_helmholtz_near_field_both_branches(n, ncut):
    """Near-field pairs on both sides of x = k r = 2, where the device switches
    from the ascending series (taken from d^2) to the large-argument
    interpolant: a root leaf evaluated directly, and a shallow tree of big
    leaves; potential and gradient against FP64 direct sums."""
    from flexmm_amd import kernels
    from flexmm_amd.kifmm2d import FMM
    x, y, tau = H.uniform(n, 80 + ncut)
    tx, ty, _ = H.uniform(500, 81)
    print("helmholtz near field", n, ncut, flush=True)
    ker, pot, grad = kernels.modified_helmholtz(K)
    f = FMM(x, y, ker, ncut, 60, bbox=BOX)
    f.register_evaluator(grad, grad, "gradient", 3)
    f.build_expansions(tau)
    g = f.evaluate_to_points(tx, ty, "gradient", False)
    if ncut >= n:   # every pair is a near-field pair: both branches occur
        r = np.hypot(tx[:, None] - x[None, :], ty[:, None] - y[None, :])
        assert (K * r > 2).mean() > 0.3 and (K * r < 2).mean() > 0.01
    mh, mhg = okern.make_modified_helmholtz(K)
    refg = odirect.direct(mhg, x, y, tau, tx, ty)
    errs = [H.rel_max_err(g[k], refg[k]) for k in range(3)]
    print("levels", f.tree.levels, "vs direct", errs)
    assert max(errs) < 1e-11
