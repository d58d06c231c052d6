"""Single + double layer Laplace sources (the 'slp+dlp' upwardor and
'slp+dlp->u' evaluator of examples/pure_kifmm_laplace_dipoles.py:32-40,
139-151): device potentials against the oracle restatement (FP64, 1e-12
relative in max norm) and against FP64 direct sums of the same source
kernel, through the reference's calling sequence (register_upwardor,
register_evaluator with the helper's kernel_app, build_expansions with
extras)."""
import numpy as np
import pytest

from oracle import direct as odirect
from oracle import kernels as okern
from oracle import kifmm2d as okifmm
from tests import helpers as H

pytestmark = pytest.mark.gpu

PARITY_TOL = 1e-12


def _densities(n, seed):
    rng = np.random.default_rng(seed)
    tau = rng.random((2, n))                 # slp, dlp (the example's np.random.rand)
    nx = rng.random(n)
    dipvec = np.stack([nx, np.sqrt(1 - nx ** 2)])
    return tau, dipvec


def _device(x, y, ncut, p, tau, dipvec, bbox=None, precision="fp64"):
    from flexmm_amd import kernels
    from flexmm_amd.kifmm2d import FMM, the_help
    the_help.clear()
    f = FMM(x, y, kernels.laplace, ncut, p, bbox=bbox, precision=precision)
    f.register_upwardor(kernels.laplace_slp_dlp, "slp+dlp")
    f.register_evaluator(f.helper.functions["kernel_app"], kernels.laplace_slp_dlp, "slp+dlp->u", 1)
    f.build_expansions(tau, "slp+dlp", dipvec)
    return f


def _oracle(x, y, ncut, p, tau, dipvec, bbox=None):
    o = okifmm.FMM(x, y, okern.laplace, ncut, p, bbox=bbox)
    o.register_evaluator(lambda *a: (okern.laplace(*a),), "slp+dlp->u", 1, src=okern.laplace_slp_dlp)
    o.build_expansions(tau, okern.laplace_slp_dlp, dipvec)
    return o


@pytest.mark.parametrize("p", [32, 60])
def test_dipole_self_and_targets_uniform(p):
    x, y, _ = H.uniform(30000, 61)
    tau, dv = _densities(x.size, 62)
    tx, ty, _ = H.uniform(5000, 63)
    print("dipole uniform p", p, flush=True)
    f = _device(x, y, 50, p, tau, dv)
    o = _oracle(x, y, 50, p, tau, dv)
    us = f.evaluate_to_points(x, y, "slp+dlp->u", True)[0]
    uso = o.evaluate_to_points(x, y, "slp+dlp->u", True)[0]
    ut = f.evaluate_to_points(tx, ty, "slp+dlp->u", False)[0]
    uto = o.evaluate_to_points(tx, ty, "slp+dlp->u", False)[0]
    es, et = H.rel_max_err(us, uso), H.rel_max_err(ut, uto)
    print("p", p, "self err %.3e target err %.3e" % (es, et))
    assert es < PARITY_TOL and et < PARITY_TOL
    if p == 60:
        sel = np.arange(0, x.size, 97)
        ref = odirect.direct_src(okern.laplace_slp_dlp, x, y, tau, dv, x[sel], y[sel], check=True)[0]
        assert H.rel_max_err(us[sel], ref) < 1e-12


def test_dipole_clustered_grid_targets():
    x, y, _ = H.clustered(40000, 5, 0.005, 64)
    tau, dv = _densities(x.size, 65)
    tx, ty = H.grid(100)
    print("dipole clustered grid", flush=True)
    f = _device(x, y, 40, 48, tau, dv, bbox=[0, 1, 0, 1])
    o = _oracle(x, y, 40, 48, tau, dv, bbox=[0, 1, 0, 1])
    u = f.evaluate_to_points(tx, ty, "slp+dlp->u", False)[0]
    uo = o.evaluate_to_points(tx, ty, "slp+dlp->u", False)[0]
    err = H.rel_max_err(u, uo)
    print("clustered err %.3e" % err)
    assert err < PARITY_TOL


def test_dipole_mixed_precision_and_coincident_target():
    x, y, _ = H.uniform(20000, 66)
    tau, dv = _densities(x.size, 67)
    tx, ty, _ = H.uniform(2000, 68)
    print("dipole mixed", flush=True)
    f64 = _device(x, y, 50, 60, tau, dv)
    fmx = _device(x, y, 50, 60, tau, dv, precision="mixed")
    u64 = f64.evaluate_to_points(tx, ty, "slp+dlp->u", False)[0]
    umx = fmx.evaluate_to_points(tx, ty, "slp+dlp->u", False)[0]
    assert H.rel_max_err(umx, u64) < 1e-7
    # a target on a source without self-exclusion: log(0) and 0/0, as the
    # reference's Laplace_Both_Eval gives
    u = f64.evaluate_to_points(x[:3], y[:3], "slp+dlp->u", False)[0]
    assert np.all(np.isnan(u))
    # the plain potential evaluator after a dipole upward pass still runs
    # (P2P with the single-layer charges, as the reference's kernel_app)
    f64.evaluate_to_points(tx[:10], ty[:10], "potential_evaluation", False)
