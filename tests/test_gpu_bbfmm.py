"""Black-box (Chebyshev) FMM on the GPU (flexmm_amd.flexmm2d.BB_FMM, the
drop-in for flexmm/flexmm2d/bbfmm.py) against the oracle's restatement of the
same FMM at the same order (the approximation is identical, so the bar is
FP64 rounding), against FP64 direct sums (the order's truncation error), and
the FP32 M2L / near-field forms."""
import numpy as np
import pytest

from oracle import bbfmm as obb
from oracle import direct as odirect
from oracle import kernels as okern
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _bb(*args, **kw):
    """A BB_FMM with a fresh helper, so that its bounding box follows the
    fresh-helper rule the oracle uses (cached widths snap later boxes)."""
    from flexmm_amd.flexmm2d import BB_FMM, the_bb_help
    the_bb_help.clear()
    return BB_FMM(*args, **kw)


def _clustered(n, seed):
    rng = np.random.default_rng(seed)
    c = rng.uniform(0.15, 0.85, (8, 2))
    pts = c[rng.integers(0, 8, n)] + 0.02 * rng.standard_normal((n, 2))
    pts = np.clip(pts, 0.0, 1.0)
    return pts[:, 0].copy(), pts[:, 1].copy(), rng.standard_normal(n)


@pytest.mark.parametrize("case,order,ncut", [("uniform", 8, 50), ("uniform", 4, 30),
                                             ("clustered", 8, 40), ("clustered", 5, 64)])
def test_bbfmm_matches_oracle(case, order, ncut):
    from flexmm_amd import kernels
    from flexmm_amd.flexmm2d import BB_FMM
    if case == "uniform":
        x, y, tau = H.uniform(20000, 300 + order)
    else:
        x, y, tau = _clustered(20000, 310 + order)
    print("bbfmm", case, "order", order, "Ncutoff", ncut, flush=True)
    f = _bb(x, y, kernels.laplace, Ncutoff=ncut, p=order)
    f.build_expansions(tau)
    # targets inside the root box: outside it the reference extrapolates the
    # Chebyshev sums (values ~1e13 for these data), which is not a parity case
    ux, uy, _ = H.uniform(4000, 320)
    b = f.bbox
    tx, ty = b[0] + (b[1] - b[0]) * ux, b[2] + (b[3] - b[2]) * uy
    us = f.source_evaluation(x, y)
    ut = f.target_evaluation(tx, ty)
    assert us.shape == (x.size,) and ut.shape == (tx.size,)
    o = obb.BBFMM(x, y, okern.laplace, Ncutoff=ncut, p=order)
    assert o.tree.levels == f.tree.levels
    o.build_expansions(tau)
    es = H.rel_max_err(us, o.source_evaluation(x, y))
    et = H.rel_max_err(ut, o.target_evaluation(tx, ty))
    print(case, order, "levels", f.tree.levels, "vs oracle source/target", es, et)
    assert max(es, et) < 1e-12


@pytest.mark.parametrize("order,tol", [(4, 3e-3), (6, 5e-5), (8, 2e-6)])
def test_bbfmm_against_direct_sum(order, tol):
    """The truncation error of order p falls geometrically (oracle: 3.7e-4,
    8.1e-6, 2.6e-7 for p = 4, 6, 8 on uniform points)."""
    from flexmm_amd import kernels
    from flexmm_amd.flexmm2d import BB_FMM
    x, y, tau = H.uniform(30000, 330)
    f = BB_FMM(x, y, kernels.laplace, Ncutoff=50, p=order, bbox=[0, 1, 0, 1])
    f.build_expansions(tau)
    sel = np.arange(0, x.size, 29)
    ref = odirect.direct(okern.laplace, x, y, tau, x[sel], y[sel], check=True)[0]
    err = H.rel_max_err(f.source_evaluation(x, y)[sel], ref)
    print("order", order, "vs direct", err)
    assert err < tol


def test_bbfmm_modified_helmholtz_matches_oracle():
    from flexmm_amd import kernels
    from flexmm_amd.flexmm2d import BB_FMM
    k0, _, _ = kernels.modified_helmholtz(10.0)
    x, y, tau = _clustered(15000, 340)
    f = _bb(x, y, k0, Ncutoff=40, p=7)
    f.build_expansions(tau)
    o = obb.BBFMM(x, y, okern.make_modified_helmholtz(10.0)[0], Ncutoff=40, p=7)
    o.build_expansions(tau)
    e = H.rel_max_err(f.source_evaluation(x, y), o.source_evaluation(x, y))
    print("helmholtz vs oracle", e)
    assert e < 1e-12


# Black-box plans keep the SVD-compressed FP64 M2L under the mixed forms (an
# FP32 M2L accumulation alone costs Chebyshev coefficients ~1e-7, with no
# dominant mode to split off); "mixed" then differs from FP64 by its FP32
# near field only, "mixed_m2l" not at all.
@pytest.mark.parametrize("precision,tol", [("mixed_m2l", 1e-12), ("mixed", 1e-7)])
def test_bbfmm_mixed(precision, tol):
    from flexmm_amd import kernels
    from flexmm_amd.flexmm2d import BB_FMM
    from flexmm_amd.flexmm2d import bbfmm
    x, y, tau = H.uniform(40000, 350)
    tx, ty, _ = H.uniform(5000, 351)
    # a cleared helper cache each time: a warm one snaps the second root box
    # to its cached widths (helpers.py:72-80), a different tree
    bbfmm.the_bb_help.clear()
    ref = BB_FMM(x, y, kernels.laplace, Ncutoff=50, p=8)
    ref.build_expansions(tau)
    bbfmm.the_bb_help.clear()
    f = BB_FMM(x, y, kernels.laplace, Ncutoff=50, p=8, precision=precision)
    assert f.bbox == ref.bbox
    f.build_expansions(tau)
    e1 = H.rel_max_err(f.source_evaluation(x, y), ref.source_evaluation(x, y))
    e2 = H.rel_max_err(f.target_evaluation(tx, ty), ref.target_evaluation(tx, ty))
    print(precision, "vs FP64 black-box", e1, e2)
    assert max(e1, e2) < tol


def test_bbfmm_rules():
    """A target on a source gives +inf without self-exclusion (the reference's
    -log(0)); other evaluators and orders above 8 are refused."""
    from flexmm_amd import _lib, kernels
    from flexmm_amd.flexmm2d import BB_FMM
    x, y, tau = H.uniform(3000, 360)
    f = BB_FMM(x, y, kernels.laplace, p=6)
    f.build_expansions(tau)
    u = f.target_evaluation(x[:5], y[:5])
    assert np.isinf(u).all()
    with pytest.raises(_lib.FMMError):
        _lib.check(f.lib.fmm2d_upward_dipole(f.handle, f.tau.data_ptr(), f.tau.data_ptr(),
                                             x.size, None))
    with pytest.raises(NotImplementedError):
        BB_FMM(x, y, kernels.laplace, p=9)
