"""Helper reuse across FMM instances (the module-level helper cache of
kifmm2d/scalar/fmm.py:11,40-43 and the warm-cache branch of Helper.get_bbox,
helpers.py:72-80): a second FMM on a non-dyadic point set snaps its root box
to the cached widths, reuses their operators without forming new ones, and
matches the oracle constructed with the same cached widths.

Also the input rules at the boundary: NaN or infinite source coordinates are
refused by the tree build (FMM2D_E_ARG); a NaN target gets NaN outputs and
leaves the others untouched (the reference's arithmetic gives the same);
targets outside the root box are located to an edge leaf and extrapolated,
as the reference's locate_points does (tree.py:703-727)."""
import numpy as np
import pytest

from oracle import kernels as okern
from oracle import kifmm2d as okifmm
from tests import helpers as H

pytestmark = pytest.mark.gpu


def test_warm_helper_snaps_bbox_and_reuses_operators():
    from flexmm_amd import kernels
    from flexmm_amd.kifmm2d import FMM, the_help
    the_help.clear()
    x1, y1, tau1 = H.uniform(60000, 201)
    f1 = FMM(x1, y1, kernels.laplace, 64, 60, bbox=[0, 1, 0, 1])
    f1.build_expansions(tau1)
    helper = f1.helper
    widths = sorted(helper.collect_widths().tolist(), reverse=True)
    assert widths[0] == 1.0 and len(widths) == f1.tree.levels
    ops_before = {k: v for k, v in helper.precomputations.items()}
    dev_before = dict(helper._device_ops)
    # a second point set on a non-dyadic box: [0.13, 0.71] x [0.22, 0.69]
    rng = np.random.default_rng(202)
    x2 = 0.13 + 0.58 * rng.random(30000)
    y2 = 0.22 + 0.47 * rng.random(30000)
    tau2 = rng.standard_normal(30000)
    print("warm helper: cached widths", widths, flush=True)
    f2 = FMM(x2, y2, kernels.laplace, 64, 60)
    exp_bbox = okifmm.get_bbox(x2, y2, None, widths=widths)
    assert f2.bbox == exp_bbox
    w2 = f2.bbox[1] - f2.bbox[0]
    assert w2 == 1.0 and abs(w2 - 0.58) > 0.1   # snapped to a cached width, not the data's
    f2.build_expansions(tau2)
    u2 = f2.source_evaluation(x2, y2)[0]
    # the operators of every level came from the cache: no new width, same objects
    for d, w in enumerate(f2.tree.widths):
        k = helper._key(w)
        assert k in ops_before, "level %d width %g was not cached" % (d, w)
        if d >= 2:
            assert helper.precomputations[k] is ops_before[k]
    assert set(helper.precomputations) == set(ops_before)
    assert all(helper._device_ops[k] is v for k, v in dev_before.items())
    o = okifmm.FMM(x2, y2, okern.laplace, 64, 60, widths=widths)
    assert o.bbox == f2.bbox
    o.build_expansions(tau2)
    uo = o.source_evaluation(x2, y2)[0]
    err = H.rel_max_err(u2, uo)
    print("warm helper: snapped FMM vs oracle %.3e" % err, flush=True)
    assert err < 1e-12
    # a fresh helper would have taken the data's own box instead
    assert okifmm.get_bbox(x2, y2, None)[1] - okifmm.get_bbox(x2, y2, None)[0] < 0.6


def test_nonfinite_sources_are_refused():
    from flexmm_amd import _lib, kernels
    from flexmm_amd.kifmm2d import FMM, the_help
    the_help.clear()
    x, y, _ = H.uniform(5000, 203)
    for bad in (np.nan, np.inf, -np.inf):
        xb = x.copy()
        xb[1234] = bad
        with pytest.raises(_lib.FMMError) as ei:
            FMM(xb, y, kernels.laplace, 64, 32, bbox=[0, 1, 0, 1])
        assert ei.value.code == _lib.E_ARG
        assert "NaN or infinite" in str(ei.value)
    # the plan is usable again afterwards
    f = FMM(x, y, kernels.laplace, 64, 32, bbox=[0, 1, 0, 1])
    f.build_expansions(np.ones_like(x))
    assert np.isfinite(f.target_evaluation(x[:10] + 1e-3, y[:10])[0]).all()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_nonfinite_targets_give_nan_outputs_only_there(precision):
    from flexmm_amd import kernels
    from flexmm_amd.kifmm2d import FMM, the_help
    the_help.clear()
    x, y, tau = H.uniform(50000, 204)
    tx, ty, _ = H.uniform(20000, 205)
    f = FMM(x, y, kernels.laplace, 64, 60, bbox=[0, 1, 0, 1], precision=precision)
    f.build_expansions(tau)
    u = f.target_evaluation(tx, ty)[0]
    bx, by = tx.copy(), ty.copy()
    bx[17], by[4242], bx[9000], by[9000] = np.nan, np.nan, np.inf, -np.inf
    ub = f.target_evaluation(bx, by)[0]
    bad = np.zeros(tx.size, dtype=bool)
    bad[[17, 4242, 9000]] = True
    assert not np.isfinite(ub[bad]).any()
    assert H.rel_max_err(ub[~bad], u[~bad]) < 1e-14


def test_targets_outside_the_root_box_follow_the_reference():
    from flexmm_amd import kernels
    from flexmm_amd.kifmm2d import FMM, the_help
    the_help.clear()
    x, y, tau = H.uniform(30000, 206)
    rng = np.random.default_rng(207)
    # just outside each side of the unit box (far outside, the extrapolation
    # multiplies the rounding of the local equivalent densities, whose
    # pseudo-inverse leaves large near-null components, by the distance)
    tx = np.concatenate([-0.01 * rng.random(200), 1 + 0.01 * rng.random(200), rng.random(400)])
    ty = np.concatenate([rng.random(400), -0.01 * rng.random(200), 1 + 0.01 * rng.random(200)])
    f = FMM(x, y, kernels.laplace, 64, 60, bbox=[0, 1, 0, 1])
    f.build_expansions(tau)
    u = f.target_evaluation(tx, ty)[0]
    o = okifmm.FMM(x, y, okern.laplace, 64, 60, bbox=[0, 1, 0, 1])
    o.build_expansions(tau)
    uo = o.target_evaluation(tx, ty)[0]
    err = H.rel_max_err(u, uo)
    print("outside targets vs oracle %.3e" % err, flush=True)
    assert err < 1e-12
