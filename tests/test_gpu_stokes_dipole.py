"""Stokeslets plus stresslets (StokesFMM.build_expansions_dipole /
evaluate_to_points_dipole, the drop-in for kifmm2d/vector/stokes_fmm.py:738-858):
velocity and pressure against FP64 direct sums of the reference's
kernel_apply_single(_check)_with_dipole, self-evaluation on clustered sources
and off-source targets, FP64 and mixed precision, and the rules that keep the
single-layer and double-layer evaluators apart."""
import numpy as np
import pytest

from oracle import direct as odirect
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _sources(n, nc, sigma, seed, forces=True):
    rng = np.random.default_rng(seed)
    c = rng.uniform(0.1, 0.9, (nc, 2))
    pts = c[np.repeat(np.arange(nc), n // nc)] + sigma * rng.standard_normal((n // nc * nc, 2))
    pts = np.clip(pts, 0.0, 1.0)
    m = pts.shape[0]
    fx, fy = rng.standard_normal(m), rng.standard_normal(m)
    if not forces:
        fx, fy = np.zeros(m), np.zeros(m)
    gx, gy = rng.standard_normal(m), rng.standard_normal(m)
    th = rng.uniform(0, 2 * np.pi, m)
    return pts[:, 0].copy(), pts[:, 1].copy(), fx, fy, gx, gy, np.cos(th), np.sin(th)


# mixed: plans with stresslets keep the FP64 M2L (their quadrupole field has no
# dominant mode for the row-mean split to remove, and an FP32 product alone
# measured 1.0e-7 to 2.0e-7 in the velocity), so the mixed form meets 1e-7
@pytest.mark.parametrize("precision,tol,forces", [("fp64", 3e-11, True), ("fp64", 3e-11, False),
                                                  ("mixed", 1e-7, True)])
def test_stokes_dipole_self_eval_clustered(precision, tol, forces):
    from flexmm_amd.kifmm2d import StokesFMM
    x, y, fx, fy, gx, gy, nx, ny = _sources(40000, 10, 0.01, 191, forces)
    print("stokes dipole self", precision, "forces", forces, flush=True)
    f = StokesFMM(x, y, 64, 100, bbox=[0, 1, 0, 1], precision=precision)
    f.precompute()
    f.build_expansions_dipole(fx, fy, gx, gy, nx, ny)
    u, v, p = f.evaluate_to_points_dipole(x, y, True)
    sel = np.arange(0, x.size, 41)
    ref = odirect.direct_stokes_dipole(x, y, fx, fy, gx, gy, nx, ny, x[sel], y[sel], check=True)
    errs = [H.rel_max_err(a[sel], r) for a, r in zip((u, v, p), ref)]
    print(precision, "levels", f.tree.levels, "u v p errors", errs)
    assert max(errs) < tol


@pytest.mark.parametrize("p,tol", [(48, 1e-8), (64, 5e-11)])
def test_stokes_dipole_targets(p, tol):
    """Off-source targets; Nequiv 48 is the reference's default."""
    from flexmm_amd.kifmm2d import StokesFMM
    x, y, fx, fy, gx, gy, nx, ny = _sources(20000, 5, 0.02, 192)
    tx, ty, _ = H.uniform(3000, 193)
    f = StokesFMM(x, y, p, 64, bbox=[0, 1, 0, 1])
    f.build_expansions_dipole(fx, fy, gx, gy, nx, ny)
    u, v, pr = f.evaluate_to_points_dipole(tx, ty)
    ref = odirect.direct_stokes_dipole(x, y, fx, fy, gx, gy, nx, ny, tx, ty)
    errs = [H.rel_max_err(a, r) for a, r in zip((u, v, pr), ref)]
    print("p", p, "targets u v p errors", errs)
    assert max(errs) < tol


def test_stokes_dipole_zero_density_matches_single_layer():
    """With sigma = 0 the double-layer path gives the Stokeslet result."""
    from flexmm_amd.kifmm2d import StokesFMM
    x, y, fx, fy, gx, gy, nx, ny = _sources(20000, 4, 0.03, 194)
    z = np.zeros_like(x)
    f = StokesFMM(x, y, 64, 80, bbox=[0, 1, 0, 1])
    f.build_expansions(fx, fy)
    a = f.evaluate_to_points(x, y, True)
    f.build_expansions_dipole(fx, fy, z, z, nx, ny)
    b = f.evaluate_to_points_dipole(x, y, True)
    for s, d in zip(a, b):
        assert H.rel_max_err(d, s) < 1e-12


def test_stokes_dipole_evaluator_rules():
    """Each evaluator needs expansions built for it; a target on a source
    without self-exclusion gives NaN, as the reference's 1/0 does."""
    from flexmm_amd import _lib
    from flexmm_amd.kifmm2d import StokesFMM
    x, y, fx, fy, gx, gy, nx, ny = _sources(4000, 2, 0.05, 195)
    f = StokesFMM(x, y, 48, 50, bbox=[0, 1, 0, 1])
    f.build_expansions(fx, fy)
    with pytest.raises(_lib.FMMError):
        f.evaluate_to_points_dipole(x[:10], y[:10])
    f.build_expansions_dipole(fx, fy, gx, gy, nx, ny)
    with pytest.raises(_lib.FMMError):
        f.evaluate_to_points(x[:10], y[:10])
    u, v, p = f.evaluate_to_points_dipole(x[:10], y[:10], False)
    assert np.isnan(u).all() and np.isnan(v).all() and np.isnan(p).all()
