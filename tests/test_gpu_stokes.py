"""Stokeslet FMM (StokesFMM, the drop-in for kifmm2d/vector/stokes_fmm.py):
velocity and pressure against FP64 direct sums on clustered sources,
self-evaluation and off-source targets, FP64 and mixed precision, and the
adjoint with respect to the forces."""
import numpy as np
import pytest

from oracle import direct as odirect
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _clusters(n, nc, sigma, seed):
    rng = np.random.default_rng(seed)
    c = rng.uniform(0.1, 0.9, (nc, 2))
    pts = c[np.repeat(np.arange(nc), n // nc)] + sigma * rng.standard_normal((n // nc * nc, 2))
    pts = np.clip(pts, 0.0, 1.0)
    m = pts.shape[0]
    return pts[:, 0].copy(), pts[:, 1].copy(), rng.standard_normal(m), rng.standard_normal(m)


# mixed: the FP32 M2L takes D_0, D_1 in the target box's frame (1.4e-8 to
# 2.0e-8 measured in the velocity, 7e-9 in the pressure)
@pytest.mark.parametrize("precision,tol,ptol", [("fp64", 1e-11, 1e-11), ("mixed", 1e-7, 1e-7)])
def test_stokes_self_eval_clustered(precision, tol, ptol):
    from flexmm_amd.kifmm2d import StokesFMM
    x, y, fx, fy = _clusters(40000, 10, 0.01, 91)
    print("stokes self", precision, flush=True)
    f = StokesFMM(x, y, 64, 100, bbox=[0, 1, 0, 1], precision=precision)
    f.precompute()
    f.build_expansions(fx, fy)
    u, v, p = f.evaluate_to_points(x, y, True)
    sel = np.arange(0, x.size, 41)
    ref = odirect.direct_stokes(x, y, fx, fy, x[sel], y[sel], check=True)
    errs = [H.rel_max_err(a[sel], r) for a, r in zip((u, v, p), ref)]
    print(precision, "levels", f.tree.levels, "u v p errors", errs)
    assert max(errs[:2]) < tol and errs[2] < ptol


@pytest.mark.parametrize("p,tol", [(48, 1e-9), (64, 1e-11)])
def test_stokes_targets_uniform(p, tol):
    """Off-source targets; Nequiv 48 is the reference's default (truncation
    error near 1e-10 there)."""
    from flexmm_amd.kifmm2d import StokesFMM
    x, y, fx, fy = _clusters(20000, 5, 0.02, 92)
    tx, ty, _ = H.uniform(3000, 93)
    f = StokesFMM(x, y, p, 64, bbox=[0, 1, 0, 1])
    f.build_expansions(fx, fy)
    u, v, pr = f.evaluate_to_points(tx, ty)
    ref = odirect.direct_stokes(x, y, fx, fy, tx, ty)
    errs = [H.rel_max_err(a, r) for a, r in zip((u, v, pr), ref)]
    print("p", p, "targets u v p errors", errs)
    assert max(errs) < tol


def test_stokes_adjoint_forces():
    import torch
    from flexmm_amd.kifmm2d import stokes_evaluate
    x, y, fx, fy = _clusters(20000, 8, 0.02, 94)
    dev = torch.device("cuda", 0)
    T = lambda a: torch.tensor(a, dtype=torch.float64, device=dev)
    X, Y, FX, FY = T(x), T(y), T(fx), T(fy)
    FX.requires_grad_(True)
    FY.requires_grad_(True)
    u, v = stokes_evaluate(X, Y, FX, FY, X, Y, 64, 100, bbox=[0, 1, 0, 1], check=True)
    rng = np.random.default_rng(95)
    wu, wv = T(rng.standard_normal(x.size)), T(rng.standard_normal(x.size))
    ((u * wu).sum() + (v * wv).sum()).backward()
    # d/df_s = sum_t G(s - t) w_t: the direct Stokeslet sum with weights as forces
    sel = np.arange(0, x.size, 53)
    ref = odirect.direct_stokes(x, y, wu.cpu().numpy(), wv.cpu().numpy(), x[sel], y[sel], check=True)
    e1 = H.rel_max_err(FX.grad.cpu().numpy()[sel], ref[0])
    e2 = H.rel_max_err(FY.grad.cpu().numpy()[sel], ref[1])
    print("stokes adjoint errors", e1, e2)
    assert max(e1, e2) < 1e-10


def test_stokes_duplicate_sources_and_coincident_targets():
    """Sources repeated at exactly the same point: self-evaluation skips every
    coincident pair (each duplicate as well as the point itself), as the
    reference's kernel_apply_single_check does; without self-exclusion a
    target on a source gives NaN (the reference's 0/0)."""
    from flexmm_amd.kifmm2d import StokesFMM
    rng = np.random.default_rng(97)
    base = rng.random((2500, 2))
    pts = np.concatenate([base, base[:700], base[:100]])   # triples and pairs
    x, y = pts[:, 0].copy(), pts[:, 1].copy()
    fx, fy = rng.standard_normal(x.size), rng.standard_normal(x.size)
    print("stokes duplicates", x.size, flush=True)
    f = StokesFMM(x, y, 64, 40, bbox=[0, 1, 0, 1])
    f.build_expansions(fx, fy)
    u, v, p = f.evaluate_to_points(x, y, True)
    sel = np.arange(0, x.size, 7)
    ref = odirect.direct_stokes(x, y, fx, fy, x[sel], y[sel], check=True)
    errs = [H.rel_max_err(a[sel], r) for a, r in zip((u, v, p), ref)]
    print("vs direct", errs)
    assert max(errs) < 1e-11
    u, v, p = f.evaluate_to_points(x[:5], y[:5], False)
    assert np.isnan(u).all() and np.isnan(v).all() and np.isnan(p).all()
