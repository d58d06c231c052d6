"""Parity at the full sizes of the benchmark workloads.

* 1M uniform sources -> 1000 x 1000 grid, Nequiv 60, Ncutoff 100 (the
  reference README's headline setup, README.md:9-33; error metric of
  examples/pure_kifmm_laplace.py:149-151): FP64 against the oracle on all 1M
  targets, the whole 8-level device tree against the oracle's, and the mixed
  form against the FP64 form and FP64 direct sums.
* 1M points in 50 Gaussian clusters (sigma 0.005) -> the same grid: the
  deep adaptive tree (balance and X-list nodes) against the oracle's, FP64
  against the oracle, mixed against direct sums.
* Modified Helmholtz K0(20 r)/(2 pi), 500K sources on a circle -> 200K
  targets: the adjoint (d/dcharges, d/dsource positions) against FP64
  direct-sum gradients at 2,000 sources and central differences (h = 1e-6).
* Stokeslet, 2M points in 100 clusters (sigma 0.01), self-evaluation, Nequiv
  64: velocity against FP64 direct sums at 4,000 sources (FP64 and
  mixed-precision M2L), and the adjoint with respect to the forces.

The FP64 direct sums at sampled targets run with torch on the GPU
(tests/helpers.py); each test pins them against oracle/direct.py's NumPy sums
on a subset first."""
import numpy as np
import pytest

from oracle import direct as odirect
from oracle import kernels as okern
from oracle import kifmm2d as okifmm
from tests import helpers as H

pytestmark = pytest.mark.gpu

PARITY_TOL = 1e-12
MIXED_TOL = 1e-7


def _dev():
    import torch
    return torch.device("cuda", 0)


def _T(a):
    import torch
    return torch.tensor(np.ascontiguousarray(a), dtype=torch.float64, device=_dev())


def _laplace_headline(seed=2, n=1_000_000):
    rng = np.random.default_rng(seed)
    x, y = rng.random(n), rng.random(n)
    tau = rng.standard_normal(n)
    tx, ty = H.grid(1000)
    return x, y, tau, tx, ty


def _pinned_laplace_direct(x, y, tau, tx, ty, sel):
    """torch direct sums at targets sel, pinned by the NumPy oracle on 16 of them."""
    import torch
    X, Y, TAU = _T(x), _T(y), _T(tau)
    ref = H.torch_direct_laplace(torch, X, Y, TAU, _T(tx[sel]), _T(ty[sel])).cpu().numpy()
    pin = odirect.direct(okern.laplace, x, y, tau, tx[sel[:16]], ty[sel[:16]])[0]
    assert H.rel_max_err(ref[:16], pin) < 1e-14
    return ref


def _laplace_case(x, y, tau, tx, ty, label, tree_levels=None):
    import torch
    from flexmm_amd import kernels
    from flexmm_amd.kifmm2d import FMM, the_help
    the_help.clear()
    box = [0.0, 1.0, 0.0, 1.0]
    print(label, "fp64 on the GPU", flush=True)
    X, Y, TAU, TX, TY = _T(x), _T(y), _T(tau), _T(tx), _T(ty)
    f = FMM(X, Y, kernels.laplace, 100, 60, bbox=box)
    f.build_expansions(TAU)
    u = f.target_evaluation(TX, TY)[0].cpu().numpy()
    print(label, "oracle", flush=True)
    o = okifmm.FMM(x, y, okern.laplace, 100, 60, bbox=box)
    n = H.compare_trees_fast(f, o.tree)
    print(label, "levels %d, %d nodes equal to the oracle's" % (f.tree.levels, n), flush=True)
    if tree_levels is not None:
        assert f.tree.levels == tree_levels
    o.build_expansions(tau)
    uo = o.target_evaluation(tx, ty)[0]
    err = H.rel_max_err(u, uo)
    print(label, "fp64 vs oracle on all %d targets %.3e" % (tx.size, err), flush=True)
    assert err < PARITY_TOL
    del o
    rng = np.random.default_rng(7)
    sel = np.sort(rng.choice(tx.size, 2000, replace=False))
    ref = _pinned_laplace_direct(x, y, tau, tx, ty, sel)
    errd = H.rel_max_err(u[sel], ref)
    print(label, "fp64 vs direct on 2000 targets %.3e" % errd, flush=True)
    assert errd < 5e-13
    f.close()
    m = FMM(X, Y, kernels.laplace, 100, 60, bbox=box, precision="mixed")
    m.build_expansions(TAU)
    um = m.target_evaluation(TX, TY)[0].cpu().numpy()
    e64 = H.rel_max_err(um, u)
    emd = H.rel_max_err(um[sel], ref)
    print(label, "mixed vs fp64 %.3e, vs direct %.3e" % (e64, emd), flush=True)
    assert e64 < MIXED_TOL and emd < MIXED_TOL
    m.close()
    torch.cuda.synchronize()


def test_laplace_1M_uniform_to_1M_grid():
    x, y, tau, tx, ty = _laplace_headline()
    _laplace_case(x, y, tau, tx, ty, "1M uniform -> 1M grid", tree_levels=8)


def test_laplace_1M_clustered_to_1M_grid():
    x, y, tau = H.clustered(1_000_000, 50, 0.005, 3)
    tx, ty = H.grid(1000)
    _laplace_case(x, y, tau, tx, ty, "1M in 50 clusters -> 1M grid")


def test_helmholtz_adjoint_500k_circle_to_200k():
    import torch
    from flexmm_amd import kernels
    from flexmm_amd.kifmm2d import fmm_evaluate
    k = 20.0
    rng = np.random.default_rng(4)
    n, nt = 500_000, 200_000
    th = rng.uniform(0, 2 * np.pi, n)
    x, y = 0.5 + 0.4 * np.cos(th), 0.5 + 0.4 * np.sin(th)
    tau, tau2 = rng.standard_normal(n), rng.standard_normal(n)
    tg = rng.random((nt, 2))
    tx, ty = tg[:, 0].copy(), tg[:, 1].copy()
    X, Y, TAU, TAU2, TX, TY = [_T(a) for a in (x, y, tau, tau2, tx, ty)]
    ker, _, _ = kernels.modified_helmholtz(k)
    box = [0.0, 1.0, 0.0, 1.0]
    print("helmholtz adjoint 500K -> 200K", flush=True)
    with torch.no_grad():
        utrue = fmm_evaluate(X, Y, TAU2, TX, TY, ker, 100, 60, bbox=box)
    Xg, Yg, Tg = X.clone().requires_grad_(True), Y.clone().requires_grad_(True), \
        TAU.clone().requires_grad_(True)
    u = fmm_evaluate(Xg, Yg, Tg, TX, TY, ker, 100, 60, bbox=box)
    ((u - utrue) ** 2).sum().backward()
    g = (2.0 * (u - utrue)).detach()
    # the forward potential itself against direct sums at 500 targets
    mh, mhg = okern.make_modified_helmholtz(k)
    st = np.sort(rng.choice(nt, 500, replace=False))
    with torch.no_grad():
        ref_u = []
        for s0 in range(0, st.size, 8):
            ss = torch.from_numpy(st[s0:s0 + 8]).to(_dev())
            r = torch.sqrt((TX[ss, None] - X[None, :]) ** 2 + (TY[ss, None] - Y[None, :]) ** 2)
            ref_u.append((torch.special.modified_bessel_k0(k * r) * TAU[None, :]).sum(1))
        ref_u = (torch.cat(ref_u) / (2 * np.pi)).cpu().numpy()
    pin = odirect.direct(mh, x, y, tau, tx[st[:8]], ty[st[:8]])[0]
    assert H.rel_max_err(ref_u[:8], pin) < 1e-13
    eu = H.rel_max_err(u.detach().cpu().numpy()[st], ref_u)
    print("forward vs direct %.3e" % eu, flush=True)
    assert eu < 1e-12
    # adjoint against direct-sum gradients at 2,000 sources
    sel = np.sort(rng.choice(n, 2000, replace=False))
    SEL = torch.from_numpy(sel).to(_dev())
    with torch.no_grad():
        rt, rx, ry = H.torch_direct_mh_adjoint(torch, k, X[SEL], Y[SEL], TAU[SEL], TX, TY, g)
    gn = g.cpu().numpy()
    pin = odirect.direct(mhg, tx, ty, gn, x[sel[:4]], y[sel[:4]])
    assert H.rel_max_err(rt[:4].cpu().numpy(), pin[0]) < 1e-12
    assert H.rel_max_err(rx[:4].cpu().numpy(), tau[sel[:4]] * pin[1]) < 1e-12
    errs = dict(tau=H.rel_max_err(Tg.grad[SEL].cpu().numpy(), rt.cpu().numpy()),
                x=H.rel_max_err(Xg.grad[SEL].cpu().numpy(), rx.cpu().numpy()),
                y=H.rel_max_err(Yg.grad[SEL].cpu().numpy(), ry.cpu().numpy()))
    print("adjoint vs direct", {kk: "%.2e" % v for kk, v in errs.items()}, flush=True)
    assert max(errs.values()) < 1e-10
    # central differences, h = 1e-6: L(+) - L(-) = sum_t (u+ - u-)(u+ + u- - 2 u_true)
    h = 1e-6
    for which, grad in (("tau", Tg.grad), ("x", Xg.grad), ("y", Yg.grad)):
        for i in rng.choice(n, 2, replace=False):
            vals = []
            for sgn in (1.0, -1.0):
                a = dict(x=X.clone(), y=Y.clone(), tau=TAU.clone())
                a[which][i] += sgn * h
                with torch.no_grad():
                    vals.append(fmm_evaluate(a["x"], a["y"], a["tau"], TX, TY, ker, 100, 60,
                                             bbox=box))
            fd = float(((vals[0] - vals[1]) * (vals[0] + vals[1] - 2 * utrue)).sum()) / (2 * h)
            ad = float(grad[i])
            print(which, int(i), "fd %.12e ad %.12e" % (fd, ad), flush=True)
            assert abs(fd - ad) <= 1e-5 * max(1.0, abs(ad)), (which, int(i), fd, ad)


def _stokes_clusters():
    rng = np.random.default_rng(5)
    nc, per = 100, 20_000
    c = rng.uniform(0.1, 0.9, (nc, 2))
    pts = np.clip(c[np.repeat(np.arange(nc), per)] + 0.01 * rng.standard_normal((nc * per, 2)),
                  0, 1)
    n = pts.shape[0]
    return pts[:, 0].copy(), pts[:, 1].copy(), rng.standard_normal(n), rng.standard_normal(n), rng


# The mixed form's velocity: the FP32 M2L takes the fields D_0, D_1 in the
# target box's frame, so the cancellation of D_i + X_i(t) D_2 at the targets
# no longer amplifies it; 4.8e-8 to 9.2e-8 measured over runs (the point
# order inside a leaf, set by the sort's atomic scatter, changes between runs
# and with it the rounding).
STOKES_MIXED_TOL = 2e-7


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-11), ("mixed", STOKES_MIXED_TOL)])
def test_stokeslet_2M_clusters_self(precision, tol):
    import torch
    from flexmm_amd.kifmm2d import stokes_evaluate
    x, y, fx, fy, rng = _stokes_clusters()
    n = x.size
    X, Y, FX, FY = _T(x), _T(y), _T(fx), _T(fy)
    print("stokeslet 2M self", precision, flush=True)
    FXg, FYg = FX.clone().requires_grad_(True), FY.clone().requires_grad_(True)
    u, v = stokes_evaluate(X, Y, FXg, FYg, X, Y, 64, 100, bbox=[0, 1, 0, 1],
                           precision=precision, check=True)
    wu, wv = _T(rng.standard_normal(n)), _T(rng.standard_normal(n))
    ((u * wu).sum() + (v * wv).sum()).backward()
    sel = np.sort(rng.choice(n, 4000, replace=False))
    SEL = torch.from_numpy(sel).to(_dev())
    with torch.no_grad():
        ru, rv = H.torch_direct_stokes(torch, X, Y, FX, FY, X[SEL], Y[SEL], check=True)
        au, av = H.torch_direct_stokes(torch, X, Y, wu, wv, X[SEL], Y[SEL], check=True)
    pin = odirect.direct_stokes(x, y, fx, fy, x[sel[:4]], y[sel[:4]], check=True)
    assert H.rel_max_err(ru[:4].cpu().numpy(), pin[0]) < 1e-13
    assert H.rel_max_err(rv[:4].cpu().numpy(), pin[1]) < 1e-13
    ev = max(H.rel_max_err(u.detach()[SEL].cpu().numpy(), ru.cpu().numpy()),
             H.rel_max_err(v.detach()[SEL].cpu().numpy(), rv.cpu().numpy()))
    ea = max(H.rel_max_err(FXg.grad[SEL].cpu().numpy(), au.cpu().numpy()),
             H.rel_max_err(FYg.grad[SEL].cpu().numpy(), av.cpu().numpy()))
    print(precision, "velocity vs direct %.3e, adjoint vs direct %.3e" % (ev, ea), flush=True)
    assert ev < tol and ea < tol
