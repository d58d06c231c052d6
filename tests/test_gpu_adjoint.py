"""Adjoint of the FMM evaluation (fmm_evaluate): gradients of a loss with
respect to charges, source positions and target positions, against FP64
direct-sum gradients and central finite differences."""
import numpy as np
import pytest

from oracle import direct as odirect
from oracle import kernels as okern
from tests import helpers as H

pytestmark = pytest.mark.gpu


def _kernels(name):
    from flexmm_amd import kernels
    if name == "laplace":
        return kernels.laplace, okern.laplace, okern.laplace_gradient
    ker, _, _ = kernels.modified_helmholtz(20.0)
    mh, mhg = okern.make_modified_helmholtz(20.0)
    return ker, mh, mhg


def _setup(n, nt, seed):
    import torch
    rng = np.random.default_rng(seed)
    th = rng.uniform(0, 2 * np.pi, n)
    x, y = 0.5 + 0.4 * np.cos(th), 0.5 + 0.4 * np.sin(th)
    tau, tau2 = rng.standard_normal(n), rng.standard_normal(n)
    tx, ty = rng.random(nt), rng.random(nt)
    dev = torch.device("cuda", 0)
    T = lambda a: torch.tensor(a, dtype=torch.float64, device=dev)
    return (x, y, tau, tau2, tx, ty), [T(a) for a in (x, y, tau, tau2, tx, ty)]


@pytest.mark.parametrize("kname", ["laplace", "modified_helmholtz"])
def test_adjoint_vs_direct(kname):
    import torch
    from flexmm_amd.kifmm2d import fmm_evaluate
    ker, K, KG = _kernels(kname)
    (x, y, tau, tau2, tx, ty), (X, Y, TAU, TAU2, TX, TY) = _setup(20000, 3000, 81)
    print("adjoint", kname, flush=True)
    box = [0.0, 1.0, 0.0, 1.0]
    with torch.no_grad():
        utrue = fmm_evaluate(X, Y, TAU2, TX, TY, ker, 64, 60, bbox=box)
    for a in (X, Y, TAU, TX, TY):
        a.requires_grad_(True)
    u = fmm_evaluate(X, Y, TAU, TX, TY, ker, 64, 60, bbox=box)
    loss = ((u - utrue) ** 2).sum()
    loss.backward()
    g = (2.0 * (u - utrue)).detach().cpu().numpy()
    # direct-sum adjoint on sampled sources / targets
    ss = np.arange(0, x.size, 97)
    v = odirect.direct(KG, tx, ty, g, x[ss], y[ss])        # sum_t g_t K(s - t), grad wrt s
    ref_tau = v[0]
    ref_x = tau[ss] * v[1]
    ref_y = tau[ss] * v[2]
    st = np.arange(0, tx.size, 7)
    w = odirect.direct(KG, x, y, tau, tx[st], ty[st])
    ref_tx = g[st] * w[1]
    ref_ty = g[st] * w[2]
    errs = dict(
        tau=H.rel_max_err(TAU.grad.cpu().numpy()[ss], ref_tau),
        x=H.rel_max_err(X.grad.cpu().numpy()[ss], ref_x),
        y=H.rel_max_err(Y.grad.cpu().numpy()[ss], ref_y),
        tx=H.rel_max_err(TX.grad.cpu().numpy()[st], ref_tx),
        ty=H.rel_max_err(TY.grad.cpu().numpy()[st], ref_ty))
    print(kname, {k: "%.2e" % v for k, v in errs.items()})
    assert max(errs.values()) < 1e-10


def test_adjoint_finite_differences():
    """Central differences (h = 1e-6) of L = ||u(t) - u_true||^2 in a few
    charges and coordinates, modified Helmholtz k = 20."""
    import torch
    from flexmm_amd.kifmm2d import fmm_evaluate
    ker, K, KG = _kernels("modified_helmholtz")
    (x, y, tau, tau2, tx, ty), (X, Y, TAU, TAU2, TX, TY) = _setup(8000, 1500, 82)
    box = [0.0, 1.0, 0.0, 1.0]
    with torch.no_grad():
        utrue = fmm_evaluate(X, Y, TAU2, TX, TY, ker, 64, 60, bbox=box)

    def loss_of(Xv, Yv, TAUv):
        with torch.no_grad():
            u = fmm_evaluate(Xv, Yv, TAUv, TX, TY, ker, 64, 60, bbox=box)
            return float(((u - utrue) ** 2).sum())

    Xg, Yg, TAUg = X.clone().requires_grad_(True), Y.clone().requires_grad_(True), TAU.clone().requires_grad_(True)
    u = fmm_evaluate(Xg, Yg, TAUg, TX, TY, ker, 64, 60, bbox=box)
    ((u - utrue) ** 2).sum().backward()
    h = 1e-6
    rng = np.random.default_rng(83)
    for which, base, grad in (("tau", TAU, TAUg.grad), ("x", X, Xg.grad), ("y", Y, Yg.grad)):
        for i in rng.choice(x.size, 3, replace=False):
            plus, minus = base.clone(), base.clone()
            plus[i] += h
            minus[i] -= h
            args_p = dict(tau=(X, Y, plus), x=(plus, Y, TAU), y=(X, plus, TAU))[which]
            args_m = dict(tau=(X, Y, minus), x=(minus, Y, TAU), y=(X, minus, TAU))[which]
            fd = (loss_of(*args_p) - loss_of(*args_m)) / (2 * h)
            ad = float(grad[i])
            print(which, int(i), "fd %.10e ad %.10e" % (fd, ad))
            assert abs(fd - ad) <= 1e-5 * max(1.0, abs(ad)), (which, i, fd, ad)
