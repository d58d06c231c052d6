"""A sweep of small configurations against the oracle at the FP64 parity bar:
deep trees from small leaf capacities (items of a few targets, several LDS
chunks of colleague sources per item), uniform, clustered and curve-like
sources, self-evaluation and off-source targets, few and many targets, for
the KIFMM and the black-box FMM."""
import numpy as np
import pytest

from oracle import bbfmm as obb
from oracle import kernels as okern
from oracle import kifmm2d as okifmm
from tests import helpers as H

pytestmark = pytest.mark.gpu

PARITY_TOL = 1e-12


def _points(kind, n, rng):
    if kind == "uniform":
        return rng.random(n), rng.random(n)
    if kind == "clustered":
        c = rng.uniform(0.2, 0.8, (4, 2))
        p = c[rng.integers(0, 4, n)] + 0.03 * rng.standard_normal((n, 2))
        p = np.clip(p, 0.0, 1.0)
        return p[:, 0].copy(), p[:, 1].copy()
    th = rng.uniform(0, 2 * np.pi, n)          # on a circle, random order
    return 0.5 + 0.35 * np.cos(th), 0.5 + 0.35 * np.sin(th)


CASES = [
    # kind, n, Ncutoff, p, targets
    ("uniform", 3000, 8, 16, "self"),
    ("clustered", 6000, 12, 32, "self"),
    ("circle", 5000, 10, 24, "self"),
    ("uniform", 12000, 40, 32, "many"),
    ("clustered", 8000, 20, 32, "few"),
    ("circle", 4000, 30, 16, "many"),
]


@pytest.mark.parametrize("kind,n,ncut,p,tgt", CASES)
def test_kifmm_sweep(kind, n, ncut, p, tgt):
    from flexmm_amd import kernels
    from flexmm_amd.kifmm2d import FMM, the_help
    rng = np.random.default_rng([n, ncut, p, len(kind), len(tgt)])
    x, y = _points(kind, n, rng)
    tau = rng.standard_normal(n)
    print("sweep", kind, n, ncut, p, tgt, flush=True)
    the_help.clear()
    f = FMM(x, y, kernels.laplace, ncut, p, bbox=[0, 1, 0, 1])
    f.build_expansions(tau)
    o = okifmm.FMM(x, y, okern.laplace, ncut, p, bbox=[0, 1, 0, 1])
    o.build_expansions(tau)
    if tgt == "self":
        u, uo = f.source_evaluation(x, y)[0], o.source_evaluation(x, y)[0]
    else:
        m = 7 if tgt == "few" else 20000
        tx, ty = rng.random(m), rng.random(m)
        u, uo = f.target_evaluation(tx, ty)[0], o.target_evaluation(tx, ty)[0]
    err = H.rel_max_err(u, uo)
    print("levels", f.tree.levels, "err %.2e" % err)
    assert err < PARITY_TOL


@pytest.mark.parametrize("kind,n,ncut,order", [("uniform", 4000, 6, 5), ("clustered", 6000, 16, 8),
                                                ("circle", 3000, 12, 3)])
def test_bbfmm_sweep(kind, n, ncut, order):
    from flexmm_amd import kernels
    from flexmm_amd.flexmm2d import BB_FMM, the_bb_help
    rng = np.random.default_rng([n, ncut, order, len(kind)])
    x, y = _points(kind, n, rng)
    tau = rng.standard_normal(n)
    print("bb sweep", kind, n, ncut, order, flush=True)
    the_bb_help.clear()
    f = BB_FMM(x, y, kernels.laplace, Ncutoff=ncut, p=order, bbox=[0, 1, 0, 1])
    f.build_expansions(tau)
    o = obb.BBFMM(x, y, okern.laplace, Ncutoff=ncut, p=order, bbox=[0, 1, 0, 1])
    o.build_expansions(tau)
    err = H.rel_max_err(f.source_evaluation(x, y), o.source_evaluation(x, y))
    print("levels", f.tree.levels, "err %.2e" % err)
    assert err < PARITY_TOL
