"""Pin the oracle FMM (restatement of flexmm kifmm2d) against exact FP64
direct sums, check the precomputed operators' closed forms and pseudo-inverse
identities, and show that plausible mistakes in the restatement are caught."""
import numpy as np
import pytest

from oracle import direct as odirect
from oracle import kernels as okern
from oracle import kifmm2d as okifmm
from tests import helpers as H


def _err(u, ref):
    return float(np.max(np.abs(u - ref)) / np.max(np.abs(ref)))


@pytest.fixture(scope="module")
def uniform5k():
    x, y, tau = H.uniform(5000, 51)
    ref = odirect.direct(okern.laplace, x, y, tau, x, y, check=True)[0]
    return x, y, tau, ref


def test_convergence_in_p(uniform5k):
    x, y, tau, ref = uniform5k
    errs = []
    for p in (16, 24, 32, 48, 64):
        f = okifmm.FMM(x, y, okern.laplace, 40, p)
        f.build_expansions(tau)
        errs.append(_err(f.source_evaluation(x, y)[0], ref))
    assert all(b < a for a, b in zip(errs, errs[1:])), errs
    assert errs[-1] < 1e-13, errs


def test_clustered_targets_p64():
    x, y, tau = H.clustered(8000, 4, 0.004, 32)
    tx, ty = H.grid(50)
    f = okifmm.FMM(x, y, okern.laplace, 25, 64, bbox=[0, 1, 0, 1])
    f.build_expansions(tau)
    ref = odirect.direct(okern.laplace, x, y, tau, tx, ty)[0]
    assert _err(f.target_evaluation(tx, ty)[0], ref) < 1e-12
    refs = odirect.direct(okern.laplace, x, y, tau, x, y, check=True)[0]
    assert _err(f.source_evaluation(x, y)[0], refs) < 1e-12


def test_gradient_p64():
    x, y, tau = H.uniform(4000, 52)
    tx, ty, _ = H.uniform(500, 53)
    f = okifmm.FMM(x, y, okern.laplace, 40, 64)
    f.register_evaluator(okern.laplace_gradient, "gradient", 3)
    f.build_expansions(tau)
    g = f.evaluate_to_points(tx, ty, "gradient", False)
    ref = odirect.direct(okern.laplace_gradient, x, y, tau, tx, ty)
    for k in range(3):
        assert _err(g[k], ref[k]) < 1e-11


def test_operator_closed_forms():
    w, p = 0.25, 24
    pr = okifmm.Precomputation(w, okern.laplace, p)
    th = 2 * np.pi * np.arange(p) / p
    r1, r2 = w * np.sqrt(2) / 2, w * (4 - np.sqrt(2)) / 2
    # S2L: K(small_i -> large_j) at distance |r2 e^{i th_j} - r1 e^{i th_i}|
    dist = np.abs(r2 * np.exp(1j * th)[:, None] - r1 * np.exp(1j * th)[None, :])
    np.testing.assert_allclose(pr.S2L, -np.log(dist) / (2 * np.pi), rtol=1e-13)
    # M2MC column block c: child (sx, sy) = (+-w/4, +-w/4) in order (-,-),(-,+),(+,-),(+,+)
    q = w / 4
    for c, (sx, sy) in enumerate([(-q, -q), (-q, q), (q, -q), (q, q)]):
        src = sx + 1j * sy + (r1 / 2) * np.exp(1j * th)
        d = np.abs(r2 * np.exp(1j * th)[:, None] - src[None, :])
        np.testing.assert_allclose(pr.M2MC[:, c * p:(c + 1) * p], -np.log(d) / (2 * np.pi), rtol=1e-13)
    np.testing.assert_array_equal(pr.L2LC, pr.M2MC.T)
    # M2L(ix, iy): source small circle shifted by ((ix-3) w, (iy-3) w)
    for ix, iy in [(0, 3), (6, 6), (3, 5), (1, 2)]:
        src = (ix - 3) * w + 1j * (iy - 3) * w + r1 * np.exp(1j * th)
        d = np.abs(r1 * np.exp(1j * th)[:, None] - src[None, :])
        np.testing.assert_allclose(pr.M2L[ix, iy], -np.log(d) / (2 * np.pi), rtol=1e-13)


def test_pinv_and_compression_identities():
    pr = okifmm.Precomputation(1.0 / 64, okern.laplace, 60)
    A = pr.S2L
    # applied in factored form, as the passes do: A pinv(A) b = b for b = A q
    q = np.random.default_rng(5).standard_normal((60, 8))
    b = A @ q
    assert np.abs(A @ pr.S2L_Solver(b) - b).max() / np.abs(b).max() < 1e-13
    for ix in range(7):
        for iy in range(7):
            ev = pr.M2LF[ix, iy]
            if ev is None:
                continue
            if ev.kind == "svd":
                err = np.abs(ev.M2 @ ev.M1 - ev.m2l).max()
                assert err < 1e-13, (ix, iy, err)


# --- planted mistakes ----------------------------------------------------------

def _fmm_err(x, y, tau, ref, p=48):
    f = okifmm.FMM(x, y, okern.laplace, 40, p)
    f.build_expansions(tau)
    return _err(f.source_evaluation(x, y)[0], ref)


@pytest.fixture
def fresh_cache():
    okifmm._PRECOMP_CACHE.clear()
    yield
    okifmm._PRECOMP_CACHE.clear()


def test_planted_m2l_offset_sign_is_caught(uniform5k, fresh_cache, monkeypatch):
    x, y, tau, ref = uniform5k
    orig = okifmm.kernel_form

    def flipped(K, sx, sy, tx, ty):
        if sx.size == tx.size and not np.shares_memory(sx, tx) and np.abs(np.mean(sx) - np.mean(tx)) > 0:
            return orig(K, 2 * tx - sx, 2 * ty - sy, tx, ty)   # source circle mirrored through the target
        return orig(K, sx, sy, tx, ty)

    assert _fmm_err(x, y, tau, ref) < 1e-10
    okifmm._PRECOMP_CACHE.clear()
    monkeypatch.setattr(okifmm, "kernel_form", flipped)
    assert _fmm_err(x, y, tau, ref) > 1e-6


def test_planted_child_order_is_caught(uniform5k, fresh_cache, monkeypatch):
    x, y, tau, ref = uniform5k
    orig_init = okifmm.Precomputation.__init__

    def swapped(self, width, K, p):
        orig_init(self, width, K, p)
        m = self.M2MC.copy()
        self.M2MC = np.concatenate([m[:, p:2 * p], m[:, :p], m[:, 2 * p:]], axis=1)
        self.L2LC = self.M2MC.T

    monkeypatch.setattr(okifmm.Precomputation, "__init__", swapped)
    assert _fmm_err(x, y, tau, ref) > 1e-6


def test_planted_child_width_l2l_is_caught(uniform5k, fresh_cache, monkeypatch):
    """L2L must use the parent level's L2LC (fmm.py:142,162); using the
    child's operator is off by a factor 2 in geometry."""
    x, y, tau, ref = uniform5k
    orig = okifmm.FMM.downwards_pass

    def wrong(self):
        saved = list(self.precs)
        try:
            shifted = saved[1:] + saved[-1:]
            for d in range(len(self.precs)):
                pr = okifmm.Precomputation.__new__(okifmm.Precomputation)
                pr.__dict__.update(saved[d].__dict__)
                pr.L2LC = shifted[d].L2LC
                self.precs[d] = pr
            orig(self)
        finally:
            self.precs[:] = saved

    monkeypatch.setattr(okifmm.FMM, "downwards_pass", wrong)
    assert _fmm_err(x, y, tau, ref) > 1e-6


def test_planted_missing_self_exclusion_is_caught(uniform5k):
    x, y, tau, ref = uniform5k
    f = okifmm.FMM(x, y, okern.laplace, 40, 48)
    f.build_expansions(tau)
    with np.errstate(all="ignore"):
        u = f.target_evaluation(x, y)[0]   # check=False at coincident points
    assert not np.all(np.isfinite(u))


def test_get_bbox_rules():
    x, y, _ = H.uniform(100, 54)
    b = okifmm.get_bbox(x, y, None)
    assert abs((b[1] - b[0]) - max(x.max() - x.min(), y.max() - y.min())) < 1e-15
    assert abs((b[1] + b[0]) / 2 - (x.max() + x.min()) / 2) < 1e-15
    # warm helper: the width snaps to a power-of-two multiple of the smallest cached width
    b2 = okifmm.get_bbox(x, y, [0.1, 0.7, 0.2, 0.5], widths=[1.0, 0.5, 0.25, 0.125])
    assert b2[1] - b2[0] == 0.125 * 2 ** np.ceil(np.log2(0.6 / 0.125))


def test_modified_helmholtz_vs_direct():
    """The oracle with the K0(k r)/(2 pi) kernel, k = 20, sources on a circle."""
    rng = np.random.default_rng(77)
    th = rng.uniform(0, 2 * np.pi, 3000)
    x, y = 0.5 + 0.4 * np.cos(th), 0.5 + 0.4 * np.sin(th)
    tau = rng.standard_normal(3000)
    tx, ty = rng.random(500), rng.random(500)
    mh, mhg = okern.make_modified_helmholtz(20.0)
    f = okifmm.FMM(x, y, mh, 40, 60, bbox=[0, 1, 0, 1])   # the targets fill the unit square
    f.register_evaluator(mhg, "gradient", 3)
    f.build_expansions(tau)
    ref = odirect.direct(mhg, x, y, tau, tx, ty)
    g = f.evaluate_to_points(tx, ty, "gradient", False)
    for k in range(3):
        assert _err(g[k], ref[k]) < 1e-11


def test_slp_dlp_upwardor_p64():
    """Upwardor with extras (the dipole example, examples/pure_kifmm_laplace_dipoles.py:
    139-151): single + double layer sources, self and target evaluation, against
    FP64 direct sums of the same source kernel."""
    x, y, _ = H.uniform(4000, 54)
    rng = np.random.default_rng(55)
    tau = rng.random((2, x.size))
    nx = rng.random(x.size)
    ex = np.stack([nx, np.sqrt(1 - nx ** 2)])
    tx, ty, _ = H.uniform(400, 56)
    f = okifmm.FMM(x, y, okern.laplace, 40, 64)
    f.register_evaluator(lambda *a: (okern.laplace(*a),), "slp+dlp->u", 1, src=okern.laplace_slp_dlp)
    f.build_expansions(tau, okern.laplace_slp_dlp, ex)
    u = f.evaluate_to_points(tx, ty, "slp+dlp->u", False)[0]
    ref = odirect.direct_src(okern.laplace_slp_dlp, x, y, tau, ex, tx, ty)[0]
    assert _err(u, ref) < 1e-12
    us = f.evaluate_to_points(x[:300], y[:300], "slp+dlp->u", True)[0]
    refs = odirect.direct_src(okern.laplace_slp_dlp, x, y, tau, ex, x[:300], y[:300], check=True)[0]
    assert _err(us, refs) < 1e-12


def test_parallel_oracle_equals_serial():
    """The forked-worker oracle (the CPU baseline of the benchmark) gives the
    serial oracle's potentials: S2M over leaf chunks and evaluation over target
    chunks only regroup additions."""
    from oracle import parallel as opar
    x, y, tau = H.uniform(20000, 31)
    tx, ty = H.grid(90)
    a = okifmm.FMM(x, y, okern.laplace, 40, 24, bbox=[0, 1, 0, 1])
    a.build_expansions(tau)
    b = opar.ParallelFMM(x, y, okern.laplace, 40, 24, bbox=[0, 1, 0, 1], workers=3)
    b.build_expansions(tau)
    for m in (a.multipoles, b.multipoles):
        assert len(m) == a.tree.levels
    assert H.rel_max_err(b.multipoles[-1], a.multipoles[-1]) < 1e-13
    assert H.rel_max_err(b.target_evaluation(tx, ty), a.target_evaluation(tx, ty)) < 1e-14
    assert H.rel_max_err(b.source_evaluation(x, y), a.source_evaluation(x, y)) < 1e-14
