"""Pins of the black-box (Chebyshev) FMM oracle (oracle/bbfmm.py) by what the
mathematics fixes: numpy's Chebyshev routines, exactness of M2M and L2L on
polynomials, the closed-form bivariate interpolation behind M2L, and FP64
direct sums with geometric convergence in the order."""
import numpy as np
import pytest

from oracle import bbfmm as ob
from oracle import direct as odirect
from oracle import kernels as okern

C = np.polynomial.chebyshev


@pytest.mark.parametrize("p", [2, 5, 8])
def test_chebyshev_basics(p):
    rng = np.random.default_rng(p)
    np.testing.assert_allclose(ob.cheb_nodes(p), np.sort(C.chebgauss(p)[0]), rtol=0, atol=1e-15)
    x, y = rng.uniform(-1, 1, 50), rng.uniform(-1, 1, 50)
    np.testing.assert_allclose(ob.chebt(p, x), C.chebvander(x, p - 1), rtol=0, atol=1e-14)
    np.testing.assert_allclose(ob.chebv2(x, y, p), C.chebvander2d(x, y, [p - 1, p - 1]),
                               rtol=0, atol=1e-14)
    c = rng.standard_normal(p * p)
    np.testing.assert_allclose(ob.cheb_eval2d(x, y, c[None, :], p),
                               C.chebval2d(x, y, c.reshape(p, p)), rtol=1e-13, atol=1e-13)


def _child_moments(info, sx, sy, tau, w):
    """moments of each child box of a box of width w centred at 0, children
    in the order (-,-), (-,+), (+,-), (+,+)."""
    p = info.p
    out = []
    for cx, cy in [(-1, -1), (-1, 1), (1, -1), (1, 1)]:
        sel = ((sx > 0) == (cx > 0)) & ((sy > 0) == (cy > 0))
        xc, yc = cx * w / 4, cy * w / 4
        out.append(ob.chebv2((sx[sel] - xc) * 4 / w, (sy[sel] - yc) * 4 / w, p).T @ tau[sel])
    return np.concatenate(out)


@pytest.mark.parametrize("p", [3, 6, 8])
def test_m2m_is_exact_on_moments(p):
    """Moments about the parent equal M2MC times the children's moments: the
    parent's T_i(x) T_j(y) are polynomials of degree < p in each child's
    coordinates, which p-point interpolation reproduces exactly."""
    rng = np.random.default_rng(30 + p)
    info = ob.BBInfo(p)
    w = 0.37
    sx, sy = rng.uniform(-w / 2, w / 2, 400), rng.uniform(-w / 2, w / 2, 400)
    tau = rng.standard_normal(400)
    parent = ob.chebv2(sx * 2 / w, sy * 2 / w, p).T @ tau
    got = info.M2MC @ _child_moments(info, sx, sy, tau, w)
    assert np.abs(got - parent).max() < 1e-12 * np.abs(parent).max()
    # planted mistake: children (-,+) and (+,-) swapped
    m = _child_moments(info, sx, sy, tau, w).reshape(4, -1)[[0, 2, 1, 3]].ravel()
    assert np.abs(info.M2MC @ m - parent).max() > 1e-3 * np.abs(parent).max()


@pytest.mark.parametrize("p", [3, 8])
def test_l2l_is_exact_on_expansions(p):
    """A child's expansion from L2LC equals the parent's expansion inside the
    child."""
    rng = np.random.default_rng(40 + p)
    info = ob.BBInfo(p)
    c = rng.standard_normal(p * p)
    kids = (info.L2LC @ c).reshape(4, p * p)
    for k, (cx, cy) in enumerate([(-1, -1), (-1, 1), (1, -1), (1, 1)]):
        x, y = rng.uniform(-1, 1, 30), rng.uniform(-1, 1, 30)      # child coordinates
        ref = ob.cheb_eval2d((x + cx) / 2, (y + cy) / 2, c[None, :], p)
        got = ob.cheb_eval2d(x, y, kids[k][None, :], p)
        np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-12)


def _interp_weights(p, xh):
    """Chebyshev interpolation weights S_p(x, n_a) = 1/p + 2/p sum_k T_k(x) T_k(n_a)."""
    n = ob.cheb_nodes(p)
    Tx = C.chebvander(np.atleast_1d(xh), p - 1)
    Tn = C.chebvander(n, p - 1)
    return (1.0 + 2.0 * (Tx[:, None, 1:] * Tn[None, :, 1:]).sum(-1)) / p


@pytest.mark.parametrize("kernel", ["laplace", "helmholtz"])
def test_m2l_is_bivariate_chebyshev_interpolation(kernel):
    """S2M -> M2L -> L2T of one source and one target equals the kernel
    interpolated at the Chebyshev nodes of both boxes (Fong & Darve, J. Comput.
    Phys. 228 (2009) 8712): sum_ab S(t, n_a) K(n_b -> n_a) S(s, n_b)."""
    K = okern.laplace if kernel == "laplace" else okern.make_modified_helmholtz(7.0)[0]
    rng = np.random.default_rng(50)
    p, w = 6, 0.25
    prec = ob.BBPrecomputation(w, K, ob.BBInfo(p))
    for ix, iy in [(0, 3), (5, 1), (6, 6), (2, 0)]:
        off = np.array([(ix - 3) * w, (iy - 3) * w])
        s = off + rng.uniform(-w / 2, w / 2, 2)
        t = rng.uniform(-w / 2, w / 2, 2)
        m = ob.chebv2([(s[0] - off[0]) * 2 / w], [(s[1] - off[1]) * 2 / w], p)[0]
        le = prec.M2L[ix, iy] @ m
        got = ob.cheb_eval2d([t[0] * 2 / w], [t[1] * 2 / w], le[None, :], p)[0]
        nv = ob.cheb_nodes(p) * w / 2
        Sx_t, Sy_t = _interp_weights(p, t[0] * 2 / w)[0], _interp_weights(p, t[1] * 2 / w)[0]
        Sx_s = _interp_weights(p, (s[0] - off[0]) * 2 / w)[0]
        Sy_s = _interp_weights(p, (s[1] - off[1]) * 2 / w)[0]
        SX, SY = np.meshgrid(nv + off[0], nv + off[1], indexing="ij")   # source nodes (c, d)
        TX, TY = np.meshgrid(nv, nv, indexing="ij")                     # target nodes (a, b)
        Kab = K(SX[None, None], SY[None, None], TX[:, :, None, None], TY[:, :, None, None])
        ref = np.einsum("a,b,abcd,c,d->", Sx_t, Sy_t, Kab, Sx_s, Sy_s)
        assert abs(got - ref) < 1e-12 * max(1.0, abs(ref))
        # planted mistake: transposed M2L
        bad = ob.cheb_eval2d([t[0] * 2 / w], [t[1] * 2 / w],
                             (prec.M2L[ix, iy].T @ m)[None, :], p)[0]
        assert abs(bad - ref) > 1e-6 * max(1.0, abs(ref))


@pytest.mark.parametrize("kernel", ["laplace", "helmholtz"])
def test_bbfmm_converges_to_direct_sum(kernel):
    """Source and target evaluation against FP64 direct sums; the error falls
    geometrically with the order (about 5.8^-p for well-separated boxes)."""
    K = okern.laplace if kernel == "laplace" else okern.make_modified_helmholtz(10.0)[0]
    rng = np.random.default_rng(60)
    n = 4000
    x, y = rng.random(n), rng.random(n)
    tau = rng.standard_normal(n)
    tx, ty = rng.random(500), rng.random(500)
    sel = np.arange(0, n, 13)
    ref_s = odirect.direct(K, x, y, tau, x[sel], y[sel], check=True)[0]
    ref_t = odirect.direct(K, x, y, tau, tx, ty)[0]
    errs = []
    for p in (4, 6, 8):
        f = ob.BBFMM(x, y, K, Ncutoff=40, p=p, bbox=[0, 1, 0, 1])
        f.build_expansions(tau)
        us = f.source_evaluation(x, y)[sel]
        ut = f.target_evaluation(tx, ty)
        e = max(np.abs(us - ref_s).max() / np.abs(ref_s).max(),
                np.abs(ut - ref_t).max() / np.abs(ref_t).max())
        errs.append(e)
    assert f.tree.levels >= 4
    assert errs[0] > errs[1] > errs[2]
    assert errs[2] < 2e-5
