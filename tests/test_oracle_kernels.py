"""Pin the oracle's kernel functions and direct sums by closed forms,
finite differences and explicit loops."""
import math

import numpy as np
from scipy import special

from oracle import direct as odirect
from oracle import kernels as okern


def test_laplace_closed_form():
    rng = np.random.default_rng(0)
    s = rng.random((2, 100))
    t = rng.random((2, 100))
    r = np.hypot(t[0] - s[0], t[1] - s[1])
    np.testing.assert_allclose(okern.laplace(s[0], s[1], t[0], t[1]), -np.log(r) / (2 * np.pi),
                               rtol=1e-14, atol=1e-15)


def test_laplace_gradient_is_target_derivative():
    rng = np.random.default_rng(1)
    s = rng.random((2, 50))
    t = rng.random((2, 50)) + 2.0
    u, ux, uy = okern.laplace_gradient(s[0], s[1], t[0], t[1])
    h = 1e-6
    fx = (okern.laplace(s[0], s[1], t[0] + h, t[1]) - okern.laplace(s[0], s[1], t[0] - h, t[1])) / (2 * h)
    fy = (okern.laplace(s[0], s[1], t[0], t[1] + h) - okern.laplace(s[0], s[1], t[0], t[1] - h)) / (2 * h)
    np.testing.assert_allclose(ux, fx, rtol=1e-7)
    np.testing.assert_allclose(uy, fy, rtol=1e-7)
    np.testing.assert_allclose(u, okern.laplace(s[0], s[1], t[0], t[1]), rtol=0)


def test_modified_helmholtz_against_scipy():
    mh, mhg = okern.make_modified_helmholtz(20.0)
    rng = np.random.default_rng(2)
    s = rng.random((2, 60))
    t = rng.random((2, 60))
    r = np.hypot(t[0] - s[0], t[1] - s[1])
    np.testing.assert_allclose(mh(s[0], s[1], t[0], t[1]), special.k0(20 * r) / (2 * np.pi), rtol=1e-14)
    h = 1e-6
    u, ux, uy = mhg(s[0], s[1], t[0], t[1])
    fx = (mh(s[0], s[1], t[0] + h, t[1]) - mh(s[0], s[1], t[0] - h, t[1])) / (2 * h)
    np.testing.assert_allclose(ux, fx, rtol=1e-6, atol=1e-10)


def test_direct_sum_matches_python_loops():
    rng = np.random.default_rng(3)
    sx, sy, tau = rng.random(40), rng.random(40), rng.standard_normal(40)
    tx, ty = rng.random(7), rng.random(7)
    got = odirect.direct(okern.laplace, sx, sy, tau, tx, ty)[0]
    for i in range(7):
        ref = sum(-math.log((tx[i] - sx[j]) ** 2 + (ty[i] - sy[j]) ** 2) / (4 * math.pi) * tau[j]
                  for j in range(40))
        assert abs(got[i] - ref) < 1e-13 * max(1.0, abs(ref))


def test_direct_sum_self_exclusion():
    rng = np.random.default_rng(4)
    sx, sy, tau = rng.random(30), rng.random(30), rng.standard_normal(30)
    got = odirect.direct(okern.laplace, sx, sy, tau, sx, sy, check=True)[0]
    for i in range(30):
        ref = sum(-math.log((sx[i] - sx[j]) ** 2 + (sy[i] - sy[j]) ** 2) / (4 * math.pi) * tau[j]
                  for j in range(30) if j != i)
        assert abs(got[i] - ref) < 1e-13 * max(1.0, abs(ref))


def test_stokeslet_solves_stokes_equations():
    """div u = 0 and Laplacian(u) = grad p away from the sources (unit
    viscosity), by central differences; a tiny case by explicit loops."""
    from oracle import direct as odirect
    rng = np.random.default_rng(11)
    sx, sy = rng.random(20) * 0.2 + 0.4, rng.random(20) * 0.2 + 0.4
    fx, fy = rng.standard_normal(20), rng.standard_normal(20)
    P = np.array([[0.1, 0.15], [0.9, 0.2], [0.2, 0.85]])
    h = 1e-3
    def f(x, y):
        return odirect.direct_stokes(sx, sy, fx, fy, np.array([x]), np.array([y]))[:, 0]
    for x, y in P:
        c = f(x, y)
        xp, xm, yp, ym = f(x + h, y), f(x - h, y), f(x, y + h), f(x, y - h)
        div = (xp[0] - xm[0]) / (2 * h) + (yp[1] - ym[1]) / (2 * h)
        lap = (xp + xm + yp + ym - 4 * c) / h ** 2
        gp = np.array([(xp[2] - xm[2]) / (2 * h), (yp[2] - ym[2]) / (2 * h)])
        scale = np.abs(gp).max()
        assert abs(div) < 1e-6 * scale
        assert np.abs(lap[:2] - gp).max() < 1e-4 * scale
    # explicit loop
    u = np.zeros(3)
    x, y = 0.3, 0.7
    for j in range(3):
        dx, dy = x - sx[j], y - sy[j]
        r2 = dx * dx + dy * dy
        u[0] += ((-0.5 * np.log(r2) + dx * dx / r2) * fx[j] + dx * dy / r2 * fy[j]) / (4 * np.pi)
        u[1] += (dx * dy / r2 * fx[j] + (-0.5 * np.log(r2) + dy * dy / r2) * fy[j]) / (4 * np.pi)
        u[2] += (dx * fx[j] + dy * fy[j]) / r2 / (2 * np.pi)
    got = odirect.direct_stokes(sx[:3], sy[:3], fx[:3], fy[:3], np.array([x]), np.array([y]))[:, 0]
    np.testing.assert_allclose(got, u, rtol=1e-13)


def test_stokeslet_laplace_decomposition():
    """The six-Laplace-density form of the Stokeslet that the device far field
    uses reproduces the direct Stokeslet sum."""
    from oracle import direct as odirect
    rng = np.random.default_rng(12)
    sx, sy = rng.random(400), rng.random(400)
    fx, fy = rng.standard_normal(400), rng.standard_normal(400)
    tx, ty = rng.random(60), rng.random(60)
    c0 = 0.5
    X, Y = sx - c0, sy - c0

    def L(g):   # sum g log|t - s| and its gradient
        dx = tx[:, None] - sx[None, :]
        dy = ty[:, None] - sy[None, :]
        d2 = dx * dx + dy * dy
        return ((0.5 * np.log(d2) * g).sum(1), (dx / d2 * g).sum(1), (dy / d2 * g).sum(1))

    Lx, Ly = L(fx), L(fy)
    B = Lx[1] + Ly[2]
    Cx = L(X * fx)[1] + L(X * fy)[2]
    Cy = L(Y * fx)[1] + L(Y * fy)[2]
    u = (-Lx[0] + (tx - c0) * B - Cx) / (4 * np.pi)
    v = (-Ly[0] + (ty - c0) * B - Cy) / (4 * np.pi)
    p = B / (2 * np.pi)
    ref = odirect.direct_stokes(sx, sy, fx, fy, tx, ty)
    for got, r in zip((u, v, p), ref):
        assert np.abs(got - r).max() < 1e-13 * np.abs(r).max()


def test_double_layer_is_source_derivative_of_single_layer():
    """(n.d)/d^2/(2 pi) with d = t - s is the derivative of the single layer
    -log(d^2)/(4 pi) along n at the source: the dipole of
    examples/pure_kifmm_laplace_dipoles.py:32-40."""
    rng = np.random.default_rng(17)
    sx, sy, tx, ty = rng.random(4)
    th = rng.uniform(0, 2 * np.pi)
    nx, ny = np.cos(th), np.sin(th)
    h = 1e-6
    fd = (okern.laplace(sx + h * nx, sy + h * ny, tx, ty)
          - okern.laplace(sx - h * nx, sy - h * ny, tx, ty)) / (2 * h)
    dl = okern.laplace_slp_dlp(sx, sy, tx, ty, np.array([0.0, 1.0]), np.array([nx, ny]))[0]
    assert abs(fd - dl) < 1e-8 * abs(dl)
    sl = okern.laplace_slp_dlp(sx, sy, tx, ty, np.array([1.0, 0.0]), np.array([nx, ny]))[0]
    assert sl == okern.laplace(sx, sy, tx, ty)


def test_direct_src_matches_python_loops():
    rng = np.random.default_rng(18)
    sx, sy = rng.random(7), rng.random(7)
    tau = rng.standard_normal((2, 7))
    ex = rng.standard_normal((2, 7))
    tx, ty = np.concatenate([rng.random(3), sx[:2]]), np.concatenate([rng.random(3), sy[:2]])
    got = odirect.direct_src(okern.laplace_slp_dlp, sx, sy, tau, ex, tx, ty, check=True)[0]
    for i in range(tx.size):
        acc = 0.0
        for j in range(sx.size):
            if sx[j] == tx[i] and sy[j] == ty[i]:
                continue
            dx, dy = tx[i] - sx[j], ty[i] - sy[j]
            d2 = dx * dx + dy * dy
            acc += -np.log(d2) / (4 * np.pi) * tau[0, j] + (ex[0, j] * dx + ex[1, j] * dy) / d2 / (2 * np.pi) * tau[1, j]
        assert abs(got[i] - acc) < 1e-13 * max(1.0, abs(acc))
