"""Pin the oracle's kernel functions and direct sums by closed forms,
finite differences and explicit loops."""
import math

import numpy as np
import pytest
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
    """The Stokeslet as Laplace potentials and gradients of six densities
    (f_x, f_y and their products with the source coordinates about a fixed
    centre) reproduces the direct Stokeslet sum; the three harmonic fields of
    the next tests regroup these terms."""
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


def _stokes_fd(fun, x, y, h):
    c = fun(x, y)
    return c, fun(x + h, y), fun(x - h, y), fun(x, y + h), fun(x, y - h)


def test_stresslet_is_minus_stokeslet_stress_times_normal():
    """The double-layer velocity of a source with density sigma and normal n
    equals -T n, T = -p I + grad u + grad u^T the stress of a Stokeslet of
    force sigma at the same source (unit viscosity), by central differences
    of direct_stokes; the double-layer pressure follows from the Stokes
    equations (next test)."""
    from oracle import direct as odirect
    rng = np.random.default_rng(21)
    for trial in range(4):
        s = rng.random(2)
        sig = rng.standard_normal(2)
        nrm = rng.standard_normal(2)
        nrm /= np.linalg.norm(nrm)
        t = rng.random(2) + np.array([1.0, 0.0]) * (trial % 2)
        h = 1e-5

        def f(x, y):
            return odirect.direct_stokes([s[0]], [s[1]], [sig[0]], [sig[1]],
                                         np.array([x]), np.array([y]))[:, 0]
        c, xp, xm, yp, ym = _stokes_fd(f, t[0], t[1], h)
        gu = np.array([[xp[0] - xm[0], yp[0] - ym[0]], [xp[1] - xm[1], yp[1] - ym[1]]]) / (2 * h)
        T = -c[2] * np.eye(2) + gu + gu.T
        got = odirect.direct_stokes_dipole([s[0]], [s[1]], [0.0], [0.0], [sig[0]], [sig[1]],
                                           [nrm[0]], [nrm[1]], t[:1], t[1:])[:, 0]
        ref = -T @ nrm
        assert np.abs(got[:2] - ref).max() < 1e-7 * np.abs(ref).max()


def test_stokes_double_layer_solves_stokes_equations():
    """div u = 0 and Laplacian(u) = grad p for Stokeslets plus stresslets away
    from the sources, by central differences; a tiny case by explicit loops in
    the reference's own form (flexmm/kifmm2d/vector/stokes_fmm.py:187-235)."""
    from oracle import direct as odirect
    rng = np.random.default_rng(22)
    m = 20
    sx, sy = rng.random(m) * 0.2 + 0.4, rng.random(m) * 0.2 + 0.4
    fx, fy, gx, gy = [rng.standard_normal(m) for _ in range(4)]
    th = rng.random(m) * 2 * np.pi
    nx, ny = np.cos(th), np.sin(th)
    h = 1e-4

    def f(x, y):
        return odirect.direct_stokes_dipole(sx, sy, fx, fy, gx, gy, nx, ny,
                                            np.array([x]), np.array([y]))[:, 0]
    for x, y in [[0.1, 0.15], [0.9, 0.2], [0.2, 0.85]]:
        c, xp, xm, yp, ym = _stokes_fd(f, x, y, h)
        div = (xp[0] - xm[0]) / (2 * h) + (yp[1] - ym[1]) / (2 * h)
        lap = (xp + xm + yp + ym - 4 * c) / h ** 2
        gp = np.array([(xp[2] - xm[2]) / (2 * h), (yp[2] - ym[2]) / (2 * h)])
        scale = np.abs(gp).max()
        assert abs(div) < 1e-6 * scale
        assert np.abs(lap[:2] - gp).max() < 1e-4 * scale
    # explicit loop over three sources
    x, y = 0.3, 0.7
    u = np.zeros(3)
    for j in range(3):
        dx, dy = x - sx[j], y - sy[j]
        r2 = dx * dx + dy * dy
        dn = 4 * (dx * nx[j] + dy * ny[j]) / r2 ** 2
        u[0] += ((-0.5 * np.log(r2) + dx * dx / r2) * fx[j] + dn * dx * dx * gx[j]
                 + dx * dy / r2 * fy[j] + dn * dx * dy * gy[j]) / (4 * np.pi)
        u[1] += (dx * dy / r2 * fx[j] + dn * dx * dy * gx[j]
                 + (-0.5 * np.log(r2) + dy * dy / r2) * fy[j] + dn * dy * dy * gy[j]) / (4 * np.pi)
        u[2] += ((dx * fx[j] + dy * fy[j]) / r2 + 2 * (-nx[j] / r2 + 0.5 * dx * dn) * gx[j]
                 + 2 * (-ny[j] / r2 + 0.5 * dy * dn) * gy[j]) / (2 * np.pi)
    got = odirect.direct_stokes_dipole(sx[:3], sy[:3], fx[:3], fy[:3], gx[:3], gy[:3], nx[:3],
                                       ny[:3], np.array([x]), np.array([y]))[:, 0]
    np.testing.assert_allclose(got, u, rtol=1e-12)
    # coincident target skipped under check
    got = odirect.direct_stokes_dipole(sx[:3], sy[:3], fx[:3], fy[:3], gx[:3], gy[:3], nx[:3],
                                       ny[:3], sx[:1], sy[:1], check=True)[:, 0]
    ref = odirect.direct_stokes_dipole(sx[1:3], sy[1:3], fx[1:3], fy[1:3], gx[1:3], gy[1:3],
                                       nx[1:3], ny[1:3], sx[:1], sy[:1])[:, 0]
    np.testing.assert_allclose(got, ref, rtol=1e-14)


def test_stresslet_laplace_decomposition():
    """The four extra Laplace fields that the device far field uses for the
    stresslet (g = n.sigma, a = n_x sigma_x - n_y sigma_y, b = n_x sigma_y +
    n_y sigma_x; Q[a, b] = sum a d_xx log|r| + b d_xy log|r|) reproduce the
    direct stresslet sum: u_i = (d_i L[g] - X_i Q[a, b] + Q[X_i a, X_i b])/(2 pi),
    p = -Q[a, b]/pi, X the coordinates about a fixed centre."""
    from oracle import direct as odirect
    rng = np.random.default_rng(23)
    m = 300
    sx, sy = rng.random(m), rng.random(m)
    gx, gy = rng.standard_normal(m), rng.standard_normal(m)
    th = rng.random(m) * 2 * np.pi
    nx, ny = np.cos(th), np.sin(th)
    tx, ty = rng.random(50) + 1.5, rng.random(50)
    c0 = np.array([0.4, 0.6])
    dx = tx[:, None] - sx[None, :]
    dy = ty[:, None] - sy[None, :]
    d2 = dx * dx + dy * dy
    g = nx * gx + ny * gy
    a = nx * gx - ny * gy
    b = nx * gy + ny * gx
    X, Y = sx - c0[0], sy - c0[1]

    def Q(aa, bb):
        return ((aa * (dy * dy - dx * dx) - 2 * bb * dx * dy) / d2 ** 2).sum(1)
    u = ((dx / d2 * g).sum(1) - (tx - c0[0]) * Q(a, b) + Q(X * a, X * b)) / (2 * np.pi)
    v = ((dy / d2 * g).sum(1) - (ty - c0[1]) * Q(a, b) + Q(Y * a, Y * b)) / (2 * np.pi)
    p = -Q(a, b) / np.pi
    z = np.zeros(m)
    ref = odirect.direct_stokes_dipole(sx, sy, z, z, gx, gy, nx, ny, tx, ty)
    for got, r in zip((u, v, p), ref):
        assert np.abs(got - r).max() < 1e-12 * np.abs(r).max()


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


def _three_fields(sx, sy, fx, fy, gx, gy, nx, ny, tx, ty, c0):
    """The three fields that the device far field carries, per source as the
    S2M writes them at a check point t (r = t - s, X = s - c0):
    D_i = sum -f_i log|r| - X_i (f.r)/|r|^2 + 2 (g r_i/|r|^2 + X_i q),
    D_2 = sum (f.r)/|r|^2 - 2 q, with g = n.sigma, a = n_x sigma_x - n_y sigma_y,
    b = n_x sigma_y + n_y sigma_x and q = (a (r_y^2 - r_x^2) - 2 b r_x r_y)/|r|^4."""
    dx = tx[:, None] - sx[None, :]
    dy = ty[:, None] - sy[None, :]
    d2 = dx * dx + dy * dy
    X, Y = sx - c0[0], sy - c0[1]
    fr = (fx * dx + fy * dy) / d2
    g = nx * gx + ny * gy
    a = nx * gx - ny * gy
    b = nx * gy + ny * gx
    q = (a * (dy * dy - dx * dx) - 2 * b * dx * dy) / d2 ** 2
    lg = 0.5 * np.log(d2)
    Dx = (-fx * lg - X * fr + 2 * (g * dx / d2 + X * q)).sum(1)
    Dy = (-fy * lg - Y * fr + 2 * (g * dy / d2 + Y * q)).sum(1)
    D2 = (fr - 2 * q).sum(1)
    return Dx, Dy, D2


@pytest.mark.parametrize("stresslets", [False, True])
def test_stokes_three_harmonic_fields(stresslets):
    """u_i = (D_i + X_i(t) D_2)/(4 pi) and p = D_2/(2 pi) reproduce the direct
    Stokeslet (+ stresslet) sums, and each field is harmonic away from the
    sources (5-point Laplacian), which is what lets single-layer equivalent
    densities represent it.  Planted mistake: the X_i(t) D_2 term with the
    coordinate about another centre fails."""
    from oracle import direct as odirect
    rng = np.random.default_rng(31 + stresslets)
    m = 300
    sx, sy = rng.random(m), rng.random(m)
    fx, fy = rng.standard_normal(m), rng.standard_normal(m)
    gx, gy = (rng.standard_normal(m), rng.standard_normal(m)) if stresslets else (np.zeros(m),) * 2
    th = rng.random(m) * 2 * np.pi
    nx, ny = np.cos(th), np.sin(th)
    tx, ty = rng.random(40) * 3 - 1, rng.random(40) + 1.2
    c0 = (0.45, 0.55)
    Dx, Dy, D2 = _three_fields(sx, sy, fx, fy, gx, gy, nx, ny, tx, ty, c0)
    u = (Dx + (tx - c0[0]) * D2) / (4 * np.pi)
    v = (Dy + (ty - c0[1]) * D2) / (4 * np.pi)
    p = D2 / (2 * np.pi)
    ref = odirect.direct_stokes_dipole(sx, sy, fx, fy, gx, gy, nx, ny, tx, ty)
    for got, r in zip((u, v, p), ref):
        assert np.abs(got - r).max() < 1e-13 * np.abs(r).max()
    bad = (Dx + (tx - 0.5) * D2) / (4 * np.pi)
    assert np.abs(bad - ref[0]).max() > 1e-3 * np.abs(ref[0]).max()
    # the 5-point Laplacian of each field is only its O(h^2) truncation, a
    # thousandth of the Laplacian 2 d_x D_2 of the non-harmonic X(t) D_2
    h = 5e-4
    c = _three_fields(sx, sy, fx, fy, gx, gy, nx, ny, tx, ty, c0)
    nb = [_three_fields(sx, sy, fx, fy, gx, gy, nx, ny, tx + ex, ty + ey, c0)
          for ex, ey in [(h, 0), (-h, 0), (0, h), (0, -h)]]
    lap_xd2 = 2 * np.abs(nb[0][2] - nb[1][2]).max() / (2 * h)
    for k in range(3):
        lap = (sum(f[k] for f in nb) - 4 * c[k]) / h ** 2
        assert np.abs(lap).max() < 1e-3 * lap_xd2
