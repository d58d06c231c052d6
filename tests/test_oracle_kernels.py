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
