"""The FP64 K0/K1 of the modified-Helmholtz kernel (flexmm_amd/csrc/bessel.h):
its two formulas restated in NumPy with the generated coefficients, checked
against scipy.special.k0/k1.  The device code itself is pinned by the
modified-Helmholtz FMM against direct sums on the GPU."""
import os
import re

import numpy as np
from scipy import special

HERE = os.path.dirname(os.path.abspath(__file__))
COEF = os.path.join(HERE, "..", "flexmm_amd", "csrc", "bessel_coef.h")
GAMMA = 0.57721566490153286061


def _coefs():
    txt = open(COEF).read()
    out = {}
    for name in ("kK0Cheb", "kK1Cheb"):
        body = re.search(name + r"\[\d+\] = \{(.*?)\};", txt, re.S).group(1)
        out[name] = np.array([float(v) for v in body.replace("\n", " ").split(",") if v.strip()])
    return out


def _clenshaw(c, u):
    b1 = np.zeros_like(u)
    b2 = np.zeros_like(u)
    for k in range(len(c) - 1, 0, -1):
        b1, b2 = c[k] + 2 * u * b1 - b2, b1
    return c[0] + u * b1 - b2


def _series(x, terms=14):
    """The ascending series as the device evaluates them: polynomials in
    q = x^2/4 with coefficients 1/(k!)^2, H_k/(k!)^2, 1/(k!(k+1)!) and
    (H_k + H_{k+1} - 2 gamma)/(k!(k+1)!), by Horner's rule."""
    c = np.zeros((4, terms))
    f = g = 1.0
    h = 0.0
    for k in range(terms):
        if k > 0:
            f /= k * k
            g /= k * (k + 1)
            h += 1.0 / k
        h1 = h + 1.0 / (k + 1)
        c[:, k] = f, h * f, g, (h + h1 - 2 * GAMMA) * g
    q = 0.25 * x * x
    lg = np.log(0.5 * x)
    i0, s0, i1, s1 = [np.full_like(x, c[j, -1]) for j in range(4)]
    for k in range(terms - 2, -1, -1):
        i0, s0, i1, s1 = i0 * q + c[0, k], s0 * q + c[1, k], i1 * q + c[2, k], s1 * q + c[3, k]
    return s0 - (lg + GAMMA) * i0, 1.0 / x + lg * 0.5 * x * i1 - 0.25 * x * s1


def test_k0_k1_against_scipy():
    c = _coefs()
    assert len(c["kK0Cheb"]) == 25 and len(c["kK1Cheb"]) == 25
    xs = np.geomspace(1e-10, 2.0, 4000)
    k0, k1 = _series(xs)
    assert np.abs(k0 / special.k0(xs) - 1).max() < 1e-14
    assert np.abs(k1 / special.k1(xs) - 1).max() < 1e-14
    xl = np.concatenate([np.linspace(2.0 + 1e-12, 40.0, 4000), np.geomspace(40.0, 700.0, 400)])
    u = 4.0 / xl - 1.0
    e = np.exp(-xl) / np.sqrt(xl)
    assert np.abs(e * _clenshaw(c["kK0Cheb"], u) / special.k0(xl) - 1).max() < 1e-14
    assert np.abs(e * _clenshaw(c["kK1Cheb"], u) / special.k1(xl) - 1).max() < 1e-14


def test_formulas_meet_at_two():
    c = _coefs()
    x = np.array([2.0])
    k0s, k1s = _series(x)
    e = np.exp(-x) / np.sqrt(x)
    assert abs(k0s[0] - e[0] * _clenshaw(c["kK0Cheb"], np.array([1.0]))[0]) < 2e-15
    assert abs(k1s[0] - e[0] * _clenshaw(c["kK1Cheb"], np.array([1.0]))[0]) < 2e-15
