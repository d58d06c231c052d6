"""Vectorised FP64 kernel functions for the oracle -- TEST INFRASTRUCTURE ONLY.

Each kernel is ``K(sx, sy, tx, ty)`` with NumPy broadcasting, the same
argument convention as the user kernels of the reference examples
(``examples/pure_kifmm_laplace.py:24-29``: ``dx = tx - sx``).

Pinned by: closed forms (``tests/test_oracle_kernels.py``) -- the Laplace
kernel against ``-log|x-y|/(2 pi)``, gradients against central finite
differences, the modified-Helmholtz kernel against ``scipy.special.k0/k1``,
the double-layer term against the finite-difference derivative of the
single layer with respect to the source position.
"""
import numpy as np
from scipy import special

FOUR_PI = 4.0 * np.pi
TWO_PI = 2.0 * np.pi


def laplace(sx, sy, tx, ty):
    """-log(d^2)/(4 pi); examples/pure_kifmm_laplace.py:24-29."""
    dx = tx - sx
    dy = ty - sy
    d2 = dx * dx + dy * dy
    return -np.log(d2) / FOUR_PI


def laplace_gradient(sx, sy, tx, ty):
    """(u, u_x, u_y) per unit charge; examples/pure_kifmm_laplace_gradient.py:32-55.

    Derivatives are with respect to the target coordinates.
    """
    dx = tx - sx
    dy = ty - sy
    d2 = dx * dx + dy * dy
    id2 = 1.0 / d2
    scale = -1.0 / TWO_PI
    return (-np.log(d2) / FOUR_PI, scale * dx * id2, scale * dy * id2)


def laplace_slp_dlp(sx, sy, tx, ty, tau, extras):
    """Single plus double layer: -log(d^2)/(4 pi) tau[0] + (n.d)/d^2/(2 pi) tau[1]
    with n = (extras[0], extras[1]) and d = t - s; the upwardor/evaluator
    kernel Laplace_Both_Eval of examples/pure_kifmm_laplace_dipoles.py:32-40.
    Returns a 1-tuple like the reference's KAS."""
    dx = tx - sx
    dy = ty - sy
    d2 = dx * dx + dy * dy
    return (-np.log(d2) / FOUR_PI * tau[0] + (extras[0] * dx + extras[1] * dy) / d2 / TWO_PI * tau[1],)


def make_modified_helmholtz(k):
    """K0(k r)/(2 pi) and its target gradient; the kernel of
    examples/pure_kifmm_modified_helmholtz_gradient.py:25-68 (which evaluates
    K0/K1 through the external ``function_generator`` package)."""
    k = float(k)

    def mh(sx, sy, tx, ty):
        dx = tx - sx
        dy = ty - sy
        r = np.sqrt(dx * dx + dy * dy)
        return special.k0(k * r) / TWO_PI

    def mh_gradient(sx, sy, tx, ty):
        dx = tx - sx
        dy = ty - sy
        r = np.sqrt(dx * dx + dy * dy)
        kr = k * r
        g = -k * special.k1(kr) / (TWO_PI * r)
        return (special.k0(kr) / TWO_PI, g * dx, g * dy)

    return mh, mh_gradient
