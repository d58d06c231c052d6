"""Kernel registry: the GPU stand-ins for the reference's user kernel functions.

The reference accepts any numba-jitted ``kernel_eval(sx, sy, tx, ty)`` and
inlines it into its loops (flexmm/kifmm2d/scalar/fmm.py:38,
helpers.py:115-153).  A GPU evaluator cannot compile arbitrary Python, so
``kernel_eval`` here must be one of the descriptors below.  Each is still a
plain callable with the reference's argument convention (FP64, NumPy
broadcasting; ``dx = tx - sx``), used on the host for the per-width
precomputation, and carries the id of its hand-written device code.
Anything else passed as ``kernel_eval`` raises ``NotImplementedError``.
"""
import numpy as np

from . import _lib

FOUR_PI = 4.0 * np.pi
TWO_PI = 2.0 * np.pi


class Kernel:
    """A scalar kernel K(s, t) with a device implementation."""

    def __init__(self, name, kernel_id, params, fn, doc):
        self.name = name
        self.kernel_id = kernel_id
        self.params = tuple(float(p) for p in params)
        self._fn = fn
        self.__doc__ = doc

    def __call__(self, sx, sy, tx, ty):
        return self._fn(np.asarray(sx, dtype=np.float64), np.asarray(sy, dtype=np.float64),
                        np.asarray(tx, dtype=np.float64), np.asarray(ty, dtype=np.float64))

    @property
    def key(self):
        return (self.kernel_id, self.params)

    def __repr__(self):
        return "Kernel(%s%s)" % (self.name, "" if not self.params else ", %s" % (self.params,))


class Evaluator:
    """A named evaluator (the pair KAS1/KAS2 of register_evaluator,
    helpers.py:176-238): which outputs the device L2T/P2P pass produces."""

    def __init__(self, name, kernel, evaluator_id, outputs, fn, doc):
        self.name = name
        self.kernel = kernel
        self.evaluator_id = evaluator_id
        self.outputs = outputs
        self._fn = fn
        self.__doc__ = doc

    def __call__(self, sx, sy, tx, ty):
        return self._fn(np.asarray(sx, dtype=np.float64), np.asarray(sy, dtype=np.float64),
                        np.asarray(tx, dtype=np.float64), np.asarray(ty, dtype=np.float64))

    def __repr__(self):
        return "Evaluator(%s, outputs=%d)" % (self.name, self.outputs)


def _laplace(sx, sy, tx, ty):
    dx = tx - sx
    dy = ty - sy
    return -np.log(dx * dx + dy * dy) / FOUR_PI


def _laplace_grad(sx, sy, tx, ty):
    dx = tx - sx
    dy = ty - sy
    d2 = dx * dx + dy * dy
    return (-np.log(d2) / FOUR_PI, -dx / (TWO_PI * d2), -dy / (TWO_PI * d2))


laplace = Kernel(
    "laplace", _lib.KERNEL_LAPLACE, (), _laplace,
    "Laplace kernel -log(|t-s|^2)/(4 pi) (examples/pure_kifmm_laplace.py:24-29).")

laplace_potential = Evaluator(
    "laplace_potential", laplace, _lib.EVAL_POTENTIAL, 1, lambda *a: (_laplace(*a),),
    "Potential u (the default 'potential_evaluation', fmm.py:63).")

laplace_gradient = Evaluator(
    "laplace_gradient", laplace, _lib.EVAL_GRADIENT, 3, _laplace_grad,
    "(u, du/dx_t, du/dy_t) (examples/pure_kifmm_laplace_gradient.py:32-55).")

class SourceKernel:
    """An upwardor kernel KAS(sx, sy, tx, ty, tau, extras) -> (value,) whose
    sources carry several densities and per-source extras (register_upwardor,
    helpers.py:155-174), with a device implementation."""

    def __init__(self, name, kernel, ntau, nextras, fn, doc):
        self.name = name
        self.kernel = kernel
        self.ntau = ntau
        self.nextras = nextras
        self._fn = fn
        self.__doc__ = doc

    def __call__(self, sx, sy, tx, ty, tau, extras):
        return self._fn(np.asarray(sx, dtype=np.float64), np.asarray(sy, dtype=np.float64),
                        np.asarray(tx, dtype=np.float64), np.asarray(ty, dtype=np.float64),
                        np.asarray(tau, dtype=np.float64), np.asarray(extras, dtype=np.float64))

    def __repr__(self):
        return "SourceKernel(%s)" % self.name


def _laplace_slp_dlp(sx, sy, tx, ty, tau, extras):
    dx = tx - sx
    dy = ty - sy
    d2 = dx * dx + dy * dy
    return (-np.log(d2) / FOUR_PI * tau[0]
            + (extras[0] * dx + extras[1] * dy) / d2 / TWO_PI * tau[1],)


laplace_slp_dlp = SourceKernel(
    "laplace_slp_dlp", laplace, 2, 2, _laplace_slp_dlp,
    "Single + double layer Laplace sources, -log(d^2)/(4 pi) tau[0] + (n.d)/d^2/(2 pi) tau[1], "
    "d = t - s, n = (extras[0], extras[1]) (Laplace_Both_Eval, "
    "examples/pure_kifmm_laplace_dipoles.py:32-40).")

laplace_slp_dlp_potential = Evaluator(
    "laplace_slp_dlp_potential", laplace, _lib.EVAL_DIPOLE, 1,
    lambda *a: (_laplace(*a),),
    "Potential of single + double layer sources: L2T with the single-layer "
    "kernel on the local density, P2P with laplace_slp_dlp (the 'slp+dlp->u' "
    "evaluator of examples/pure_kifmm_laplace_dipoles.py:140).")


def _stokeslet_u(sx, sy, tx, ty):
    """First row of the Stokeslet for a unit x-force (stokes_fmm.py:33-61)."""
    dx = tx - sx
    dy = ty - sy
    d2 = dx * dx + dy * dy
    return (-0.5 * np.log(d2) + dx * dx / d2) / FOUR_PI


stokeslet = Kernel(
    "stokeslet", _lib.KERNEL_STOKESLET, (), _stokeslet_u,
    "2D Stokeslet (-log r I + r r^T / r^2) / (4 pi) and pressure r / (2 pi r^2) "
    "(kifmm2d/vector/stokes_fmm.py:33-61); used through flexmm_amd.kifmm2d.StokesFMM.")

KERNELS = {k.name: k for k in (laplace,)}
EVALUATORS = {e.name: e for e in (laplace_potential, laplace_gradient,
                                   laplace_slp_dlp_potential)}
SOURCE_KERNELS = {k.name: k for k in (laplace_slp_dlp,)}

_MH = {}


def modified_helmholtz(k):
    """Modified-Helmholtz kernel K0(k |t-s|)/(2 pi), k > 0, with its potential
    and gradient evaluators (examples/pure_kifmm_modified_helmholtz_gradient.py:
    36-68; the device evaluates K0/K1 in FP64, the example through a 1e-8
    interpolant).  Returns (kernel, potential_evaluator, gradient_evaluator);
    the same k returns the same objects."""
    k = float(k)
    if not k > 0.0:
        raise ValueError("modified Helmholtz needs k > 0, got %r" % k)
    if k in _MH:
        return _MH[k]
    from scipy import special

    def mh(sx, sy, tx, ty):
        dx = tx - sx
        dy = ty - sy
        return special.k0(k * np.sqrt(dx * dx + dy * dy)) / TWO_PI

    def mh_grad(sx, sy, tx, ty):
        dx = tx - sx
        dy = ty - sy
        r = np.sqrt(dx * dx + dy * dy)
        g = -k * special.k1(k * r) / (TWO_PI * r)
        return (special.k0(k * r) / TWO_PI, g * dx, g * dy)

    ker = Kernel("modified_helmholtz", _lib.KERNEL_MODIFIED_HELMHOLTZ, (k,), mh,
                 "K0(%g |t-s|)/(2 pi)" % k)
    pot = Evaluator("modified_helmholtz_potential", ker, _lib.EVAL_POTENTIAL, 1,
                    lambda *a: (mh(*a),), "Potential u.")
    grad = Evaluator("modified_helmholtz_gradient", ker, _lib.EVAL_GRADIENT, 3, mh_grad,
                     "(u, du/dx_t, du/dy_t).")
    _MH[k] = (ker, pot, grad)
    EVALUATORS[pot.name + "(%g)" % k] = pot
    EVALUATORS[grad.name + "(%g)" % k] = grad
    return _MH[k]


def resolve_kernel(kernel_eval):
    if isinstance(kernel_eval, Kernel):
        return kernel_eval
    if isinstance(kernel_eval, str) and kernel_eval in KERNELS:
        return KERNELS[kernel_eval]
    raise NotImplementedError(
        "kernel_eval=%r has no device implementation; pass one of flexmm_amd.kernels (%s)"
        % (kernel_eval, ", ".join(sorted(KERNELS))))


def gradient_evaluator(kernel):
    """The (u, du/dx_t, du/dy_t) evaluator of a kernel."""
    for e in EVALUATORS.values():
        if e.kernel is kernel and e.evaluator_id == _lib.EVAL_GRADIENT:
            return e
    raise NotImplementedError("no gradient evaluator for %r" % (kernel,))


def default_evaluator(kernel):
    for e in EVALUATORS.values():
        if e.kernel is kernel and e.evaluator_id == _lib.EVAL_POTENTIAL:
            return e
    raise NotImplementedError("no potential evaluator for %r" % (kernel,))


def resolve_evaluator(kernel, KAS1, KAS2, outputs):
    """Map register_evaluator's (KAS1, KAS2, outputs) to a device evaluator.
    KAS2 may be a SourceKernel (P2P with the sources' own densities and
    extras); KAS1 is then the kernel's potential evaluator (the L2T kernel)."""
    if isinstance(KAS2, SourceKernel):
        if KAS2.kernel is not kernel:
            raise ValueError("source kernel %s belongs to %s, not %s" % (KAS2.name, KAS2.kernel.name, kernel.name))
        if KAS1 is not None and KAS1 is not kernel and not (
                isinstance(KAS1, Evaluator) and KAS1.kernel is kernel
                and KAS1.evaluator_id == _lib.EVAL_POTENTIAL):
            raise NotImplementedError("with a source kernel, KAS1 must be the kernel's potential "
                                      "evaluator (the L2T kernel), got %r" % (KAS1,))
        if outputs != 1:
            raise ValueError("%s produces 1 output, not %d" % (KAS2.name, outputs))
        for e in EVALUATORS.values():
            if e.kernel is kernel and e.evaluator_id == _lib.EVAL_DIPOLE:
                return e
        raise NotImplementedError("no device evaluator for %r" % (KAS2,))
    cands = [e for e in (KAS1, KAS2) if e is not None]
    if not cands or not all(isinstance(e, Evaluator) for e in cands):
        raise NotImplementedError(
            "register_evaluator needs flexmm_amd.kernels evaluators (%s), got %r"
            % (", ".join(sorted(EVALUATORS)), (KAS1, KAS2)))
    ev = cands[0]
    if any(e is not ev for e in cands):
        raise NotImplementedError("KAS1 and KAS2 must be the same device evaluator")
    if ev.kernel is not kernel:
        raise ValueError("evaluator %s belongs to kernel %s, not %s" % (ev.name, ev.kernel.name, kernel.name))
    if outputs != ev.outputs:
        raise ValueError("evaluator %s produces %d outputs, not %d" % (ev.name, ev.outputs, outputs))
    return ev
