"""Drop-in MI355X black-box FMM with the calling convention of flexmm's
``flexmm2d/bbfmm.py`` (class BB_FMM, :200-231, on the generic passes of
``flexmm2d/fmm.py``): ``BB_FMM(x, y, kernel_eval, Ncutoff=50, p=8, dtype,
bbox, helper, verbose)``, ``build_expansions(tau)``,
``source_evaluation(x, y)`` / ``target_evaluation(x, y)`` /
``evaluate_to_points(x, y, name)`` returning a 1-D potential, and
``register_neighbor_evaluator``.

The expansions are Chebyshev interpolants of order p on each box: the
multipole of a box is the Chebyshev moments of its sources, the local
expansion the coefficients of a p x p Chebyshev sum, and the operators
(M2M, L2L, the 40 M2L matrices) follow BB_Helper.prepare and
BB_Precomputation (bbfmm.py:95-166).  They run through the same device tree
and MFMA passes as the KIFMM, with identity pseudo-inverse blocks; only the
S2M (moments) and the L2T (Chebyshev sum) differ.

Differences from the reference:
* ``kernel_eval`` must be a ``flexmm_amd.kernels`` descriptor of the Laplace
  or modified Helmholtz kernel;
* the order is limited to p <= 8 (p^2 <= 64 coefficients per box; 8 is the
  reference's default);
* helpers are kept per (kernel, p); the reference's default ``BB_Helper()``
  instance (bbfmm.py:202) is shared by every BB_FMM whatever its kernel or
  order, so a second kernel would reuse the first one's M2L matrices there.
"""
import time

import numpy as np
import torch

from .. import _lib
from .. import kernels as _kernels
from ..kifmm2d.fmm import (FMM as _KIFMM, Helper, _PRECISIONS, _acquire_plan, _print,
                           _release_plan, _stream, _to_device)
from ..kifmm2d.precompute import LevelOperators, compress_m2l, kernel_matrix, m2l_slots

MAX_ORDER = 8


def cheb_nodes(p):
    """The p Gauss-Chebyshev nodes in ascending order (bbfmm.py:131-132)."""
    k = np.arange(1, p + 1)
    return np.cos((2 * k - 1) * np.pi / (2 * p))[::-1].copy()


def cheb_vander2(x, y, p):
    """(n, p^2): column i p + j holds T_i(x) T_j(y) (chebV2, bbfmm.py:14-24)."""
    xv = np.polynomial.chebyshev.chebvander(np.asarray(x, dtype=np.float64), p - 1)
    yv = np.polynomial.chebyshev.chebvander(np.asarray(y, dtype=np.float64), p - 1)
    return (xv[:, :, None] * yv[:, None, :]).reshape(xv.shape[0], p * p)


class ChebOperators(LevelOperators):
    """FP64 operators of one box width for a Chebyshev order p, in the
    fmm2d_set_operators layout of LevelOperators (p^2 rows):

    * M2MC = chebV2(children's nodes in parent coordinates)^T blockdiag(VI^T)
      (bbfmm.py:138-155), L2LC = M2MC^T;
    * M2L of slot (ox, oy): VI K(nodes w/2 + (ox, oy) w -> nodes w/2) VI^T
      (bbfmm.py:97-117), VI the inverse of the 2-D Chebyshev Vandermonde
      matrix at the nodes;
    * identity pseudo-inverse blocks (flexmm2d/fmm.py:58-61);
    * 2/w (bbfmm.py:119) as the first entry of the circle area."""

    def __init__(self, kernel, order, width):
        p = order
        n = p * p
        self.order = p
        self.p = n
        self.width = float(width)
        nv = cheb_nodes(p)
        nx, ny = np.meshgrid(nv, nv, indexing="ij")
        nx, ny = nx.ravel(), ny.ravel()
        VI = np.linalg.inv(cheb_vander2(nx, ny, p))
        sx, sy = nx / 2, ny / 2
        cx = np.concatenate([sx - 0.5, sx - 0.5, sx + 0.5, sx + 0.5])
        cy = np.concatenate([sy - 0.5, sy + 0.5, sy - 0.5, sy + 0.5])
        W1 = np.zeros((4 * n, 4 * n))
        for i in range(4):
            W1[i * n:(i + 1) * n, i * n:(i + 1) * n] = VI.T
        self.M2MC = cheb_vander2(cx, cy, p).T @ W1
        snx, sny = nx * 0.5 * width, ny * 0.5 * width
        self.m2l = [VI @ kernel_matrix(kernel, snx + ox * width, sny + oy * width, snx, sny) @ VI.T
                    for ox, oy in m2l_slots()]
        # compress(): the M2L_Evaluator rule of flexmm2d/helpers.py:5-33
        # (rank at 1e-14, compressed if 2 rank < p^2), bbfmm.py:99
        self.m2l_svd = [compress_m2l(m, n) for m in self.m2l]
        self.pinv1 = np.eye(n)
        self.pinv2 = np.eye(n)
        self.pinv_rank = n
        self.width_adj = 2.0 / width
        wa = np.zeros(n)
        wa[0] = self.width_adj
        self.large = (wa, np.zeros(n))
        self.small = self.large


class BBHelper(Helper):
    """BB_Helper (bbfmm.py:122-199): operators per round(width, 12) for one
    (kernel, order)."""

    def __init__(self, kernel, order):
        super().__init__(kernel, order * order)
        self.order = order
        self.functions["kernel_apply_single"] = ("neighbor", False)
        self.functions["kernel_apply_single_check"] = ("neighbor", True)

    def device_operators(self, width, device):
        k = self._key(width)
        dk = (k, device)
        t = self._device_ops.get(dk)
        if t is None:
            ops = self.precomputations.get(k)
            if ops is None:
                ops = ChebOperators(self.kernel, self.order, width)
                self.precomputations[k] = ops
            t = torch.from_numpy(ops.pack()).to(device)
            self._device_ops[dk] = t
        return t


the_bb_help = {}


class BB_FMM(_KIFMM):
    """flexmm2d ``BB_FMM`` (bbfmm.py:200-231; passes flexmm2d/fmm.py:33-177) on the GPU."""

    def __init__(self, x, y, kernel_eval, Ncutoff=50, p=8, dtype=float, bbox=None, helper=None,
                 verbose=False, precision="fp64", m2l="dense"):
        kernel = _kernels.resolve_kernel(kernel_eval)
        if kernel.kernel_id not in (_lib.KERNEL_LAPLACE, _lib.KERNEL_MODIFIED_HELMHOLTZ):
            raise NotImplementedError("black-box FMM kernels: Laplace, modified Helmholtz")
        if np.dtype(dtype) != np.float64:
            raise NotImplementedError("dtype must be float64 (the reference computes in float64)")
        if precision not in _PRECISIONS:
            raise ValueError("precision must be one of %s" % sorted(_PRECISIONS))
        p = int(p)
        if p < 2 or p > MAX_ORDER:
            raise NotImplementedError("Chebyshev order p = %d outside [2, %d]" % (p, MAX_ORDER))
        if not torch.cuda.is_available():
            raise RuntimeError("flexmm_amd needs a ROCm GPU; no device is visible")
        self.lib = _lib.load()
        self.kernel = kernel
        self.kernel_eval = kernel_eval
        self.p = p
        self.Nequiv = p * p
        self.Ncutoff = int(Ncutoff)
        self.dtype = dtype
        self.precision = precision
        self.verbose = verbose
        self.print = _print if verbose else (lambda *a: None)
        if helper is None:
            key = (kernel.key, p)
            if key not in the_bb_help:
                the_bb_help[key] = BBHelper(kernel, p)
            helper = the_bb_help[key]
        elif not (isinstance(helper, BBHelper) and helper.kernel is kernel and helper.order == p):
            raise ValueError("helper belongs to another kernel or order")
        self.helper = helper
        self._numpy_io = not isinstance(x, torch.Tensor)
        if isinstance(x, torch.Tensor) and x.device.type == "cuda":
            self.device = x.device
        else:
            self.device = torch.device("cuda", torch.cuda.current_device())
        self.x = _to_device(x, self.device, "x")
        self.y = _to_device(y, self.device, "y")
        if self.x.numel() != self.y.numel():
            raise ValueError("x and y differ in length")
        if bbox is None:
            mm = torch.stack([self.x.min(), self.x.max(), self.y.min(), self.y.max()]).cpu().numpy()
            self.bbox = self.helper.get_bbox(*[float(v) for v in mm], None)
        else:
            self.bbox = self.helper.get_bbox(None, None, None, None, [float(b) for b in bbox])
        self._dev_key = str(self.device)
        self._prec = _PRECISIONS[precision]
        self.m2l = m2l
        self._plan = _acquire_plan(self.lib, kernel, self.Nequiv, self._prec, self._dev_key,
                                   cheb=p, m2l=m2l)
        self.tau = None
        self.neighbor_evaluators = {}
        self.build_tree()
        self.register_neighbor_evaluator(self.helper.functions["kernel_apply_single"],
                                         "neighbor_potential_target_evaluation")
        self.register_neighbor_evaluator(self.helper.functions["kernel_apply_single_check"],
                                         "neighbor_potential_source_evaluation")
        self.prepare()

    def close(self):
        plan = getattr(self, "_plan", None)
        if plan is not None:
            self._plan = None
            _release_plan(plan, self.kernel, self.Nequiv, self._prec, self._dev_key, cheb=self.p)

    # -- passes --------------------------------------------------------------
    def upwards_pass(self, tau):
        """flexmm2d/fmm.py:62-97 with the Chebyshev S2M (bbfmm.py:173-184, 210-216)."""
        t = _to_device(tau, self.device, "tau")
        n = self.x.numel()
        if t.numel() != n:
            raise ValueError("tau has %d entries for %d sources" % (t.numel(), n))
        st = time.time()
        self.tau = t
        _lib.check(self.lib.fmm2d_upward(self.handle, t.data_ptr(), n, _stream()))
        if self.verbose:
            torch.cuda.synchronize()
        self.print("....Time for upwards pass:      {:0.2f}".format(1000 * (time.time() - st)))

    def build_expansions(self, tau):
        """flexmm2d/fmm.py:127-129."""
        self.upwards_pass(tau)
        self.downwards_pass()

    def register_neighbor_evaluator(self, kernel_apply_single, name):
        """flexmm2d/fmm.py:131-132: the P2P part of an evaluation.  The device
        evaluates the kernel's own potential, with (helper.functions
        'kernel_apply_single_check') or without ('kernel_apply_single')
        self-exclusion."""
        f = kernel_apply_single
        if isinstance(f, tuple) and len(f) == 2 and f[0] == "neighbor":
            check = bool(f[1])
        elif f is self.kernel or (isinstance(f, _kernels.Evaluator) and f.kernel is self.kernel
                                  and f.evaluator_id == _lib.EVAL_POTENTIAL):
            check = False
        else:
            raise NotImplementedError("neighbor evaluator %r has no device implementation" % (f,))
        self.neighbor_evaluators.setdefault(name, check)

    def source_evaluation(self, x, y):
        return self.evaluate_to_points(x, y, "neighbor_potential_source_evaluation")

    def target_evaluation(self, x, y):
        return self.evaluate_to_points(x, y, "neighbor_potential_target_evaluation")

    def evaluate_to_points(self, x, y, name):
        """flexmm2d/fmm.py:138-152: the Chebyshev local expansion plus the
        named neighbor evaluation; a 1-D potential in target input order."""
        check = self.neighbor_evaluators[name]
        numpy_out = not isinstance(x, torch.Tensor)
        tx = _to_device(x, self.device, "x")
        ty = _to_device(y, self.device, "y")
        if tx.numel() != ty.numel():
            raise ValueError("x and y differ in length")
        out = torch.empty(tx.numel(), dtype=torch.float64, device=self.device)
        _lib.check(self.lib.fmm2d_evaluate(self.handle, tx.data_ptr(), ty.data_ptr(), tx.numel(),
                                           _lib.EVAL_POTENTIAL, 1 if check else 0,
                                           out.data_ptr(), _stream()))
        if numpy_out:
            return out.cpu().numpy()
        return out
