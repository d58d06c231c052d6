"""Drop-in MI355X Stokeslet FMM with the calling convention of flexmm's
``kifmm2d/vector/stokes_fmm.py`` (class FMM, :598-842): ``FMM(x, y, Nequiv,
Ncutoff, bbox, precomputations, verbose)``, ``precompute()``,
``build_expansions(taux, tauy)`` and ``evaluate_to_points(x, y, check_self)``
returning the velocity and pressure ``(u, v, p)``.

The reference runs a vector KIFMM with (2p+1)-long bordered densities
(stokes_fmm.py:473-596).  Here the far field is written as three harmonic
fields (sums of log, d log and d d log terms of the forces, and of the
double-layer densities when present), run through the scalar MFMA passes on
one quadtree as Laplace equivalent densities, and the near field evaluates
the Stokeslet and stresslet directly (the identity is spelled out in
flexmm_amd/csrc/fmm.hip).
Both approximate the same direct sum; this form reuses the Laplace
operators and reaches the Laplace FMM's accuracy.
"""
import sys
import time

import numpy as np
import torch

from .. import _lib
from .. import kernels as _kernels
from .fmm import (Helper, TreeInfo, _PRECISIONS, _acquire_plan, _release_plan, _stream,
                  _to_device, padded_dim, the_help)


class FMM:
    """stokes_fmm.FMM (stokes_fmm.py:598-842) on the GPU."""

    def __init__(self, x, y, Nequiv=48, Ncutoff=50, bbox=None, precomputations=None,
                 verbose=False, precision="fp64", m2l="dense"):
        if precision not in _PRECISIONS:
            raise ValueError("precision must be one of %s" % sorted(_PRECISIONS))
        if not torch.cuda.is_available():
            raise RuntimeError("flexmm_amd needs a ROCm GPU; no device is visible")
        self.lib = _lib.load()
        self.Nequiv = int(Nequiv)
        self.Ncutoff = int(Ncutoff)
        padded_dim(self.Nequiv)
        self.verbose = verbose
        self.print = (lambda *a: (print(*a), sys.stdout.flush())) if verbose else (lambda *a: None)
        self.precision = precision
        self.input_precomputations = precomputations
        key = (_kernels.laplace.key, self.Nequiv, "float64")
        if key not in the_help:
            the_help[key] = Helper(_kernels.laplace, self.Nequiv)
        self.helper = the_help[key]
        if isinstance(x, torch.Tensor) and x.device.type == "cuda":
            self.device = x.device
        else:
            self.device = torch.device("cuda", torch.cuda.current_device())
        self._numpy_io = not isinstance(x, torch.Tensor)
        self.x = _to_device(x, self.device, "x")
        self.y = _to_device(y, self.device, "y")
        if self.x.numel() != self.y.numel():
            raise ValueError("x and y differ in length")
        self.bbox = bbox
        self.adjust_bbox()
        self._dev_key = str(self.device)
        self._prec = _PRECISIONS[precision]
        self.m2l = m2l
        self._plan = _acquire_plan(self.lib, _kernels.stokeslet, self.Nequiv, self._prec,
                                   self._dev_key, m2l=m2l)
        self._prepared = False
        self.build_tree()

    def adjust_bbox(self):
        """stokes_fmm.py:610-641: a square root box; with stored precomputations
        its width snaps to a power-of-two multiple of their smallest width.
        Without a bbox the reference's Tree cannot start (tree.py:473-476);
        here the sources' square bounding box is used."""
        if self.input_precomputations is not None:
            for w in self.input_precomputations:
                self.helper.register_width(w)
        mm = torch.stack([self.x.min(), self.x.max(), self.y.min(), self.y.max()]).cpu().numpy()
        if self.bbox is None or self.input_precomputations is not None:
            self.bbox = self.helper.get_bbox(*[float(v) for v in mm],
                                             None if self.bbox is None else list(self.bbox))
        self.bbox = [float(b) for b in self.bbox]

    def close(self):
        plan = getattr(self, "_plan", None)
        if plan is not None:
            self._plan = None
            _release_plan(plan, _kernels.stokeslet, self.Nequiv, self._prec, self._dev_key)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._plan.h

    def set_profiling(self, enable=True):
        """Record HIP events around each phase on the current stream."""
        _lib.check(self.lib.fmm2d_set_profiling(self.handle, 1 if enable else 0))

    def phase_times(self):
        """Milliseconds of the last recorded run of each phase (waits for it)."""
        out = (_lib.ctypes.c_double * len(_lib.PHASES))()
        _lib.check(self.lib.fmm2d_phase_times(self.handle, out, len(out)))
        return {k: float(v) for k, v in zip(_lib.PHASES, out) if v >= 0}

    def build_tree(self):
        """stokes_fmm.py:642-646."""
        st = time.time()
        bb = (_lib.ctypes.c_double * 4)(*self.bbox)
        _lib.check(self.lib.fmm2d_tree_build(self.handle, self.x.data_ptr(), self.y.data_ptr(),
                                             self.x.numel(), bb, self.Ncutoff, _stream()))
        info = (_lib.ctypes.c_int64 * (9 + 3 * 32))()
        _lib.check(self.lib.fmm2d_tree_info(self.handle, info, len(info)))
        self.tree = TreeInfo(list(info), self.bbox)
        self.print("....Tree formed in:             {:0.1f}".format((time.time() - st) * 1000))

    def precompute(self):
        """stokes_fmm.py:647-659: operators for every level width (cached)."""
        s = _stream()
        n = self.lib.fmm2d_operator_size(self.Nequiv)
        for d in range(2, self.tree.levels):
            w = self.tree.widths[d]
            self.helper.register_width(w)
            if self._plan.level_width.get(d) == w:
                continue
            ops = self.helper.device_operators(w, self.device)
            _lib.check(self.lib.fmm2d_set_operators(self.handle, d, ops.data_ptr(), n, s))
            self._plan.level_width[d] = w
        self.precomputations = {w: None for w in self.tree.widths}
        self._prepared = True

    def build_expansions(self, taux, tauy):
        """stokes_fmm.py:660-737: forces (taux, tauy) in input order."""
        if not self._prepared:
            self.precompute()
        fx = _to_device(taux, self.device, "taux")
        fy = _to_device(tauy, self.device, "tauy")
        n = self.x.numel()
        if fx.numel() != n or fy.numel() != n:
            raise ValueError("forces must have one entry per source")
        self.taux, self.tauy = fx, fy
        tau = torch.cat([fx, fy])
        st = time.time()
        _lib.check(self.lib.fmm2d_upward(self.handle, tau.data_ptr(), n, _stream()))
        _lib.check(self.lib.fmm2d_downward(self.handle, _stream()))
        self._tau = tau
        self.print("....Time for expansions:        {:0.2f}".format(1000 * (time.time() - st)))

    def build_expansions_dipole(self, taux, tauy, sigmax, sigmay, dipx, dipy):
        """stokes_fmm.py:738-826: forces (taux, tauy), double-layer densities
        (sigmax, sigmay) and normals (dipx, dipy), all in input order.
        Evaluate with evaluate_to_points_dipole."""
        if not self._prepared:
            self.precompute()
        n = self.x.numel()
        arrs = []
        for name, a in (("taux", taux), ("tauy", tauy), ("sigmax", sigmax), ("sigmay", sigmay),
                        ("dipx", dipx), ("dipy", dipy)):
            t = _to_device(a, self.device, name)
            if t.numel() != n:
                raise ValueError("%s must have one entry per source" % name)
            arrs.append(t)
        self.taux, self.tauy, self.sigmax, self.sigmay, self.dipx, self.dipy = arrs
        tau = torch.cat(arrs[:4])
        dvec = torch.cat(arrs[4:])
        st = time.time()
        _lib.check(self.lib.fmm2d_upward_dipole(self.handle, tau.data_ptr(), dvec.data_ptr(), n,
                                                _stream()))
        _lib.check(self.lib.fmm2d_downward(self.handle, _stream()))
        self._tau, self._dvec = tau, dvec
        self.print("....Time for expansions:        {:0.2f}".format(1000 * (time.time() - st)))

    def evaluate_to_points_dipole(self, x, y, check_self=False):
        """stokes_fmm.py:844-858: (u, v, p) of the Stokeslets and stresslets
        given to build_expansions_dipole."""
        return self._evaluate(x, y, check_self, _lib.EVAL_STOKES_DIPOLE)

    def evaluate_to_points(self, x, y, check_self=False):
        """stokes_fmm.py:827-840: returns (u, v, p) at the points."""
        return self._evaluate(x, y, check_self, _lib.EVAL_STOKES)

    def _evaluate(self, x, y, check_self, evaluator):
        numpy_out = not isinstance(x, torch.Tensor)
        tx = _to_device(x, self.device, "x")
        ty = _to_device(y, self.device, "y")
        if tx.numel() != ty.numel():
            raise ValueError("x and y differ in length")
        out = torch.empty((3, tx.numel()), dtype=torch.float64, device=self.device)
        _lib.check(self.lib.fmm2d_evaluate(self.handle, tx.data_ptr(), ty.data_ptr(), tx.numel(),
                                           evaluator, 1 if check_self else 0,
                                           out.data_ptr(), _stream()))
        if numpy_out:
            out = out.cpu().numpy()
        return out[0], out[1], out[2]
