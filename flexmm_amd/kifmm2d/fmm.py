"""Drop-in MI355X KIFMM with the calling convention of flexmm's kifmm2d.

Mirrors ``flexmm/kifmm2d/scalar/fmm.py`` (class FMM, :37-199): the same
constructor, ``build_expansions``, ``source_evaluation``,
``target_evaluation``, ``evaluate_to_points``, ``register_evaluator`` and
``register_upwardor``, and the same per-(kernel, Nequiv, dtype) helper cache
(``the_help``, :11 and :40-43) whose cached box widths snap later bounding
boxes (helpers.py:68-88).  Every step runs in libflexmm_amd on the GPU; the
host only computes per-width FP64 operators once (precompute.py).

Differences from the reference:
* ``kernel_eval`` must be a ``flexmm_amd.kernels`` descriptor (a device
  kernel cannot be compiled from an arbitrary numba function);
* inputs may be NumPy arrays or torch tensors; results come back as NumPy
  arrays for NumPy inputs and as device tensors otherwise;
* two live FMMs never share mutable per-level state (the reference's
  Helper.prepare overwrites helper-wide surface lists, helpers.py:95-108).
"""
import math
import sys
import time

import numpy as np
import torch

from .. import _lib
from .. import kernels as _kernels
from .precompute import LevelOperators, padded_dim

the_help = {}


def _print(*args):
    print(*args)
    sys.stdout.flush()


class Helper:
    """Per-(kernel, Nequiv, dtype) state shared by FMMs (helpers.py:47-114):
    precomputations keyed by round(width, 12) (float_dict.py:23-24) and the
    registered evaluators/upwardors."""

    def __init__(self, kernel, p):
        self.kernel = kernel
        self.p = p
        self.precomputations = {}
        self._device_ops = {}
        self.evaluators = {}
        self.upwardors = {}
        # the reference's jitted helper functions (helpers.py:115-128); here
        # 'kernel_app' (the L2T kernel of register_evaluator) is the kernel's
        # potential evaluator
        self.functions = {}
        try:
            self.functions["kernel_app"] = _kernels.default_evaluator(kernel)
        except NotImplementedError:
            pass

    @staticmethod
    def _key(width):
        return round(float(width), 12)

    def collect_widths(self):
        return np.array(list(self.precomputations.keys()), dtype=float)

    def has_precomputation(self, width):
        return self._key(width) in self.precomputations

    def get_bbox(self, px_min, px_max, py_min, py_max, bbox):
        """helpers.py:68-88."""
        if bbox is None:
            bbox = [px_min, px_max, py_min, py_max]
        if self.precomputations:
            small_width = min(self.precomputations)   # the keys are the widths
            bw = max(bbox[1] - bbox[0], bbox[3] - bbox[2])
            width = 2 ** np.ceil(np.log2(bw / small_width)) * small_width
        else:
            width = max(bbox[1] - bbox[0], bbox[3] - bbox[2])
        cx = (bbox[1] + bbox[0]) / 2
        cy = (bbox[3] + bbox[2]) / 2
        return [cx - width / 2, cx + width / 2, cy - width / 2, cy + width / 2]

    def register_width(self, width):
        k = self._key(width)
        if k not in self.precomputations:
            self.precomputations[k] = None   # operators are formed on first use

    def device_operators(self, width, device):
        k = self._key(width)
        dk = (k, device)
        t = self._device_ops.get(dk)
        if t is None:
            ops = self.precomputations.get(k)
            if ops is None:
                ops = LevelOperators(self.kernel, self.p, width)
                self.precomputations[k] = ops
            t = torch.from_numpy(ops.pack()).to(device)
            self._device_ops[dk] = t
        return t


class _Plan:
    """Owner of one libflexmm_amd plan handle."""

    def __init__(self, lib, kernel, p, precision, cheb=0):
        h = _lib.ctypes.c_void_p()
        params = (_lib.ctypes.c_double * max(1, len(kernel.params)))(*kernel.params) if kernel.params else None
        if cheb:
            # black-box plan of Chebyshev order cheb (p = cheb^2 coefficients)
            _lib.check(lib.fmm2d_plan_create_cheb(kernel.kernel_id, params, len(kernel.params),
                                                  cheb, precision, _lib.ctypes.byref(h)))
        else:
            _lib.check(lib.fmm2d_plan_create(kernel.kernel_id, params, len(kernel.params), p,
                                             precision, _lib.ctypes.byref(h)))
        self.lib = lib
        self.h = h
        self.level_width = {}    # level -> width whose operators are loaded

    def close(self):
        if self.h:
            self.lib.fmm2d_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# plans are pooled per (kernel, p, precision, device) so that constructing an
# FMM reuses the device buffers sized by earlier trees
_plan_pool = {}


_M2L_MODES = {"svd": _lib.M2L_SVD, "dense": _lib.M2L_DENSE}


def _m2l_mode(m2l):
    if m2l not in _M2L_MODES:
        raise ValueError("m2l must be one of %s" % sorted(_M2L_MODES))
    return _M2L_MODES[m2l]


def _acquire_plan(lib, kernel, p, precision, device, cheb=0, m2l="dense"):
    key = (kernel.key, p, precision, device, cheb)
    pool = _plan_pool.setdefault(key, [])
    plan = None
    while pool:
        cand = pool.pop()
        # a plan finalised by the cyclic garbage collector together with the
        # FMM that returned it to the pool is no longer usable
        if cand.h:
            plan = cand
            break
    if plan is None:
        plan = _Plan(lib, kernel, p, precision, cheb)
    _lib.check(lib.fmm2d_set_m2l_mode(plan.h, _m2l_mode(m2l)))
    return plan


def _release_plan(plan, kernel, p, precision, device, cheb=0):
    key = (kernel.key, p, precision, device, cheb)
    _plan_pool.setdefault(key, []).append(plan)


_PRECISIONS = {"fp64": _lib.PREC_FP64, "mixed": _lib.PREC_MIXED,
               "mixed_m2l": _lib.PREC_MIXED_M2L}


# the current device's current stream as a raw handle: the direct accessor
# costs a fraction of building a torch.cuda.Stream object (~4 us per call)
_raw_stream = getattr(torch._C, "_cuda_getCurrentRawStream", None)
_raw_device = getattr(torch._C, "_cuda_getDevice", None)


def _stream():
    if _raw_stream is not None and _raw_device is not None:
        return _lib.ctypes.c_void_p(_raw_stream(_raw_device()))
    return _lib.ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _to_device(a, device, name):
    if (isinstance(a, torch.Tensor) and a.is_cuda and a.dtype == torch.float64 and a.dim() == 1
            and a.is_contiguous()):
        return a.detach()
    if isinstance(a, torch.Tensor):
        t = a.detach()
        if t.device.type != "cuda":
            t = t.to(device)
        t = t.to(torch.float64).reshape(-1)
    else:
        t = torch.as_tensor(np.ascontiguousarray(np.asarray(a, dtype=np.float64)).reshape(-1),
                            device=device)
    if not t.is_contiguous():
        t = t.contiguous()
    return t


class TreeInfo:
    """Host view of the device tree: per-level node counts, widths and row
    ranges (the analogue of Tree.Levels, tree.py:275-308)."""

    def __init__(self, info, bbox):
        self.levels = int(info[0])
        self.n_nodes = int(info[1])
        self.n_rows = int(info[2])
        self.n_leaves = int(info[3])
        self.n_s2m = int(info[4])
        self.N = int(info[5])
        self.PD = int(info[6])
        self.points_per_leaf = int(info[7])
        L = self.levels
        self.node_off = [int(info[8 + 3 * d]) for d in range(L)]
        self.level_nodes = [int(info[9 + 3 * d]) for d in range(L)]
        self.row_off = [int(info[10 + 3 * d]) for d in range(L)] + [int(info[8 + 3 * L])]
        self.bbox = list(bbox)
        w0 = bbox[1] - bbox[0]
        self.widths = [math.ldexp(w0, -d) for d in range(L)]


class FMM:
    """flexmm kifmm2d ``FMM`` (fmm.py:37-199) on the GPU."""

    def __init__(self, x, y, kernel_eval, Ncutoff, Nequiv, dtype=float, bbox=None, verbose=False,
                 precision="fp64", m2l="dense"):
        kernel = _kernels.resolve_kernel(kernel_eval)
        if np.dtype(dtype) != np.float64:
            raise NotImplementedError("dtype must be float64 (the reference computes in float64)")
        if precision not in _PRECISIONS:
            raise ValueError("precision must be one of %s" % sorted(_PRECISIONS))
        if not torch.cuda.is_available():
            raise RuntimeError("flexmm_amd needs a ROCm GPU; no device is visible")
        self.lib = _lib.load()
        self.kernel = kernel
        self.kernel_eval = kernel_eval
        self.Nequiv = int(Nequiv)
        self.Ncutoff = int(Ncutoff)
        self.dtype = dtype
        self.precision = precision
        self.verbose = verbose
        self.print = _print if verbose else (lambda *a: None)
        padded_dim(self.Nequiv)
        self.helper_key = (kernel.key, self.Nequiv, "float64")
        if self.helper_key not in the_help:
            the_help[self.helper_key] = Helper(kernel, self.Nequiv)
        self.helper = the_help[self.helper_key]
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
        _m2l_mode(m2l)
        self.m2l = m2l
        self._plan = _acquire_plan(self.lib, kernel, self.Nequiv, self._prec, self._dev_key,
                                   m2l=m2l)
        self.tau = None
        self.build_tree()
        self.register_evaluator(_kernels.default_evaluator(kernel), None, "potential_evaluation", 1)
        self.register_upwardor(kernel, "slp")
        self.prepare()

    # -- lifecycle -----------------------------------------------------------
    def close(self):
        """Return the device plan to the pool (also done on garbage collection)."""
        plan = getattr(self, "_plan", None)
        if plan is not None:
            self._plan = None
            _release_plan(plan, self.kernel, self.Nequiv, self._prec, self._dev_key)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._plan.h

    # -- tree ----------------------------------------------------------------
    def build_tree(self):
        """fmm.py:68-72."""
        st = time.time()
        bb = (_lib.ctypes.c_double * 4)(*self.bbox)
        _lib.check(self.lib.fmm2d_tree_build(self.handle, self.x.data_ptr(), self.y.data_ptr(),
                                             self.x.numel(), bb, self.Ncutoff, _stream()))
        info = (_lib.ctypes.c_int64 * (9 + 3 * 32))()
        _lib.check(self.lib.fmm2d_tree_info(self.handle, info, len(info)))
        self.tree = TreeInfo(info[:], self.bbox)
        if self.verbose:
            self.print("....Tree formed in:             {:0.1f}".format((time.time() - st) * 1000))

    def prepare(self):
        """Helper.prepare (helpers.py:93-108): operators of every level width."""
        lw, widths = self._plan.level_width, self.tree.widths
        if (all(lw.get(d) == widths[d] for d in range(2, self.tree.levels))
                and all(self.helper.has_precomputation(w) for w in widths)):
            return   # every width registered and every level's operators on the plan
        for w in self.tree.widths:
            self.helper.register_width(w)
        s = _stream()
        n = self.lib.fmm2d_operator_size(self.Nequiv)
        for d in range(2, self.tree.levels):
            w = self.tree.widths[d]
            if self._plan.level_width.get(d) == w:
                continue
            ops = self.helper.device_operators(w, self.device)
            _lib.check(self.lib.fmm2d_set_operators(self.handle, d, ops.data_ptr(), n, s))
            self._plan.level_width[d] = w

    def tree_array(self, name):
        """A device copy of one tree array (see FMM2D_TREE_* in include/flexmm_amd.h)."""
        t = self.tree
        n, NN = t.N, t.n_nodes
        spec = dict(keys=(n, torch.int64), perm=(n, torch.int32), sx=(n, torch.float64),
                    sy=(n, torch.float64), code=(NN, torch.int64), level=(NN, torch.int32),
                    bot=(NN, torch.int32), top=(NN, torch.int32), parent=(NN, torch.int32),
                    child=(NN, torch.int32), flags=(NN, torch.int32),
                    colleagues=(NN * 9, torch.int32), xmid=(NN, torch.float64),
                    ymid=(NN, torch.float64), row=(NN, torch.int32), m2l=(NN * 40, torch.int32),
                    evalid=(NN, torch.int32))
        cnt, dt = spec[name]
        out = torch.empty(max(1, cnt), dtype=dt, device=self.device)
        _lib.check(self.lib.fmm2d_tree_export(self.handle, _lib.TREE_ARRAYS[name], out.data_ptr(),
                                              out.numel() * out.element_size(), _stream()))
        return out[:cnt]

    def expansion_array(self, name):
        """A device copy of an expansion array (FMM2D_EXP_*), rows of PD doubles."""
        t = self.tree
        cnt = dict(check=t.n_rows * t.PD, multipole=t.n_rows * t.PD, local=t.n_nodes * t.PD,
                   tau_sorted=t.N)[name]
        out = torch.empty(max(1, cnt), dtype=torch.float64, device=self.device)
        _lib.check(self.lib.fmm2d_export_expansions(self.handle, _lib.EXP_ARRAYS[name],
                                                    out.data_ptr(), out.numel(), _stream()))
        return out[:cnt]

    def work_counts(self):
        """Algorithmic work of the tree and the last evaluation (dict)."""
        out = (_lib.ctypes.c_int64 * 9)()
        _lib.check(self.lib.fmm2d_work_counts(self.handle, out, 9))
        return dict(zip(_lib.WORK_COUNTS, [int(v) for v in out]))

    def set_profiling(self, enable=True):
        """Record HIP events around each phase on the current stream."""
        _lib.check(self.lib.fmm2d_set_profiling(self.handle, 1 if enable else 0))

    def phase_times(self):
        """Mean milliseconds of each phase over its runs recorded since
        profiling was enabled or since the previous call (waits for them)."""
        out = (_lib.ctypes.c_double * len(_lib.PHASES))()
        _lib.check(self.lib.fmm2d_phase_times(self.handle, out, len(out)))
        return {k: float(v) for k, v in zip(_lib.PHASES, out) if v >= 0}

    # -- passes --------------------------------------------------------------
    def upwards_pass(self, tau, name, extras):
        """fmm.py:73-124.  With a SourceKernel upwardor, tau is (ntau, n) and
        extras (nextras, n) (fmm.py:80-90); the single-layer upwardor ignores
        extras, as the reference's kernel_app does."""
        if name not in self.helper.upwardors:
            raise KeyError("no upwardor registered under %r" % (name,))
        up = self.helper.upwardors[name]
        n = self.x.numel()
        st = time.time()
        if isinstance(up, _kernels.SourceKernel):
            if extras is None:
                raise ValueError("upwardor %r needs extras (%d x n)" % (name, up.nextras))
            t = _to_device(tau, self.device, "tau")
            e = _to_device(extras, self.device, "extras")
            if t.numel() != up.ntau * n or e.numel() != up.nextras * n:
                raise ValueError("upwardor %r needs tau (%d x %d) and extras (%d x %d)"
                                 % (name, up.ntau, n, up.nextras, n))
            self.tau, self.extras = t, e
            _lib.check(self.lib.fmm2d_upward_dipole(self.handle, t.data_ptr(), e.data_ptr(), n,
                                                    _stream()))
        else:
            t = _to_device(tau, self.device, "tau")
            if t.numel() != n:
                raise ValueError("tau has %d entries for %d sources" % (t.numel(), n))
            self.tau = t
            _lib.check(self.lib.fmm2d_upward(self.handle, t.data_ptr(), t.numel(), _stream()))
        if self.verbose:
            torch.cuda.synchronize()
        self.print("....Time for upwards pass:      {:0.2f}".format(1000 * (time.time() - st)))

    def downwards_pass(self):
        """fmm.py:125-166."""
        st = time.time()
        _lib.check(self.lib.fmm2d_downward(self.handle, _stream()))
        if self.verbose:
            torch.cuda.synchronize()
        self.print("....Time for downwards pass:    {:0.2f}".format(1000 * (time.time() - st)))

    def build_expansions(self, tau, name="slp", extras=None):
        """fmm.py:168-170."""
        self.upwards_pass(tau, name, extras)
        self.downwards_pass()

    def register_evaluator(self, KAS1, KAS2, name, outputs):
        """fmm.py:172-173 / helpers.py:176-238: KAS1/KAS2 are flexmm_amd evaluators."""
        if name in self.helper.evaluators:
            return
        self.helper.evaluators[name] = _kernels.resolve_evaluator(self.kernel, KAS1, KAS2, outputs)

    def register_upwardor(self, KAS, name):
        """fmm.py:174-175 / helpers.py:155-174: the kernel's single-layer
        upwardor (charges) or a flexmm_amd.kernels SourceKernel of this kernel
        (e.g. laplace_slp_dlp, charges + dipoles with extras = normals)."""
        if name in self.helper.upwardors:
            return
        ok = (KAS is self.kernel
              or (isinstance(KAS, _kernels.Evaluator) and KAS.kernel is self.kernel
                  and KAS.evaluator_id == _lib.EVAL_POTENTIAL)
              or (isinstance(KAS, _kernels.SourceKernel) and KAS.kernel is self.kernel))
        if not ok:
            raise NotImplementedError("upwardor %r has no device implementation" % (KAS,))
        self.helper.upwardors[name] = KAS

    def source_evaluation(self, x, y):
        return self.evaluate_to_points(x, y, "potential_evaluation", True)

    def target_evaluation(self, x, y):
        return self.evaluate_to_points(x, y, "potential_evaluation", False)

    def evaluate_to_points(self, x, y, name, check):
        """fmm.py:182-199: returns (outputs, n_targets) in target input order."""
        ev = self.helper.evaluators[name]
        numpy_out = not isinstance(x, torch.Tensor)
        tx = _to_device(x, self.device, "x")
        ty = _to_device(y, self.device, "y")
        if tx.numel() != ty.numel():
            raise ValueError("x and y differ in length")
        out = torch.empty((ev.outputs, tx.numel()), dtype=torch.float64, device=self.device)
        _lib.check(self.lib.fmm2d_evaluate(self.handle, tx.data_ptr(), ty.data_ptr(), tx.numel(),
                                           ev.evaluator_id, 1 if check else 0, out.data_ptr(),
                                           _stream()))
        if numpy_out:
            return out.cpu().numpy()
        return out
