"""fmm2d_evaluate_noalloc (evaluation whose scratch is the library's static
device array) against fmm2d_evaluate on the same plan, through the C-ABI:
two calls in a row, the second with fewer targets, for the FP64 potential,
the gradient evaluator, self-evaluation and the mixed-precision plan; sizes
past the array's limit are refused before any GPU work."""
import ctypes

import numpy as np
import pytest
import torch

from tests import helpers as H

pytestmark = pytest.mark.gpu

# Same kernels, so the bar is parity itself: FP64 to 1e-13.  Targets reach
# their leaf in an order set by atomics, so a target can land in a different
# work item of its leaf from call to call; in the mixed form an item of at
# most 32 targets takes the FP64 near field, a full one the FP32 one, so two
# calls of the same entry point differ at the mixed form's own level (1e-7).
TOL = {"fp64": 1e-13, "mixed": 1e-7}


def _plan(n, seed, precision="fp64", ncut=64, p=60):
    from flexmm_amd import kernels
    from flexmm_amd.kifmm2d import FMM, the_help
    the_help.clear()   # the root box is the data's (a warm cache would snap it)
    x, y, tau = H.uniform(n, seed)
    dev = torch.device("cuda", 0)
    T = lambda a: torch.from_numpy(a).to(dev)
    f = FMM(T(x), T(y), kernels.laplace, ncut, p, precision=precision)
    f.build_expansions(T(tau))
    return f, T(x), T(y)


def _both(f, tx, ty, evaluator, outputs, check):
    from flexmm_amd import _lib
    lib = _lib.load()
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    n = tx.numel()
    ref = torch.empty((outputs, n), dtype=torch.float64, device=tx.device)
    out = torch.full((outputs, n), float("nan"), dtype=torch.float64, device=tx.device)
    _lib.check(lib.fmm2d_evaluate(f.handle, tx.data_ptr(), ty.data_ptr(), n, evaluator, check,
                                  ref.data_ptr(), s))
    _lib.check(lib.fmm2d_evaluate_noalloc(f.handle, tx.data_ptr(), ty.data_ptr(), n, evaluator,
                                          check, out.data_ptr(), s))
    torch.cuda.synchronize()
    return out, ref


def _err(out, ref):
    assert bool(torch.isfinite(out).all())
    return float((out - ref).abs().max() / ref.abs().max())


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_noalloc_matches_evaluate_two_calls(precision):
    from flexmm_amd import _lib
    print("noalloc potential", precision, flush=True)
    f, _, _ = _plan(200_000, 41, precision)
    rng = np.random.default_rng(42)
    dev = torch.device("cuda", 0)
    for nt in (300_000, 7_000):
        tx = torch.from_numpy(rng.random(nt)).to(dev)
        ty = torch.from_numpy(rng.random(nt)).to(dev)
        out, ref = _both(f, tx, ty, _lib.EVAL_POTENTIAL, 1, 0)
        err = _err(out, ref)
        print("nt", nt, "noalloc vs evaluate %.3e" % err)
        assert err < TOL[precision]


def test_noalloc_gradient_and_self_evaluation():
    from flexmm_amd import _lib, kernels
    print("noalloc gradient / self", flush=True)
    f, x, y = _plan(120_000, 43)
    f.register_evaluator(kernels.laplace_gradient, kernels.laplace_gradient, "gradient", 3)
    out, ref = _both(f, x, y, _lib.EVAL_POTENTIAL, 1, 1)
    assert _err(out, ref) < TOL["fp64"]
    out, ref = _both(f, x[:5000], y[:5000], _lib.EVAL_GRADIENT, 3, 1)
    for k in range(3):
        assert _err(out[k], ref[k]) < TOL["fp64"]
    # the oracle agrees with the no-alloc self-evaluation
    from oracle import kernels as okern
    from oracle import kifmm2d as okifmm
    xs, ys, tau = H.uniform(120_000, 43)
    o = okifmm.FMM(xs, ys, okern.laplace, 64, 60)
    assert o.bbox == f.bbox
    o.build_expansions(tau)
    out, _ = _both(f, x, y, _lib.EVAL_POTENTIAL, 1, 1)
    assert H.rel_max_err(out[0].cpu().numpy(), o.source_evaluation(xs, ys)[0]) < 1e-12


def test_noalloc_limits_are_refused_before_gpu_work():
    from flexmm_amd import _lib
    f, x, y = _plan(20_000, 44)
    lib = _lib.load()
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    too_many = (1 << 20) + 1
    rc = lib.fmm2d_evaluate_noalloc(f.handle, x.data_ptr(), y.data_ptr(), too_many,
                                    _lib.EVAL_POTENTIAL, 0, x.data_ptr(), s)
    assert rc == _lib.E_LIMIT
    assert b"at most 1048576 targets" in lib.fmm2d_last_error()
    rc = lib.fmm2d_evaluate_noalloc(f.handle, None, None, 0, _lib.EVAL_POTENTIAL, 0, None, s)
    assert rc == 0
