"""Differentiable FMM evaluation: u(t) = sum_s K(|t - s|) tau_s as a
torch.autograd.Function whose backward is itself computed by FMMs.

For a radial kernel K and upstream gradient g = dL/du:

* dL/dtau_s = sum_t g_t K(|s - t|): the transposed FMM, i.e. an FMM with the
  targets as sources carrying charges g, evaluated at the sources;
* dL/ds   = tau_s * grad_s sum_t g_t K(|s - t|): the same transposed FMM with
  the gradient evaluator (outputs u, du/dx, du/dy), so one pass gives both;
* dL/dt   = g_t * grad u(t): the forward FMM's gradient evaluator, computed
  in the forward pass when target positions need gradients.

The gradient evaluator is the reference's ``'gradient'`` evaluator
(examples/pure_kifmm_laplace_gradient.py:32-55,
examples/pure_kifmm_modified_helmholtz_gradient.py:45-68); the reference has
no adjoint of its own.  Self-exclusion (check) carries over: the transposed
FMM skips the same coincident pairs.
"""
import torch
from torch.autograd.function import once_differentiable

from .. import kernels as _kernels
from .fmm import FMM


def _square_bbox(*pts):
    """Smallest square [xmin, xmax, ymin, ymax] holding every point set,
    centred on their joint bounding box (the rule of helpers.py:81-88)."""
    xs = torch.stack([torch.stack([p.min(), p.max()]) for p in pts[0::2]])
    ys = torch.stack([torch.stack([p.min(), p.max()]) for p in pts[1::2]])
    xmin, xmax = float(xs[:, 0].min()), float(xs[:, 1].max())
    ymin, ymax = float(ys[:, 0].min()), float(ys[:, 1].max())
    w = max(xmax - xmin, ymax - ymin)
    xc, yc = 0.5 * (xmin + xmax), 0.5 * (ymin + ymax)
    return [xc - 0.5 * w, xc + 0.5 * w, yc - 0.5 * w, yc + 0.5 * w]


class _Evaluate(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, y, tau, tx, ty, opts):
        kernel, ncut, p, bbox, precision, check = opts
        grad_ev = _kernels.gradient_evaluator(kernel)
        f = FMM(x.detach(), y.detach(), kernel, ncut, p, bbox=bbox, precision=precision)
        f.build_expansions(tau.detach())
        want_t = ctx.needs_input_grad[3] or ctx.needs_input_grad[4]
        if want_t:
            f.register_evaluator(grad_ev, grad_ev, "gradient", 3)
            out = f.evaluate_to_points(tx.detach(), ty.detach(), "gradient", check)
            u, ux, uy = out[0], out[1], out[2]
        else:
            u = f.evaluate_to_points(tx.detach(), ty.detach(), "potential_evaluation", check)[0]
            ux = uy = None
        f.close()
        ctx.opts = opts
        ctx.grad_ev = grad_ev
        ctx.save_for_backward(x, y, tau, tx, ty)
        ctx.tgrad = (ux, uy)
        return u.clone()

    @staticmethod
    @once_differentiable
    def backward(ctx, gu):
        kernel, ncut, p, bbox, precision, check = ctx.opts
        x, y, tau, tx, ty = ctx.saved_tensors
        need_pos = ctx.needs_input_grad[0] or ctx.needs_input_grad[1]
        need_tau = ctx.needs_input_grad[2]
        dx = dy = dtau = dtx = dty = None
        gu = gu.contiguous().to(torch.float64)
        if need_pos or need_tau:
            ft = FMM(tx.detach(), ty.detach(), kernel, ncut, p, bbox=bbox, precision=precision)
            ft.build_expansions(gu)
            if need_pos:
                ft.register_evaluator(ctx.grad_ev, ctx.grad_ev, "gradient", 3)
                v = ft.evaluate_to_points(x.detach(), y.detach(), "gradient", check)
                if need_tau:
                    dtau = v[0].clone()
                dx = tau.detach() * v[1]
                dy = tau.detach() * v[2]
            else:
                dtau = ft.evaluate_to_points(x.detach(), y.detach(), "potential_evaluation",
                                             check)[0].clone()
            ft.close()
        ux, uy = ctx.tgrad
        if ctx.needs_input_grad[3]:
            dtx = gu * ux
        if ctx.needs_input_grad[4]:
            dty = gu * uy
        return dx, dy, dtau, dtx, dty, None


def fmm_evaluate(x, y, tau, tx, ty, kernel, Ncutoff, Nequiv, bbox=None, precision="fp64",
                 check=False):
    """u[t] = sum_s K(|t_t - s_s|) tau_s by the FMM, differentiable in x, y,
    tau, tx and ty (float64 HIP tensors).

    bbox defaults to the smallest square holding both the sources and the
    targets (the forward and the transposed FMM share it).  check=True skips
    coincident source/target pairs (self-evaluation, fmm.py:177-178)."""
    for name, a in (("x", x), ("y", y), ("tau", tau), ("tx", tx), ("ty", ty)):
        if not isinstance(a, torch.Tensor) or a.device.type != "cuda" or a.dtype != torch.float64:
            raise TypeError("%s must be a float64 tensor on a HIP device" % name)
    kernel = _kernels.resolve_kernel(kernel)
    if bbox is None:
        bbox = _square_bbox(x.detach(), y.detach(), tx.detach(), ty.detach())
    opts = (kernel, int(Ncutoff), int(Nequiv), [float(b) for b in bbox], precision, bool(check))
    return _Evaluate.apply(x, y, tau, tx, ty, opts)


class _StokesEvaluate(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, y, fx, fy, tx, ty, opts):
        from .stokes import FMM as StokesFMM
        ncut, p, bbox, precision, check = opts
        f = StokesFMM(x.detach(), y.detach(), p, ncut, bbox=bbox, precision=precision)
        f.build_expansions(fx.detach(), fy.detach())
        u, v, _ = f.evaluate_to_points(tx.detach(), ty.detach(), check)
        f.close()
        ctx.opts = opts
        ctx.save_for_backward(x, y, tx, ty)
        return u.clone(), v.clone()

    @staticmethod
    @once_differentiable
    def backward(ctx, gu, gv):
        from .stokes import FMM as StokesFMM
        ncut, p, bbox, precision, check = ctx.opts
        x, y, tx, ty = ctx.saved_tensors
        if not (ctx.needs_input_grad[2] or ctx.needs_input_grad[3]):
            return None, None, None, None, None, None, None
        # G(t - s) is a symmetric matrix, even in t - s: the adjoint is the
        # same Stokeslet sum with the targets as sources
        ft = StokesFMM(tx.detach(), ty.detach(), p, ncut, bbox=bbox, precision=precision)
        ft.build_expansions(gu.contiguous(), gv.contiguous())
        du, dv, _ = ft.evaluate_to_points(x.detach(), y.detach(), check)
        ft.close()
        return None, None, du.clone(), dv.clone(), None, None, None


def stokes_evaluate(x, y, fx, fy, tx, ty, Nequiv, Ncutoff, bbox=None, precision="fp64",
                    check=False):
    """Velocity (u, v) of the Stokeslet sum by the FMM, differentiable in the
    forces fx, fy (float64 HIP tensors).  bbox defaults to the smallest square
    holding sources and targets; check=True skips coincident pairs."""
    for name, a in (("x", x), ("y", y), ("fx", fx), ("fy", fy), ("tx", tx), ("ty", ty)):
        if not isinstance(a, torch.Tensor) or a.device.type != "cuda" or a.dtype != torch.float64:
            raise TypeError("%s must be a float64 tensor on a HIP device" % name)
    if bbox is None:
        bbox = _square_bbox(x.detach(), y.detach(), tx.detach(), ty.detach())
    opts = (int(Ncutoff), int(Nequiv), [float(b) for b in bbox], precision, bool(check))
    return _StokesEvaluate.apply(x, y, fx, fy, tx, ty, opts)
