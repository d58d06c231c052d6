"""Host-side rules of the drop-in API (no GPU needed)."""
import numpy as np
import pytest

from flexmm_amd import kernels
from flexmm_amd.kifmm2d.fmm import Helper
from oracle import kifmm2d as okifmm


def test_unknown_kernel_is_rejected():
    def Laplace_Eval(sx, sy, tx, ty):
        return -np.log((tx - sx) ** 2 + (ty - sy) ** 2) / (4 * np.pi)
    with pytest.raises(NotImplementedError):
        kernels.resolve_kernel(Laplace_Eval)
    assert kernels.resolve_kernel(kernels.laplace) is kernels.laplace
    assert kernels.resolve_kernel("laplace") is kernels.laplace


def test_kernel_descriptor_is_the_reference_function():
    rng = np.random.default_rng(0)
    a = rng.random((4, 50))
    ref = -np.log((a[2] - a[0]) ** 2 + (a[3] - a[1]) ** 2) / (4 * np.pi)
    np.testing.assert_allclose(kernels.laplace(*a), ref, rtol=1e-15)


def test_register_evaluator_rules():
    ev = kernels.resolve_evaluator(kernels.laplace, kernels.laplace_gradient, kernels.laplace_gradient, 3)
    assert ev.outputs == 3
    with pytest.raises(ValueError):
        kernels.resolve_evaluator(kernels.laplace, kernels.laplace_gradient, None, 1)
    with pytest.raises(NotImplementedError):
        kernels.resolve_evaluator(kernels.laplace, lambda *a: a, None, 3)


@pytest.mark.parametrize("bbox", [None, [0.1, 0.7, 0.2, 0.5]])
def test_helper_bbox_matches_reference_rule(bbox):
    rng = np.random.default_rng(1)
    x, y = rng.random(300) * 0.5 + 0.2, rng.random(300) * 0.3 + 0.1
    h = Helper(kernels.laplace, 32)
    got = h.get_bbox(x.min(), x.max(), y.min(), y.max(), bbox)
    assert got == okifmm.get_bbox(x, y, bbox)
    for w in (1.0, 0.5, 0.25, 0.125):
        h.register_width(w)
    got = h.get_bbox(x.min(), x.max(), y.min(), y.max(), bbox)
    assert got == okifmm.get_bbox(x, y, bbox, widths=[1.0, 0.5, 0.25, 0.125])


def test_modified_helmholtz_descriptors():
    from scipy import special
    ker, pot, grad = kernels.modified_helmholtz(20.0)
    assert kernels.modified_helmholtz(20.0)[0] is ker
    assert ker.params == (20.0,) and ker.kernel_id == 1
    assert kernels.default_evaluator(ker) is pot
    assert kernels.gradient_evaluator(ker) is grad
    assert kernels.gradient_evaluator(kernels.laplace) is kernels.laplace_gradient
    rng = np.random.default_rng(2)
    a = rng.random((4, 40))
    r = np.hypot(a[2] - a[0], a[3] - a[1])
    np.testing.assert_allclose(ker(*a), special.k0(20 * r) / (2 * np.pi), rtol=1e-14)
    u, ux, uy = grad(*a)
    h = 1e-6
    fd = (ker(a[0], a[1], a[2] + h, a[3]) - ker(a[0], a[1], a[2] - h, a[3])) / (2 * h)
    np.testing.assert_allclose(ux, fd, rtol=1e-6, atol=1e-9)
    with pytest.raises(ValueError):
        kernels.modified_helmholtz(0.0)


def test_fmm_evaluate_rejects_host_arrays():
    import torch
    from flexmm_amd.kifmm2d import fmm_evaluate
    z = torch.zeros(4, dtype=torch.float64)
    with pytest.raises(TypeError):
        fmm_evaluate(z, z, z, z, z, kernels.laplace, 10, 16)


def test_square_bbox_holds_both_sets():
    import torch
    from flexmm_amd.kifmm2d.adjoint import _square_bbox
    f64 = dict(dtype=torch.float64)
    x = torch.tensor([0.2, 0.4], **f64); y = torch.tensor([0.3, 0.35], **f64)
    tx = torch.tensor([0.1, 0.9], **f64); ty = torch.tensor([0.5, 0.6], **f64)
    b = _square_bbox(x, y, tx, ty)
    assert b[0] <= 0.1 and b[1] >= 0.9 and b[2] <= 0.3 and b[3] >= 0.6
    assert abs((b[1] - b[0]) - (b[3] - b[2])) < 1e-15


def test_slp_dlp_source_kernel_matches_oracle_and_resolves():
    """kernels.laplace_slp_dlp is the reference's Laplace_Both_Eval
    (examples/pure_kifmm_laplace_dipoles.py:32-40) and register_evaluator's
    (kernel_app, Laplace_Both_Eval) pair maps to the dipole device evaluator."""
    from flexmm_amd import _lib, kernels
    from oracle import kernels as okern
    rng = np.random.default_rng(5)
    sx, sy, tx, ty = rng.random((4, 50))
    tau = rng.random((2, 50))
    ex = rng.standard_normal((2, 50))
    a = kernels.laplace_slp_dlp(sx, sy, tx, ty, tau, ex)[0]
    b = okern.laplace_slp_dlp(sx, sy, tx, ty, tau, ex)[0]
    assert np.allclose(a, b, rtol=1e-15, atol=0)
    ev = kernels.resolve_evaluator(kernels.laplace, kernels.laplace_potential,
                                   kernels.laplace_slp_dlp, 1)
    assert ev.evaluator_id == _lib.EVAL_DIPOLE and ev.outputs == 1
    assert kernels.resolve_evaluator(kernels.laplace, None, kernels.laplace_slp_dlp, 1) is ev
    with pytest.raises(ValueError):
        kernels.resolve_evaluator(kernels.laplace, kernels.laplace_potential,
                                  kernels.laplace_slp_dlp, 3)
    with pytest.raises(NotImplementedError):
        kernels.resolve_evaluator(kernels.laplace, kernels.laplace_gradient,
                                  kernels.laplace_slp_dlp, 1)
    mh, _, _ = kernels.modified_helmholtz(3.0)
    with pytest.raises(ValueError):
        kernels.resolve_evaluator(mh, None, kernels.laplace_slp_dlp, 1)
