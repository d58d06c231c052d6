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
