"""flexmm_amd: an MI355X-native 2D kernel-independent FMM evaluator.

A drop-in for the ``kifmm2d`` path of flexmm (``FMM(x, y, kernel_eval,
Ncutoff, Nequiv, ...)``), running the quadtree build, the upward and
downward passes and the evaluation in hand-written HIP kernels for gfx950
behind a C-ABI (include/flexmm_amd.h) loaded with ctypes.
"""
from . import kernels
from .kernels import laplace, laplace_gradient

__all__ = ["kernels", "laplace", "laplace_gradient", "kifmm2d", "flexmm2d"]


def __getattr__(name):
    if name == "kifmm2d":
        from . import kifmm2d
        return kifmm2d
    if name == "flexmm2d":
        from . import flexmm2d
        return flexmm2d
    raise AttributeError(name)
