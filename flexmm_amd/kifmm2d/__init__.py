"""MI355X drop-in for flexmm's ``kifmm2d`` path: the scalar ``FMM``
(kifmm2d/scalar/fmm.py), the Stokeslet ``StokesFMM``
(kifmm2d/vector/stokes_fmm.py) and the differentiable ``fmm_evaluate``."""
from .adjoint import fmm_evaluate, stokes_evaluate
from .fmm import FMM, Helper, the_help, TreeInfo
from .stokes import FMM as StokesFMM

__all__ = ["FMM", "Helper", "the_help", "TreeInfo", "fmm_evaluate", "stokes_evaluate", "StokesFMM"]
