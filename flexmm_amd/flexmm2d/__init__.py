"""MI355X drop-in for flexmm's ``flexmm2d`` black-box FMM
(flexmm2d/bbfmm.py: ``BB_FMM``, Chebyshev interpolation operators on the
quadtree shared with the KIFMM)."""
from .bbfmm import BB_FMM, BBHelper, ChebOperators, the_bb_help

__all__ = ["BB_FMM", "BBHelper", "ChebOperators", "the_bb_help"]
