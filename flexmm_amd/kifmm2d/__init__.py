"""MI355X drop-in for flexmm's ``kifmm2d`` (scalar) path."""
from .fmm import FMM, Helper, the_help, TreeInfo

__all__ = ["FMM", "Helper", "the_help", "TreeInfo"]
