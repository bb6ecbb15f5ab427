"""General-purpose computations used by (and around) the FFT path."""

from .pureparallel import PureParallel
from .transpose import Transpose

__all__ = ["PureParallel", "Transpose"]
