"""General-purpose computations used by (and around) the FFT path."""

from .pureparallel import PureParallel

__all__ = ["PureParallel"]
