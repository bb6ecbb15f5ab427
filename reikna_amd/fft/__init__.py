"""FFT computations (reikna.fft)."""

from .fft import FFT
from .native import NativeFftPlan

__all__ = ["FFT", "NativeFftPlan"]
