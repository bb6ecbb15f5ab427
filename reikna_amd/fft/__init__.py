"""FFT computations (reikna.fft)."""

from .batch import BatchFFT, BatchItem, BatchOptions
from .fft import FFT
from .native import NativeFftPlan

__all__ = ["FFT", "BatchFFT", "BatchItem", "BatchOptions", "NativeFftPlan"]
