"""FFT computations (reikna.fft)."""

from .batch import BatchFFT, BatchItem, BatchOptions
from .fft import FFT
from .fftshift import FFTShift, fftshift_transformation
from .native import NativeFftPlan, PlanTuning

__all__ = ["FFT", "FFTShift", "BatchFFT", "BatchItem", "BatchOptions", "NativeFftPlan", "fftshift_transformation"]
