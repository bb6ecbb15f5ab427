"""reikna_amd: reikna's FFT Computation/Transformation API on hand-written
gfx950 (MI355X) HIP kernels, reached through the C-ABI in include/reikna_amd.h.

Importing the package does not touch the GPU and does not need torch."""

from . import dtypes, functions, helpers, transformations
from .core import (
    Annotation,
    Computation,
    ComputationParameter,
    ComputationPlan,
    Indices,
    KernelArgument,
    KernelParameter,
    Parameter,
    Signature,
    Transformation,
    Type,
)
from .runtime import Array, ArrayMetadata, Device, DeviceParameters, Event, Queue

__version__ = "0.1.0"

__all__ = [
    "Annotation",
    "Array",
    "ArrayMetadata",
    "Computation",
    "ComputationParameter",
    "ComputationPlan",
    "Device",
    "DeviceParameters",
    "Event",
    "Indices",
    "KernelArgument",
    "KernelParameter",
    "Parameter",
    "Queue",
    "Signature",
    "Transformation",
    "Type",
    "dtypes",
    "functions",
    "helpers",
    "transformations",
]
