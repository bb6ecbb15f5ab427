"""The Computation / Transformation API (reikna.core)."""

from .computation import (
    Computation,
    ComputationCallable,
    ComputationParameter,
    ComputationPlan,
    KernelArgument,
    KernelArguments,
)
from .signature import Annotation, Parameter, Signature, Type
from .transformation import Indices, KernelParameter, Transformation, TransformationParameter

__all__ = [
    "Annotation",
    "Computation",
    "ComputationCallable",
    "ComputationParameter",
    "ComputationPlan",
    "Indices",
    "KernelArgument",
    "KernelArguments",
    "KernelParameter",
    "Parameter",
    "Signature",
    "Transformation",
    "TransformationParameter",
    "Type",
]
