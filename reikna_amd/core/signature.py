"""Typed computation parameters: Type, Annotation, Parameter, Signature.

Same contract as reikna/core/signature.py:37-327: a ``Type`` is an array
(``ArrayMetadata``) or a scalar dtype; an ``Annotation`` adds a role
(``'i'``, ``'o'``, ``'io'``, ``'s'``) and a constant flag; ``Parameter`` and
``Signature`` are ``inspect`` objects, so ``bind_with_defaults`` casts scalars
and fills defaults exactly as the reference's callables do (:312-327).
"""

from __future__ import annotations

import inspect
from collections.abc import Mapping, Sequence
from typing import Any

import numpy

from .. import dtypes
from ..helpers import product
from ..runtime import Array, ArrayMetadata, AsArrayMetadata


def is_compatible_with(lhs: ArrayMetadata, rhs: ArrayMetadata) -> bool:
    """
    True if the two arrays are the same up to outermost dimensions of length 1
    (reikna/core/signature.py:17-34).
    """
    if lhs.dtype != rhs.dtype:
        return False
    n = min(len(lhs.shape), len(rhs.shape))
    if n > 0:
        if lhs.shape[-n:] != rhs.shape[-n:] or lhs.strides[-n:] != rhs.strides[-n:]:
            return False
    if product(lhs.shape[: len(lhs.shape) - n]) != 1 or product(rhs.shape[: len(rhs.shape) - n]) != 1:
        return False
    return lhs.first_element_offset == rhs.first_element_offset


class Type(AsArrayMetadata):
    """An array type (shape, dtype, strides, offset) or a scalar dtype."""

    def __init__(self, metadata: Any):
        if isinstance(metadata, Type):
            self._metadata = metadata._metadata
        elif isinstance(metadata, AsArrayMetadata):
            self._metadata = metadata.as_array_metadata()
        else:
            self._metadata = numpy.dtype(metadata)
        self.ctype = dtypes.ctype(self.dtype)

    def _raw_metadata(self) -> ArrayMetadata | numpy.dtype:
        return self._metadata

    def as_array_metadata(self) -> ArrayMetadata:
        return self.array_metadata

    @property
    def array_metadata(self) -> ArrayMetadata:
        if isinstance(self._metadata, ArrayMetadata):
            return self._metadata
        raise ValueError("This is a scalar type")

    @property
    def dtype(self) -> numpy.dtype:
        return self._metadata.dtype if isinstance(self._metadata, ArrayMetadata) else self._metadata

    @property
    def shape(self) -> tuple[int, ...]:
        return self.array_metadata.shape

    @property
    def strides(self) -> tuple[int, ...]:
        return self.array_metadata.strides

    @property
    def offset(self) -> int:
        return self.array_metadata.first_element_offset

    @property
    def nbytes(self) -> int:
        return self.array_metadata.buffer_size

    def __eq__(self, other: object) -> bool:
        return isinstance(other, Type) and self._metadata == other._metadata

    def __hash__(self) -> int:
        return hash((Type, self._metadata))

    def is_scalar(self) -> bool:
        return isinstance(self._metadata, numpy.dtype)

    def is_array(self) -> bool:
        return not self.is_scalar()

    def compatible_with(self, other: Type) -> bool:
        if self.is_scalar() and other.is_scalar():
            return self._metadata == other._metadata
        if self.is_scalar() != other.is_scalar():
            return False
        return is_compatible_with(self._metadata, other._metadata)

    def with_dtype(self, dtype: Any) -> Type:
        if isinstance(self._metadata, ArrayMetadata):
            return Type(self._metadata.with_(dtype=numpy.dtype(dtype)))
        return Type(numpy.dtype(dtype))

    @classmethod
    def from_value(cls, val: Any) -> Type:
        if isinstance(val, Type):
            return cls(val)
        if isinstance(val, Array):
            return cls(val.metadata)
        if isinstance(val, (ArrayMetadata, numpy.dtype)):
            return cls(val)
        if isinstance(val, type) and issubclass(val, numpy.generic):
            return cls(numpy.dtype(val))
        if hasattr(val, "dtype") and hasattr(val, "shape"):
            if len(val.shape) == 0 and not isinstance(val, numpy.ndarray):
                return cls(numpy.dtype(val.dtype))
            return cls(ArrayMetadata.from_arraylike(val))
        return cls(dtypes.result_type(val))

    def cast_scalar(self, val: Any) -> numpy.generic:
        if not self.is_scalar():
            raise ValueError("Can only cast scalars to a scalar type, this is an array")
        return numpy.asarray(val, dtype=self.dtype).flat[0]

    def __repr__(self) -> str:
        return f"Type({self._metadata})"


class Annotation(Type):
    """
    A computation parameter annotation: a Type plus a role, one of
    ``'i'``, ``'o'``, ``'io'`` (arrays) or ``'s'`` (scalars); the default role is
    ``'s'`` for scalars, ``'io'`` for arrays and ``'i'`` for constant arrays.
    """

    def __init__(self, type_: Any, role: str | None = None, *, constant: bool = False):
        super().__init__(type_)
        if role is None:
            role = "s" if self.is_scalar() else ("i" if constant else "io")
        if role not in ("i", "o", "io", "s"):
            raise ValueError(f"Invalid role: {role}")
        self.role = role
        self.constant = constant
        if role == "s":
            if not self.is_scalar():
                raise ValueError("Only scalars can have the scalar role")
            self.input = False
            self.output = False
        else:
            if self.is_scalar():
                raise ValueError("Scalars cannot have input or output role")
            self.input = "i" in role
            self.output = "o" in role

    def can_be_argument_for(self, annotation: Annotation) -> bool:
        if not self.compatible_with(annotation):
            return False
        if self.role == annotation.role:
            return True
        return self.role == "io" and annotation.is_array()

    def _as_type(self) -> Type:
        return Type(self)

    def __eq__(self, other: object) -> bool:
        return (
            isinstance(other, Annotation)
            and self._metadata == other._metadata
            and self.role == other.role
            and self.constant == other.constant
        )

    def __hash__(self) -> int:
        return hash((Annotation, self._metadata, self.role, self.constant))

    def __repr__(self) -> str:
        const = ", constant" if self.constant else ""
        return f"Annotation({self._metadata}, role={self.role}{const})"


class Parameter(inspect.Parameter):
    """A computation parameter (always POSITIONAL_OR_KEYWORD); only scalars have defaults."""

    def __init__(self, name: str, annotation: Annotation, default: Any = inspect.Parameter.empty):
        if default is not inspect.Parameter.empty:
            if annotation.is_array():
                raise ValueError("Array parameters cannot have default values")
            default = annotation.cast_scalar(default)
        super().__init__(name, inspect.Parameter.POSITIONAL_OR_KEYWORD, annotation=annotation, default=default)

    def rename(self, new_name: str) -> Parameter:
        return Parameter(new_name, self.annotation, default=self.default)


class Signature(inspect.Signature):
    """An ordered list of Parameters."""

    def __init__(self, parameters: Sequence[Parameter]):
        super().__init__(list(parameters))

    @property
    def _reikna_parameters(self) -> Mapping[str, Parameter]:
        return self.parameters

    def bind_with_defaults(
        self, args: Sequence[Any], kwds: Mapping[str, Any], *, cast: bool = False
    ) -> inspect.BoundArguments:
        bound = self.bind(*args, **kwds)
        for param in self.parameters.values():
            if param.name not in bound.arguments:
                bound.arguments[param.name] = param.default
            elif cast and param.annotation.is_scalar():
                bound.arguments[param.name] = param.annotation.cast_scalar(bound.arguments[param.name])
        return bound
