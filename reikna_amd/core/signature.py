"""Typed parameters of computations and transformations.

The public names, the role rules and the error texts are reikna's
(reikna/core/signature.py:37-327): user code constructs ``Annotation(arr,
'o')`` and ``Parameter(name, annotation, default)`` directly, and reikna's
tests pin the messages.  The representation is this project's:

* a :class:`Type` is an immutable pair ``(dtype, layout)``, where ``layout``
  is the array's :class:`ArrayMetadata` (shape, strides, offset, buffer
  size) or ``None`` for a scalar;
* an :class:`Annotation` is a Type plus a role, kept as the pair of flags
  ``(input, output)`` looked up from the role letter;
* :class:`Parameter` / :class:`Signature` are ``inspect`` objects, so a
  computation's signature prints and binds like a Python function's; binding
  fills defaults with ``inspect``'s own ``apply_defaults`` and then casts
  scalars to their declared types.
"""

from __future__ import annotations

import inspect
from collections.abc import Mapping, Sequence
from typing import Any

import numpy

from .. import dtypes
from ..helpers import product
from ..runtime import Array, ArrayMetadata, AsArrayMetadata

# role letter -> (the parameter is read, the parameter is written)
_ROLE_FLAGS = {"i": (True, False), "o": (False, True), "io": (True, True), "s": (False, False)}


def is_compatible_with(lhs: ArrayMetadata, rhs: ArrayMetadata) -> bool:
    """Two array layouts address the same elements the same way, up to
    outermost dimensions of length 1 (reikna/core/signature.py:17-34): equal
    types and offsets, equal extents and strides over the dimensions both
    have (counted from the innermost), and every extra outer dimension of
    the longer one of length 1."""
    if lhs.dtype != rhs.dtype or lhs.first_element_offset != rhs.first_element_offset:
        return False
    common = min(len(lhs.shape), len(rhs.shape))
    for k in range(1, common + 1):
        if lhs.shape[-k] != rhs.shape[-k] or lhs.strides[-k] != rhs.strides[-k]:
            return False
    extra = lhs.shape[: len(lhs.shape) - common] + rhs.shape[: len(rhs.shape) - common]
    return product(extra) == 1


def _dtype_and_layout(value: Any) -> tuple[numpy.dtype, ArrayMetadata | None]:
    if isinstance(value, Type):
        return value._dtype, value._layout
    if isinstance(value, AsArrayMetadata):
        meta = value.as_array_metadata()
        return meta.dtype, meta
    return numpy.dtype(value), None


class Type(AsArrayMetadata):
    """An array type (shape, dtype, strides, offset) or a scalar dtype."""

    __slots__ = ("_dtype", "_layout", "ctype")

    def __init__(self, value: Any):
        self._dtype, self._layout = _dtype_and_layout(value)
        self.ctype = dtypes.ctype(self._dtype)

    # -- kind ----------------------------------------------------------------
    def is_array(self) -> bool:
        return self._layout is not None

    def is_scalar(self) -> bool:
        return self._layout is None

    # -- layout --------------------------------------------------------------
    @property
    def array_metadata(self) -> ArrayMetadata:
        if self._layout is None:
            raise ValueError("This is a scalar type")
        return self._layout

    def as_array_metadata(self) -> ArrayMetadata:
        return self.array_metadata

    @property
    def dtype(self) -> numpy.dtype:
        return self._dtype

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

    # -- value semantics -----------------------------------------------------
    def __eq__(self, other: object) -> bool:
        return isinstance(other, Type) and (self._dtype, self._layout) == (other._dtype, other._layout)

    def __hash__(self) -> int:
        return hash((self._dtype, self._layout))

    def compatible_with(self, other: Type) -> bool:
        """Scalars of the same type, or arrays whose layouts match up to
        outer dimensions of length 1."""
        if self.is_array() and other.is_array():
            return is_compatible_with(self._layout, other._layout)
        return self.is_scalar() and other.is_scalar() and self._dtype == other._dtype

    def with_dtype(self, dtype: Any) -> Type:
        dtype = numpy.dtype(dtype)
        return Type(self._layout.with_(dtype=dtype) if self.is_array() else dtype)

    @classmethod
    def from_value(cls, val: Any) -> Type:
        """The type of a value: a Type, an array or its metadata, a dtype or
        numpy scalar class, a numpy scalar, an array-like, or a Python number
        (which takes the smallest numpy type that holds it)."""
        if isinstance(val, (Type, AsArrayMetadata, numpy.dtype)):
            return cls(val)
        if isinstance(val, type) and issubclass(val, numpy.generic):
            return cls(numpy.dtype(val))
        if isinstance(val, numpy.generic):
            return cls(val.dtype)
        if hasattr(val, "dtype") and hasattr(val, "shape"):
            return cls(ArrayMetadata.from_arraylike(val))
        return cls(dtypes.result_type(val))

    def cast_scalar(self, val: Any) -> numpy.generic:
        if self.is_array():
            raise ValueError("Can only cast scalars to a scalar type, this is an array")
        return self._dtype.type(val)

    def __repr__(self) -> str:
        return f"Type({self._layout if self.is_array() else self._dtype})"


class Annotation(Type):
    """
    A parameter's Type and role: ``'i'`` (read), ``'o'`` (written), ``'io'``
    (both) for arrays, ``'s'`` for scalars.  Without a role, scalars get
    ``'s'``, arrays ``'io'``, and constant arrays ``'i'``.
    """

    __slots__ = ("role", "input", "output", "constant")

    def __init__(self, type_: Any, role: str | None = None, *, constant: bool = False):
        Type.__init__(self, type_)
        if role is None:
            role = "s" if self.is_scalar() else ("i" if constant else "io")
        flags = _ROLE_FLAGS.get(role)
        if flags is None:
            raise ValueError(f"Invalid role: {role}")
        if role == "s" and self.is_array():
            raise ValueError("Only scalars can have the scalar role")
        if role != "s" and self.is_scalar():
            raise ValueError("Scalars cannot have input or output role")
        self.role = role
        self.input, self.output = flags
        self.constant = constant

    @property
    def type(self) -> Type:
        """The bare type, without role."""
        return Type(self)

    def can_be_argument_for(self, annotation: Annotation) -> bool:
        """An argument of this annotation may be passed for a parameter of
        ``annotation``: compatible types, and the same role, or a read-write
        array passed where any array is expected."""
        if not self.compatible_with(annotation):
            return False
        return self.role == annotation.role or (self.role == "io" and annotation.is_array())

    def __eq__(self, other: object) -> bool:
        return (
            isinstance(other, Annotation)
            and Type.__eq__(self, other)
            and (self.role, self.constant) == (other.role, other.constant)
        )

    def __hash__(self) -> int:
        return hash((Type.__hash__(self), self.role, self.constant))

    def __repr__(self) -> str:
        inner = self._layout if self.is_array() else self._dtype
        return f"Annotation({inner}, role={self.role}{', constant' if self.constant else ''})"


class Parameter(inspect.Parameter):
    """A named, annotated parameter (positional or keyword); only scalars
    may have a default, which is cast to the scalar's type."""

    def __init__(self, name: str, annotation: Annotation, default: Any = inspect.Parameter.empty):
        if default is not inspect.Parameter.empty:
            if annotation.is_array():
                raise ValueError("Array parameters cannot have default values")
            default = annotation.cast_scalar(default)
        super().__init__(name, inspect.Parameter.POSITIONAL_OR_KEYWORD, annotation=annotation, default=default)

    def rename(self, new_name: str) -> Parameter:
        return Parameter(new_name, self.annotation, default=self.default)


class Signature(inspect.Signature):
    """The ordered parameters of a computation or transformation."""

    def __init__(self, parameters: Sequence[Parameter]):
        super().__init__(list(parameters))

    def bind_with_defaults(
        self, args: Sequence[Any], kwds: Mapping[str, Any], *, cast: bool = False
    ) -> inspect.BoundArguments:
        """Binds like a call, fills in defaults, and (with ``cast``) converts
        scalar arguments to their parameters' types."""
        bound = self.bind(*args, **kwds)
        given = set(bound.arguments)
        bound.apply_defaults()
        if cast:
            for name in given:
                ann = self.parameters[name].annotation
                if ann.is_scalar():
                    bound.arguments[name] = ann.cast_scalar(bound.arguments[name])
        return bound
