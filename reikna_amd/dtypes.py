"""numpy dtype <-> HIP C type helpers (the grunnur.dtypes calls reikna makes:
is_complex, real_for, ctype, c_constant, min_scalar_type, result_type)."""

from __future__ import annotations

from typing import Any

import numpy

_CTYPES = {
    numpy.dtype(numpy.bool_): "bool",
    numpy.dtype(numpy.int8): "signed char",
    numpy.dtype(numpy.uint8): "unsigned char",
    numpy.dtype(numpy.int16): "short",
    numpy.dtype(numpy.uint16): "unsigned short",
    numpy.dtype(numpy.int32): "int",
    numpy.dtype(numpy.uint32): "unsigned int",
    numpy.dtype(numpy.int64): "long long",
    numpy.dtype(numpy.uint64): "unsigned long long",
    numpy.dtype(numpy.float32): "float",
    numpy.dtype(numpy.float64): "double",
    numpy.dtype(numpy.complex64): "float2",
    numpy.dtype(numpy.complex128): "double2",
}


def normalize_type(dtype: Any) -> numpy.dtype:
    return numpy.dtype(dtype)


def is_complex(dtype: Any) -> bool:
    return numpy.dtype(dtype).kind == "c"


def is_double(dtype: Any) -> bool:
    return numpy.dtype(dtype) in (numpy.dtype(numpy.float64), numpy.dtype(numpy.complex128))


def is_integer(dtype: Any) -> bool:
    return numpy.dtype(dtype).kind in "iub"


def is_real(dtype: Any) -> bool:
    return numpy.dtype(dtype).kind == "f"


def real_for(dtype: Any) -> numpy.dtype:
    d = numpy.dtype(dtype)
    if d == numpy.complex64:
        return numpy.dtype(numpy.float32)
    if d == numpy.complex128:
        return numpy.dtype(numpy.float64)
    return d


def complex_for(dtype: Any) -> numpy.dtype:
    d = numpy.dtype(dtype)
    if d == numpy.float32:
        return numpy.dtype(numpy.complex64)
    if d == numpy.float64:
        return numpy.dtype(numpy.complex128)
    if is_complex(d):
        return d
    raise ValueError(f"{d} does not have a complex counterpart")


def ctype(dtype: Any) -> str:
    d = numpy.dtype(dtype)
    if d in _CTYPES:
        return _CTYPES[d]
    raise ValueError(f"dtype {d} has no HIP C type")


def min_scalar_type(val: Any) -> numpy.dtype:
    if isinstance(val, numpy.generic):
        return val.dtype
    if isinstance(val, complex):
        return numpy.dtype(numpy.complex64) if abs(complex(numpy.complex64(val)) - val) == 0 else numpy.dtype(numpy.complex128)
    if isinstance(val, float):
        return numpy.dtype(numpy.float32) if float(numpy.float32(val)) == val else numpy.dtype(numpy.float64)
    if isinstance(val, (int, numpy.integer)):
        return numpy.min_scalar_type(val) if val >= 0 else numpy.dtype(numpy.int32)
    return numpy.asarray(val).dtype


def result_type(*vals: Any) -> numpy.dtype:
    return numpy.result_type(*[min_scalar_type(v) if not isinstance(v, numpy.dtype) else v for v in vals])


def _float_literal(x: float, dtype: numpy.dtype) -> str:
    r = repr(float(x))
    if r in ("inf", "-inf", "nan"):
        raise ValueError(f"cannot write {r} as a constant")
    if "e" not in r and "." not in r:
        r += ".0"
    return r + ("f" if dtype == numpy.float32 else "")


def c_constant(val: Any, dtype: Any = None) -> str:
    """A C literal of ``val`` cast to ``dtype`` (e.g. ``make_float2(1.0f, 0.0f)``)."""
    if dtype is None:
        dtype = min_scalar_type(val)
    d = numpy.dtype(dtype)
    if is_complex(d):
        v = complex(val)
        rd = real_for(d)
        return f"make_{ctype(d)}({_float_literal(v.real, rd)}, {_float_literal(v.imag, rd)})"
    if is_real(d):
        return _float_literal(float(numpy.asarray(val).real), d)
    if d == numpy.bool_:
        return "true" if bool(val) else "false"
    v = int(numpy.asarray(val).real)
    suffix = {"unsigned int": "u", "long long": "ll", "unsigned long long": "ull"}.get(ctype(d), "")
    return f"({ctype(d)}){v}{suffix}"
