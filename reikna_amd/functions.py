"""Typed arithmetic helpers for kernel/transformation code (the
grunnur.functions calls reikna uses: mul, div, add, cast, polar_unit, norm,
conj, exp, pow; e.g. reikna/transformations.py:91, 117, 152 and
reikna/fft/fft.py:247-249).  Each returns a Module: ``${mul}(a, b)`` calls a
generated ``__device__`` function with complex-aware arithmetic."""

from __future__ import annotations

from typing import Any

import numpy

from . import dtypes
from .template import Module


def _ct(d: Any) -> str:
    return dtypes.ctype(d)


def _to(out: numpy.dtype, expr: str, src: numpy.dtype) -> str:
    """C expression converting `expr` of dtype src to dtype out."""
    if out == src:
        return expr
    if dtypes.is_complex(out):
        r = _ct(dtypes.real_for(out))
        if dtypes.is_complex(src):
            return f"make_{_ct(out)}(({r})({expr}).x, ({r})({expr}).y)"
        return f"make_{_ct(out)}(({r})({expr}), ({r})0)"
    if dtypes.is_complex(src):
        return f"({_ct(out)})(({expr}).x)"
    return f"({_ct(out)})({expr})"


def _module(body_fn: Any) -> Module:
    return Module(body_fn)


def cast(out_dtype: Any, in_dtype: Any) -> Module:
    out, src = numpy.dtype(out_dtype), numpy.dtype(in_dtype)

    def render(prefix: str) -> str:
        return (
            f"__device__ __forceinline__ {_ct(out)} {prefix}({_ct(src)} x) "
            f"{{ return {_to(out, 'x', src)}; }}\n"
        )

    return _module(render)


def _binary(op: str, in_dtypes: tuple[numpy.dtype, ...], out_dtype: Any) -> Module:
    ins = [numpy.dtype(d) for d in in_dtypes]
    out = numpy.dtype(out_dtype) if out_dtype is not None else numpy.result_type(*ins)

    def render(prefix: str) -> str:
        params = ", ".join(f"{_ct(d)} a{k}" for k, d in enumerate(ins))
        acc_t = out
        acc = _to(acc_t, "a0", ins[0])
        lines = [f"    {_ct(acc_t)} r = {acc};"]
        for k, d in enumerate(ins[1:], start=1):
            b = _to(acc_t, f"a{k}", d) if dtypes.is_complex(acc_t) or not dtypes.is_complex(d) else f"a{k}"
            if dtypes.is_complex(acc_t):
                if op == "+":
                    lines.append(f"    r = make_{_ct(acc_t)}(r.x + ({b}).x, r.y + ({b}).y);")
                elif op == "*":
                    if dtypes.is_complex(d):
                        lines.append(f"    {{ {_ct(acc_t)} b = {b}; r = make_{_ct(acc_t)}(r.x * b.x - r.y * b.y, r.x * b.y + r.y * b.x); }}")
                    else:
                        lines.append(f"    r = make_{_ct(acc_t)}(r.x * a{k}, r.y * a{k});")
                elif op == "/":
                    if dtypes.is_complex(d):
                        lines.append(
                            f"    {{ {_ct(acc_t)} b = {b}; {_ct(dtypes.real_for(acc_t))} n = b.x * b.x + b.y * b.y; "
                            f"r = make_{_ct(acc_t)}((r.x * b.x + r.y * b.y) / n, (r.y * b.x - r.x * b.y) / n); }}"
                        )
                    else:
                        lines.append(f"    r = make_{_ct(acc_t)}(r.x / a{k}, r.y / a{k});")
            else:
                lines.append(f"    r = r {op} ({_ct(acc_t)})({b});")
        body = "\n".join(lines)
        return f"__device__ __forceinline__ {_ct(out)} {prefix}({params})\n{{\n{body}\n    return r;\n}}\n"

    return _module(render)


def mul(*in_dtypes: Any, out_dtype: Any = None) -> Module:
    return _binary("*", tuple(numpy.dtype(d) for d in in_dtypes), out_dtype)


def add(*in_dtypes: Any, out_dtype: Any = None) -> Module:
    return _binary("+", tuple(numpy.dtype(d) for d in in_dtypes), out_dtype)


def div(in_dtype1: Any, in_dtype2: Any, out_dtype: Any = None) -> Module:
    return _binary("/", (numpy.dtype(in_dtype1), numpy.dtype(in_dtype2)), out_dtype)


def conj(dtype: Any) -> Module:
    d = numpy.dtype(dtype)

    def render(prefix: str) -> str:
        return f"__device__ __forceinline__ {_ct(d)} {prefix}({_ct(d)} x) {{ return make_{_ct(d)}(x.x, -x.y); }}\n"

    return _module(render)


def polar_unit(dtype: Any) -> Module:
    """exp(i * theta) for a real theta of `dtype`."""
    d = numpy.dtype(dtype)
    c = dtypes.complex_for(d)

    def render(prefix: str) -> str:
        fn = "sincosf" if d == numpy.float32 else "sincos"
        return (
            f"__device__ __forceinline__ {_ct(c)} {prefix}({_ct(d)} theta) "
            f"{{ {_ct(d)} s, co; {fn}(theta, &s, &co); return make_{_ct(c)}(co, s); }}\n"
        )

    return _module(render)


def polar(dtype: Any) -> Module:
    d = numpy.dtype(dtype)
    c = dtypes.complex_for(d)

    def render(prefix: str) -> str:
        fn = "sincosf" if d == numpy.float32 else "sincos"
        return (
            f"__device__ __forceinline__ {_ct(c)} {prefix}({_ct(d)} rho, {_ct(d)} theta) "
            f"{{ {_ct(d)} s, co; {fn}(theta, &s, &co); return make_{_ct(c)}(rho * co, rho * s); }}\n"
        )

    return _module(render)


def norm(dtype: Any) -> Module:
    """|x|^2 (reikna's norm: squared absolute value)."""
    d = numpy.dtype(dtype)
    r = dtypes.real_for(d)

    def render(prefix: str) -> str:
        if dtypes.is_complex(d):
            return f"__device__ __forceinline__ {_ct(r)} {prefix}({_ct(d)} x) {{ return x.x * x.x + x.y * x.y; }}\n"
        return f"__device__ __forceinline__ {_ct(d)} {prefix}({_ct(d)} x) {{ return x * x; }}\n"

    return _module(render)


def exp(dtype: Any) -> Module:
    d = numpy.dtype(dtype)

    def render(prefix: str) -> str:
        if dtypes.is_complex(d):
            r = _ct(dtypes.real_for(d))
            e, sc = ("expf", "sincosf") if d == numpy.complex64 else ("exp", "sincos")
            return (
                f"__device__ __forceinline__ {_ct(d)} {prefix}({_ct(d)} x) "
                f"{{ {r} m = {e}(x.x), s, c; {sc}(x.y, &s, &c); return make_{_ct(d)}(m * c, m * s); }}\n"
            )
        e = "expf" if d == numpy.float32 else "exp"
        return f"__device__ __forceinline__ {_ct(d)} {prefix}({_ct(d)} x) {{ return {e}(x); }}\n"

    return _module(render)


def pow(dtype: Any, exponent_dtype: Any = None, out_dtype: Any = None) -> Module:  # noqa: A001
    d = numpy.dtype(dtype)
    e = numpy.dtype(exponent_dtype) if exponent_dtype is not None else d
    out = numpy.dtype(out_dtype) if out_dtype is not None else d
    if dtypes.is_complex(d):
        raise NotImplementedError("complex pow")

    def render(prefix: str) -> str:
        fn = "powf" if out == numpy.float32 else "pow"
        return (
            f"__device__ __forceinline__ {_ct(out)} {prefix}({_ct(d)} x, {_ct(e)} p) "
            f"{{ return {fn}(({_ct(out)})x, ({_ct(out)})p); }}\n"
        )

    return _module(render)
