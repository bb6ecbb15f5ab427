"""Ready-made transformations (reikna/transformations.py:11-347): copy,
copy_broadcasted, cast, add/mul/div by a constant or a parameter,
split/combine complex, norms, ignore, broadcast a constant or a parameter."""

from __future__ import annotations

from typing import Any

import numpy

from . import dtypes, functions
from .core import Annotation, Parameter, Transformation
from .runtime import ArrayMetadata


def _meta(x: Any) -> ArrayMetadata:
    return x.as_array_metadata() if hasattr(x, "as_array_metadata") else ArrayMetadata.from_arraylike(x)


def copy(arr_t: Any, out_arr_t: Any = None) -> Transformation:
    """``output = input``; ``out_arr_t`` may have other strides, same shape and dtype."""
    inp = _meta(arr_t)
    out = _meta(out_arr_t) if out_arr_t is not None else inp
    if out.shape != inp.shape or out.dtype != inp.dtype:
        raise ValueError("Input and output arrays must have the same shape and data type")
    return Transformation(
        [Parameter("output", Annotation(out, "o")), Parameter("input", Annotation(inp, "i"))],
        "${output.store_same}(${input.load_same});",
    )


def _broadcastable_to(src: ArrayMetadata, dst: ArrayMetadata) -> bool:
    if len(src.shape) > len(dst.shape):
        return False
    return all(src.shape[-i] in (1, dst.shape[-i]) for i in range(1, len(src.shape) + 1))


def copy_broadcasted(arr_t: Any, out_arr_t: Any = None) -> Transformation:
    """``output = input`` with numpy broadcasting of ``input`` (an input-only transformation)."""
    inp = _meta(arr_t)
    out = _meta(out_arr_t) if out_arr_t is not None else inp
    if out.dtype != inp.dtype:
        raise ValueError("Input and output arrays must have the same data type")
    if not _broadcastable_to(inp, out):
        raise ValueError("Input is not broadcastable to output")
    return Transformation(
        [Parameter("output", Annotation(out, "o")), Parameter("input", Annotation(inp, "i"))],
        """
        <%
            lead = len(output.shape) - len(input.shape)
            ids = ["0" if input.shape[i] == 1 else idxs[i + lead] for i in range(len(input.shape))]
        %>
        ${output.store_same}(${input.load_idx}(${", ".join(ids)}));
        """,
        connectors=["output"],
    )


def cast(arr_t: Any, dtype: Any) -> Transformation:
    """``output = dtype(input)``."""
    inp = _meta(arr_t)
    out = inp.with_(dtype=numpy.dtype(dtype))
    return Transformation(
        [Parameter("output", Annotation(out, "o")), Parameter("input", Annotation(inp, "i"))],
        "${output.store_same}(${cast}(${input.load_same}));",
        render_kwds=dict(cast=functions.cast(dtype, inp.dtype)),
    )


def add_param(arr_t: Any, param_dtype: Any) -> Transformation:
    """``output = input + param``."""
    inp = _meta(arr_t)
    return Transformation(
        [
            Parameter("output", Annotation(inp, "o")),
            Parameter("input", Annotation(inp, "i")),
            Parameter("param", Annotation(param_dtype)),
        ],
        "${output.store_same}(${add}(${input.load_same}, ${param}));",
        render_kwds=dict(add=functions.add(inp.dtype, param_dtype, out_dtype=inp.dtype)),
    )


def add_const(arr_t: Any, param: Any) -> Transformation:
    """``output = input + param`` with a constant ``param``."""
    inp = _meta(arr_t)
    pdt = dtypes.min_scalar_type(param)
    return Transformation(
        [Parameter("output", Annotation(inp, "o")), Parameter("input", Annotation(inp, "i"))],
        "${output.store_same}(${add}(${input.load_same}, ${param}));",
        render_kwds=dict(add=functions.add(inp.dtype, pdt, out_dtype=inp.dtype), param=dtypes.c_constant(param, pdt)),
    )


def mul_param(arr_t: Any, param_dtype: Any) -> Transformation:
    """``output = input * param``."""
    inp = _meta(arr_t)
    return Transformation(
        [
            Parameter("output", Annotation(inp, "o")),
            Parameter("input", Annotation(inp, "i")),
            Parameter("param", Annotation(param_dtype)),
        ],
        "${output.store_same}(${mul}(${input.load_same}, ${param}));",
        render_kwds=dict(mul=functions.mul(inp.dtype, param_dtype, out_dtype=inp.dtype)),
    )


def mul_const(arr_t: Any, param: Any) -> Transformation:
    """``output = input * param`` with a constant ``param``."""
    inp = _meta(arr_t)
    pdt = dtypes.min_scalar_type(param)
    return Transformation(
        [Parameter("output", Annotation(inp, "o")), Parameter("input", Annotation(inp, "i"))],
        "${output.store_same}(${mul}(${input.load_same}, ${param}));",
        render_kwds=dict(mul=functions.mul(inp.dtype, pdt, out_dtype=inp.dtype), param=dtypes.c_constant(param, pdt)),
    )


def div_param(arr_t: Any, param_dtype: Any) -> Transformation:
    """``output = input / param``."""
    inp = _meta(arr_t)
    return Transformation(
        [
            Parameter("output", Annotation(inp, "o")),
            Parameter("input", Annotation(inp, "i")),
            Parameter("param", Annotation(param_dtype)),
        ],
        "${output.store_same}(${div}(${input.load_same}, ${param}));",
        render_kwds=dict(div=functions.div(inp.dtype, param_dtype, out_dtype=inp.dtype)),
    )


def div_const(arr_t: Any, param: Any) -> Transformation:
    """``output = input / param`` with a constant ``param``."""
    inp = _meta(arr_t)
    pdt = dtypes.min_scalar_type(param)
    return Transformation(
        [Parameter("output", Annotation(inp, "o")), Parameter("input", Annotation(inp, "i"))],
        "${output.store_same}(${div}(${input.load_same}, ${param}));",
        render_kwds=dict(div=functions.div(inp.dtype, pdt, out_dtype=inp.dtype), param=dtypes.c_constant(param, pdt)),
    )


def split_complex(input_arr_t: Any) -> Transformation:
    """``real = Re(input), imag = Im(input)``."""
    inp = _meta(input_arr_t)
    out = ArrayMetadata(inp.shape, dtypes.real_for(inp.dtype))
    return Transformation(
        [
            Parameter("real", Annotation(out, "o")),
            Parameter("imag", Annotation(out, "o")),
            Parameter("input", Annotation(inp, "i")),
        ],
        """
        ${input.ctype} _v = ${input.load_same};
        ${real.store_same}(_v.x);
        ${imag.store_same}(_v.y);
        """,
    )


def combine_complex(output_arr_t: Any) -> Transformation:
    """``output = real + 1j * imag``."""
    out = _meta(output_arr_t)
    inp = ArrayMetadata(out.shape, dtypes.real_for(out.dtype))
    return Transformation(
        [
            Parameter("output", Annotation(out, "o")),
            Parameter("real", Annotation(inp, "i")),
            Parameter("imag", Annotation(inp, "i")),
        ],
        "${output.store_same}(COMPLEX_CTR(${output.ctype})(${real.load_same}, ${imag.load_same}));",
    )


def norm_const(arr_t: Any, order: float) -> Transformation:
    """``output = abs(input) ** order``."""
    inp = _meta(arr_t)
    out_dtype = dtypes.real_for(inp.dtype)
    return Transformation(
        [
            Parameter("output", Annotation(ArrayMetadata(inp.shape, out_dtype), "o")),
            Parameter("input", Annotation(inp, "i")),
        ],
        """
        ${output.ctype} _n = ${norm}(${input.load_same});
        %if order == 1:
        _n = sqrt(_n);
        %elif order != 2:
        _n = ${powf}(_n, ${dtypes.c_constant(order / 2, output.dtype)});
        %endif
        ${output.store_same}(_n);
        """,
        render_kwds=dict(norm=functions.norm(inp.dtype), order=order, dtypes=dtypes,
                         powf="powf" if out_dtype == numpy.float32 else "pow"),
    )


def norm_param(arr_t: Any) -> Transformation:
    """``output = abs(input) ** order`` with a run-time ``order``."""
    inp = _meta(arr_t)
    out_dtype = dtypes.real_for(inp.dtype)
    return Transformation(
        [
            Parameter("output", Annotation(ArrayMetadata(inp.shape, out_dtype), "o")),
            Parameter("input", Annotation(inp, "i")),
            Parameter("order", Annotation(out_dtype)),
        ],
        """
        ${output.ctype} _n = ${norm}(${input.load_same});
        _n = ${powf}(_n, ${order} / 2);
        ${output.store_same}(_n);
        """,
        render_kwds=dict(norm=functions.norm(inp.dtype), powf="powf" if out_dtype == numpy.float32 else "pow"),
    )


def ignore(arr_t: Any) -> Transformation:
    """Discards the output it is attached to."""
    return Transformation([Parameter("input", Annotation(_meta(arr_t), "i"))], "// ignored\n")


def broadcast_const(arr_t: Any, val: Any) -> Transformation:
    """``output = val``."""
    out = _meta(arr_t)
    val = numpy.asarray(val, out.dtype)
    if val.shape != ():
        raise ValueError("The constant must be a scalar")
    return Transformation(
        [Parameter("output", Annotation(out, "o"))],
        "${output.store_same}(${const});",
        render_kwds=dict(const=dtypes.c_constant(val, out.dtype)),
    )


def broadcast_param(arr_t: Any) -> Transformation:
    """``output = param``."""
    out = _meta(arr_t)
    return Transformation(
        [Parameter("output", Annotation(out, "o")), Parameter("param", Annotation(out.dtype))],
        "${output.store_same}(${param});",
    )
