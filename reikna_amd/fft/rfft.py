"""Real-to-complex and complex-to-real FFTs through a half-length complex FFT
(examples/demo_specialized_fft.py:156-331, demo_specialized_fft.mako:1-62).

``RFFT(arr_t, dont_store_last=False)``: signature ``(output, input)``; a real
array of even last-axis length N gives the N/2 + 1 non-negative frequencies
of numpy.fft.rfft (N/2 without the last one).  The real input is read as
N/2 complex numbers x[n] = a[2n] + i a[2n+1] by an input transformation fused
into the FFT's load; one kernel then combines Z = FFT_{N/2}(x) into

    X[k] = Z[k] A[k] + conj(Z[N/2 - k]) B[k],  A = (1 - i W^k)/2,
    B = (1 + i W^k)/2,  W = exp(-2 pi i / N),  X[N/2] = Re Z[0] - Im Z[0]

(Rabiner 1979, the algorithm the reference's demo uses).  ``IRFFT(arr_t)``
runs the steps backwards: one kernel forms the half-length spectrum
conj(A) X[k] + conj(B) conj(X[N/2 - k]) (imaginary parts of X[0] and X[N/2]
ignored, as numpy.fft.irfft), an inverse FFT of N/2 points follows, and an
output transformation writes each complex result as two reals.  A and B are
computed in FP64 and rounded once to complex64.
"""

from __future__ import annotations

from collections.abc import Callable

import numpy

from .. import dtypes, functions, helpers
from ..core import Annotation, Computation, ComputationPlan, KernelArguments, Parameter, Transformation
from ..runtime import ArrayMetadata, DeviceParameters
from ..template import DefTemplate
from .fft import FFT

__all__ = ["RFFT", "IRFFT"]


def _meta(arr_t) -> ArrayMetadata:
    return arr_t.as_array_metadata() if hasattr(arr_t, "as_array_metadata") else ArrayMetadata.from_arraylike(arr_t)


def _coefficients(n: int) -> tuple[numpy.ndarray, numpy.ndarray]:
    w = numpy.exp(-2j * numpy.pi * numpy.arange(n // 2) / n)
    return 0.5 * (1 - 1j * w), 0.5 * (1 + 1j * w)


def _pairs_as_complex(real_arr: ArrayMetadata) -> Transformation:
    """(..., N) reals read as (..., N/2) complex numbers (a[2n], a[2n+1])."""
    out = ArrayMetadata((*real_arr.shape[:-1], real_arr.shape[-1] // 2), dtypes.complex_for(real_arr.dtype))
    return Transformation(
        [Parameter("output", Annotation(out, "o")), Parameter("input", Annotation(real_arr, "i"))],
        """
        <% lead = "".join(i + ", " for i in idxs[:-1]) %>
        ${output.store_same}(COMPLEX_CTR(${output.ctype})(
            ${input.load_idx}(${lead}${idxs[-1]} * 2),
            ${input.load_idx}(${lead}${idxs[-1]} * 2 + 1)));
        """,
        connectors=["output"],
    )


def _complex_as_pairs(cplx_arr: ArrayMetadata) -> Transformation:
    """(..., N/2) complex numbers written as (..., N) reals."""
    out = ArrayMetadata((*cplx_arr.shape[:-1], cplx_arr.shape[-1] * 2), dtypes.real_for(cplx_arr.dtype))
    return Transformation(
        [Parameter("output", Annotation(out, "o")), Parameter("input", Annotation(cplx_arr, "i"))],
        """
        <% lead = "".join(i + ", " for i in idxs[:-1]) %>
        ${input.ctype} v = ${input.load_same};
        ${output.store_idx}(${lead}${idxs[-1]} * 2, v.x);
        ${output.store_idx}(${lead}${idxs[-1]} * 2 + 1, v.y);
        """,
        connectors=["input"],
    )


_RFFT_OUT = r"""
${kernel_declaration}
{
    if (${static.skip}()) return;
    const VSIZE_T b = ${static.global_id}(0);
    const VSIZE_T k = ${static.global_id}(1);
    const VSIZE_T kr = k == 0 ? 0 : ${half} - k;
    const ${z.ctype} zk = ${z.load_combined_idx(slices)}(b, k);
    const ${z.ctype} zr = ${z.load_combined_idx(slices)}(b, kr);
    %if not dont_store_last:
    if (k == 0)
        ${output.store_combined_idx(slices)}(b, ${half}, COMPLEX_CTR(${output.ctype})(zk.x - zk.y, 0));
    %endif
    const ${A.ctype} ca = ${A.load_idx}(k);
    const ${B.ctype} cb = ${B.load_idx}(k);
    ${output.store_combined_idx(slices)}(b, k, ${add}(${mul}(zk, ca), ${mul}(${conj}(zr), cb)));
}
"""

_IRFFT_IN = r"""
${kernel_declaration}
{
    if (${static.skip}()) return;
    const VSIZE_T b = ${static.global_id}(0);
    const VSIZE_T k = ${static.global_id}(1);
    ${input.ctype} xk = ${input.load_combined_idx(slices)}(b, k);
    ${input.ctype} xr = ${input.load_combined_idx(slices)}(b, ${half} - k);
    if (k == 0) {
        xk.y = 0;
        xr.y = 0;
    }
    const ${A.ctype} ca = ${A.load_idx}(k);
    const ${B.ctype} cb = ${B.load_idx}(k);
    ${output.store_combined_idx(slices)}(b, k, ${add}(${mul}(xk, ca), ${mul}(${conj}(xr), cb)));
}
"""


class RFFT(Computation):
    """numpy.fft.rfft over the last axis of a real array of even length.
    Signature ``(output:o, input:i)``; output (..., N/2 + 1) complex, or
    (..., N/2) with ``dont_store_last``."""

    def __init__(self, arr_t, *, dont_store_last: bool = False):
        meta = _meta(arr_t)
        if dtypes.is_complex(meta.dtype):
            raise ValueError("RFFT takes a real array")
        if meta.shape[-1] % 2:
            raise ValueError("the last axis of an RFFT must have even length")
        self._dont_store_last = dont_store_last
        n_out = meta.shape[-1] // 2 + (0 if dont_store_last else 1)
        out = ArrayMetadata((*meta.shape[:-1], n_out), dtypes.complex_for(meta.dtype))
        Computation.__init__(
            self, [Parameter("output", Annotation(out, "o")), Parameter("input", Annotation(meta, "i"))]
        )

    def _build_plan(
        self, plan_factory: Callable[[], ComputationPlan], _device_params: DeviceParameters, args: KernelArguments
    ) -> ComputationPlan:
        plan = plan_factory()
        inp, out = args.input, args.output
        n = inp.shape[-1]
        a, b = _coefficients(n)
        a_arr = plan.persistent_array(a.astype(out.dtype))
        b_arr = plan.persistent_array(b.astype(out.dtype))
        half_arr = ArrayMetadata((*inp.shape[:-1], n // 2), out.dtype)
        cfft = FFT(half_arr, axes=(len(inp.shape) - 1,))
        pairs = _pairs_as_complex(_meta(inp))
        cfft.parameter.input.connect(pairs, pairs.output, real_input=pairs.input)
        z = plan.temp_array_like(cfft.parameter.output)
        plan.computation_call(cfft, z, inp)
        batch = helpers.product(inp.shape[:-1])
        template = DefTemplate.from_string("rfft_combine", ["kernel_declaration", "output", "z", "A", "B"], _RFFT_OUT)
        plan.kernel_call(
            template,
            [out, z, a_arr, b_arr],
            kernel_name="kernel_rfft_combine",
            global_size=(batch, n // 2),
            render_kwds=dict(
                slices=(len(inp.shape) - 1, 1),
                half=n // 2,
                dont_store_last=self._dont_store_last,
                mul=functions.mul(out.dtype, out.dtype),
                add=functions.add(out.dtype, out.dtype),
                conj=functions.conj(out.dtype),
            ),
        )
        return plan


class IRFFT(Computation):
    """numpy.fft.irfft over the last axis: (..., N/2 + 1) complex to (..., N)
    real.  Signature ``(output:o, input:i)``."""

    def __init__(self, arr_t):
        meta = _meta(arr_t)
        if not dtypes.is_complex(meta.dtype):
            raise ValueError("IRFFT takes a complex array")
        out = ArrayMetadata((*meta.shape[:-1], (meta.shape[-1] - 1) * 2), dtypes.real_for(meta.dtype))
        Computation.__init__(
            self, [Parameter("output", Annotation(out, "o")), Parameter("input", Annotation(meta, "i"))]
        )

    def _build_plan(
        self, plan_factory: Callable[[], ComputationPlan], _device_params: DeviceParameters, args: KernelArguments
    ) -> ComputationPlan:
        plan = plan_factory()
        inp, out = args.input, args.output
        n = (inp.shape[-1] - 1) * 2
        a, b = _coefficients(n)
        a_arr = plan.persistent_array(a.conj().astype(inp.dtype))
        b_arr = plan.persistent_array(b.conj().astype(inp.dtype))
        half_arr = ArrayMetadata((*inp.shape[:-1], n // 2), inp.dtype)
        cfft = FFT(half_arr, axes=(len(inp.shape) - 1,))
        pairs = _complex_as_pairs(half_arr)
        cfft.parameter.output.connect(pairs, pairs.input, real_output=pairs.output)
        spec = plan.temp_array_like(cfft.parameter.input)
        batch = helpers.product(inp.shape[:-1])
        template = DefTemplate.from_string("irfft_split", ["kernel_declaration", "output", "input", "A", "B"],
                                           _IRFFT_IN)
        plan.kernel_call(
            template,
            [spec, inp, a_arr, b_arr],
            kernel_name="kernel_irfft_split",
            global_size=(batch, n // 2),
            render_kwds=dict(
                slices=(len(inp.shape) - 1, 1),
                half=n // 2,
                mul=functions.mul(inp.dtype, inp.dtype),
                add=functions.add(inp.dtype, inp.dtype),
                conj=functions.conj(inp.dtype),
            ),
        )
        plan.computation_call(cfft, out, spec, 1)
        return plan
