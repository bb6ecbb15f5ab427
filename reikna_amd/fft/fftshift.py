"""FFTShift (reikna/fft/fftshift.py:14-110) and the same shift as a
transformation (examples/demo_fftshift_transformation.py:24-72).

``FFTShift(arr_t, axes=None)`` with signature ``(output, input, inverse=0)``
is numpy.fft.fftshift (``inverse=0``) or ifftshift (``inverse=1``) over
``axes``.  When every shifted axis has even length the shift is an involution
that swaps element pairs, done in place by one kernel over half the array
(each thread exchanges one pair, so ``output`` may be ``input``); otherwise
the array is shifted out of place into a temporary and copied back, as in
reikna.  :func:`fftshift_transformation` gives the forward shift as an
output transformation (it moves elements, so the computation it is attached
to must not run in place).
"""

from __future__ import annotations

from collections.abc import Callable, Iterable, Sequence

import numpy

from .. import helpers
from ..algorithms import PureParallel
from ..core import Annotation, Computation, ComputationPlan, KernelArguments, Parameter, Transformation
from ..runtime import ArrayMetadata, DeviceParameters
from ..template import DefTemplate
from ..transformations import copy

__all__ = ["FFTShift", "fftshift_transformation"]


_INPLACE = r"""
${kernel_declaration}
{
    if (${static.skip}()) return;
    %for d in range(ndim):
    const VSIZE_T i${d} = ${static.global_id}(${d});
    %endfor
    %for d in range(ndim):
    <% n = shape[d] %>
    %if d in axes:
    const VSIZE_T j${d} = i${d} + ${n // 2} < ${n} ? i${d} + ${n // 2} : i${d} + ${n // 2} - ${n};
    %else:
    const VSIZE_T j${d} = i${d};
    %endif
    %endfor
    const ${output.ctype} a = ${input.load_idx}(${", ".join("i%d" % d for d in range(ndim))});
    const ${output.ctype} b = ${input.load_idx}(${", ".join("j%d" % d for d in range(ndim))});
    ${output.store_idx}(${", ".join("i%d" % d for d in range(ndim))}, b);
    ${output.store_idx}(${", ".join("j%d" % d for d in range(ndim))}, a);
}
"""

_OUTPLACE = r"""
${kernel_declaration}
{
    if (${static.skip}()) return;
    %for d in range(ndim):
    const VSIZE_T i${d} = ${static.global_id}(${d});
    %endfor
    %for d in range(ndim):
    <% n = shape[d] %>
    %if d in axes:
    const VSIZE_T s${d} = ${inverse} ? ${n - n // 2} : ${n // 2};
    const VSIZE_T j${d} = i${d} + s${d} < ${n} ? i${d} + s${d} : i${d} + s${d} - ${n};
    %else:
    const VSIZE_T j${d} = i${d};
    %endif
    %endfor
    ${output.store_idx}(${", ".join("j%d" % d for d in range(ndim))},
        ${input.load_idx}(${", ".join("i%d" % d for d in range(ndim))}));
}
"""


class FFTShift(Computation):
    """
    Shift the zero-frequency component to the centre of the spectrum, as
    numpy.fft.fftshift (``inverse=0``) or numpy.fft.ifftshift (``inverse=1``).

    :param arr_t: an array-like defining the problem array.
    :param axes: the axes to shift (all if not given).

    Signature: ``(output:o, input:i, inverse:int32=0)``; ``output`` and
    ``input`` may be the same array.
    """

    def __init__(self, arr_t, axes: Iterable[int] | None = None):
        arr = arr_t.as_array_metadata() if hasattr(arr_t, "as_array_metadata") else ArrayMetadata.from_arraylike(arr_t)
        Computation.__init__(
            self,
            [
                Parameter("output", Annotation(arr, "o")),
                Parameter("input", Annotation(arr, "i")),
                Parameter("inverse", Annotation(numpy.int32), default=0),
            ],
        )
        self._axes = tuple(range(len(arr.shape)) if axes is None else axes)

    def _build_plan(
        self, plan_factory: Callable[[], ComputationPlan], _device_params: DeviceParameters, args: KernelArguments
    ) -> ComputationPlan:
        plan = plan_factory()
        output, inp, inverse = args.output, args.input, args.inverse
        shape = tuple(inp.shape)
        if helpers.product([shape[a] for a in self._axes]) == 1:
            # nothing moves; a copy still runs attached transformations
            trf = copy(inp, out_arr_t=output)
            plan.computation_call(PureParallel.from_trf(trf, trf.input), output, inp)
            return plan
        axes = tuple(sorted(self._axes))
        ndim = len(shape)
        if all(shape[a] % 2 == 0 for a in axes):
            half = list(shape)
            half[axes[0]] //= 2
            template = DefTemplate.from_string("fftshift_inplace", ["kernel_declaration", "output", "input"], _INPLACE)
            plan.kernel_call(template, [output, inp], kernel_name="kernel_fftshift_inplace", global_size=half,
                             render_kwds=dict(ndim=ndim, shape=shape, axes=axes))
        else:
            temp = plan.temp_array_like(output)
            template = DefTemplate.from_string("fftshift_outplace",
                                               ["kernel_declaration", "output", "input", "inverse"], _OUTPLACE)
            plan.kernel_call(template, [temp, inp, inverse], kernel_name="kernel_fftshift_outplace",
                             global_size=shape, render_kwds=dict(ndim=ndim, shape=shape, axes=axes))
            trf = copy(inp, out_arr_t=output)
            plan.computation_call(PureParallel.from_trf(trf, trf.input), output, temp)
        return plan


def fftshift_transformation(arr_t, axes: Sequence[int] | None = None) -> Transformation:
    """An output transformation ``output = numpy.fft.fftshift(input, axes)``
    (connector: ``input``).  It writes each element to its shifted position,
    so the computation it is attached to must not run in place."""
    meta = arr_t.as_array_metadata() if hasattr(arr_t, "as_array_metadata") else ArrayMetadata.from_arraylike(arr_t)
    norm_axes = tuple(range(len(meta.shape)) if axes is None else sorted(axes))
    code = r"""
        %for d in range(len(output.shape)):
        <% n = output.shape[d] %>
        VSIZE_T rk_shift${d} = ${idxs[d]}
        %if d in axes:
            + (${idxs[d]} < ${n - n // 2} ? ${n // 2} : ${-(n - n // 2)})
        %endif
        ;
        %endfor
        ${output.store_idx}(${", ".join("rk_shift%d" % d for d in range(len(output.shape)))}, ${input.load_same});
    """
    return Transformation(
        [Parameter("output", Annotation(meta, "o")), Parameter("input", Annotation(meta, "i"))],
        code,
        connectors=["input"],
        render_kwds=dict(axes=norm_axes),
    )
