"""Spectrogram: matplotlib's ``specgram(mode="magnitude")`` scaled by the
window sum, as a chain of transformations around one FFT and a Transpose
(examples/demo_specgram.py:196-267).

The element-wise steps are transformations fused into the FFT's kernels:
on the input side a rolling frame (1-D signal -> overlapping rows, zero
padded to ``pad_to``), a von Hann window, and a real-to-complex combination
with a broadcast zero imaginary part; on the output side the magnitude and a
crop to the non-negative frequencies.  The (times, frequencies) result is then
transposed to (frequencies, times).  Signature ``(output:o, input:i)``.
"""

from __future__ import annotations

from collections.abc import Callable

import numpy

from .. import dtypes, functions, transformations
from ..algorithms import Transpose
from ..core import Annotation, Computation, ComputationPlan, KernelArguments, Parameter, Transformation
from ..runtime import ArrayMetadata, DeviceParameters
from .fft import FFT

__all__ = ["Spectrogram", "hanning_window", "rolling_frame", "crop_frequencies"]


def _meta(arr_t) -> ArrayMetadata:
    return arr_t.as_array_metadata() if hasattr(arr_t, "as_array_metadata") else ArrayMetadata.from_arraylike(arr_t)


def hanning_window(arr_t, fft_size: int) -> Transformation:
    """Rows multiplied by the von Hann window 0.5 (1 - cos(2 pi n / (fft_size - 1)))
    over their first fft_size elements (the zero padding after them is left as is)."""
    meta = _meta(arr_t)
    coeff_dtype = dtypes.real_for(meta.dtype) if dtypes.is_complex(meta.dtype) else meta.dtype
    return Transformation(
        [Parameter("output", Annotation(meta, "o")), Parameter("input", Annotation(meta, "i"))],
        """
        ${dtypes.ctype(coeff_dtype)} coeff = 1;
        if (${idxs[-1]} < ${fft_size})
            coeff = 0.5 * (1 - cos(${2 * numpy.pi} * ${idxs[-1]} / (${fft_size} - 1)));
        ${output.store_same}(${mul}(${input.load_same}, coeff));
        """,
        render_kwds=dict(dtypes=dtypes, numpy=numpy, coeff_dtype=coeff_dtype, fft_size=fft_size,
                         mul=functions.mul(meta.dtype, coeff_dtype, out_dtype=meta.dtype)),
    )


def rolling_frame(arr_t, fft_size: int, noverlap: int, pad_to: int | None) -> Transformation:
    """A 1-D signal as rows of fft_size samples starting every fft_size - noverlap
    samples, zero padded to pad_to (connector: output)."""
    meta = _meta(arr_t)
    if len(meta.shape) != 1:
        raise ValueError("the signal must be one-dimensional")
    step = fft_size - noverlap
    frames = (meta.shape[0] - noverlap) // step
    size = fft_size if pad_to is None else pad_to
    out = ArrayMetadata((frames, size), meta.dtype)
    return Transformation(
        [Parameter("output", Annotation(out, "o")), Parameter("input", Annotation(meta, "i"))],
        """
        if (${idxs[1]} >= ${fft_size})
            ${output.store_same}(0);
        else
            ${output.store_same}(${input.load_idx}(${idxs[0]} * ${step} + ${idxs[1]}));
        """,
        render_kwds=dict(step=step, fft_size=fft_size),
        connectors=["output"],
    )


def crop_frequencies(arr_t) -> Transformation:
    """Keeps columns 0 .. n/2 of (rows, n) (connector: input)."""
    meta = _meta(arr_t)
    out = ArrayMetadata((meta.shape[0], meta.shape[1] // 2 + 1), meta.dtype)
    return Transformation(
        [Parameter("output", Annotation(out, "o")), Parameter("input", Annotation(meta, "i"))],
        """
        if (${idxs[1]} < ${input.shape[1] // 2 + 1})
            ${output.store_idx}(${idxs[0]}, ${idxs[1]}, ${input.load_same});
        """,
        connectors=["input"],
    )


class Spectrogram(Computation):
    def __init__(
        self,
        x,
        fft_size: int = 256,
        noverlap: int = 128,
        pad_to: int | None = None,
        window: Callable = hanning_window,
    ):
        meta = _meta(x)
        if not dtypes.is_real(meta.dtype) or len(meta.shape) != 1:
            raise ValueError("the signal must be a real one-dimensional array")
        frame = rolling_frame(meta, fft_size, noverlap, pad_to)
        cdtype = dtypes.complex_for(meta.dtype)
        fft_arr = ArrayMetadata(frame.output.shape, cdtype)
        real_arr = ArrayMetadata(frame.output.shape, meta.dtype)
        win = window(real_arr, fft_size)
        zero = transformations.broadcast_const(real_arr, 0)
        to_complex = transformations.combine_complex(fft_arr)
        amplitude = transformations.norm_const(fft_arr, 1)
        crop = crop_frequencies(amplitude.output)

        fft = FFT(fft_arr, axes=(1,))
        fft.parameter.input.connect(to_complex, to_complex.output, input_real=to_complex.real,
                                    input_imag=to_complex.imag)
        fft.parameter.input_imag.connect(zero, zero.output)
        fft.parameter.input_real.connect(win, win.output, unwindowed_input=win.input)
        fft.parameter.unwindowed_input.connect(frame, frame.output, flat_input=frame.input)
        fft.parameter.output.connect(amplitude, amplitude.input, amplitude=amplitude.output)
        fft.parameter.amplitude.connect(crop, crop.input, cropped_amplitude=crop.output)
        self._fft = fft
        self._transpose = Transpose(fft.parameter.cropped_amplitude)
        Computation.__init__(
            self,
            [
                Parameter("output", Annotation(self._transpose.parameter.output, "o")),
                Parameter("input", Annotation(fft.parameter.flat_input, "i")),
            ],
        )

    def _build_plan(
        self, plan_factory: Callable[[], ComputationPlan], _device_params: DeviceParameters, args: KernelArguments
    ) -> ComputationPlan:
        plan = plan_factory()
        temp = plan.temp_array_like(self._fft.parameter.cropped_amplitude)
        plan.computation_call(self._fft, temp, args.input)
        plan.computation_call(self._transpose, args.output, temp)
        return plan
