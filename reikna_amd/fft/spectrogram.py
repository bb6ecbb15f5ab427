"""Spectrogram: the magnitude of the short-time Fourier transform of a real
signal, as ``matplotlib.mlab.specgram(mode="magnitude")`` computes it times
the window's sum (the quantity examples/demo_specgram.py:196-267 builds with
reikna), laid out (frequencies, frames).

One FFT does the work; everything around it is fused into that FFT's first
and last pass, so the signal is read from HBM by the FFT kernel itself and
the spectrum never exists in complex form in memory:

* input side, one transformation (:func:`framed_signal`): output element
  ``(f, p)`` of the FFT's complex input is ``signal[f * step + p] *
  window[p]`` for ``p < fft_size`` and 0 in the zero padding up to
  ``pad_to``, with a zero imaginary part.  The window is an array parameter,
  held by the computation as a persistent device array (any window;
  default the symmetric von Hann window, computed in FP64 and rounded once);
* output side, one transformation (:func:`half_spectrum_magnitude`): the
  FFT's (frames, pad_to) complex output is consumed as ``|X|`` of the
  non-negative frequencies only, a real (frames, pad_to // 2 + 1) array;
* a :class:`Transpose` then lays it out (frequencies, frames).

Signature ``(output:o, input:i)``.
"""

from __future__ import annotations

from collections.abc import Callable

import numpy

from .. import dtypes
from ..algorithms import Transpose
from ..core import Annotation, Computation, ComputationPlan, KernelArguments, Parameter, Transformation
from ..runtime import ArrayMetadata, DeviceParameters
from .fft import FFT

__all__ = ["Spectrogram", "hann_window", "framed_signal", "half_spectrum_magnitude"]


def _meta(arr_t) -> ArrayMetadata:
    return arr_t.as_array_metadata() if hasattr(arr_t, "as_array_metadata") else ArrayMetadata.from_arraylike(arr_t)


def hann_window(n: int) -> numpy.ndarray:
    """The symmetric von Hann window 0.5 - 0.5 cos(2 pi k / (n - 1)), k < n, in
    FP64 (numpy.hanning, matplotlib's default specgram window)."""
    if n == 1:
        return numpy.ones(1)
    k = numpy.arange(n, dtype=numpy.float64)
    return 0.5 - 0.5 * numpy.cos(2.0 * numpy.pi * k / (n - 1))


def framed_signal(signal_t, fft_size: int, noverlap: int, pad_to: int | None = None) -> Transformation:
    """Transformation (connector ``output``) from a real 1-D signal and a real
    window of ``fft_size`` values to the complex (frames, pad_to) input of a
    row FFT: frame f is ``signal[f * step : f * step + fft_size] * window``
    (step = fft_size - noverlap) followed by zeros up to ``pad_to``.
    Parameters: ``output``, ``signal``, ``window``."""
    sig = _meta(signal_t)
    if len(sig.shape) != 1 or not dtypes.is_real(sig.dtype):
        raise ValueError("the signal must be a real one-dimensional array")
    if not 0 <= noverlap < fft_size:
        raise ValueError(f"noverlap must be in [0, fft_size), got {noverlap} for fft_size {fft_size}")
    size = fft_size if pad_to is None else int(pad_to)
    if size < fft_size:
        raise ValueError(f"pad_to ({size}) must be at least fft_size ({fft_size})")
    step = fft_size - noverlap
    frames = (sig.shape[0] - noverlap) // step
    if frames < 1:
        raise ValueError(f"a signal of {sig.shape[0]} samples holds no frame of {fft_size}")
    cdtype = dtypes.complex_for(sig.dtype)
    out = ArrayMetadata((frames, size), cdtype)
    win = ArrayMetadata((fft_size,), sig.dtype)
    return Transformation(
        [
            Parameter("output", Annotation(out, "o")),
            Parameter("signal", Annotation(sig, "i")),
            Parameter("window", Annotation(win, "i")),
        ],
        """
        ${rtype} v = 0;
        if (${idxs[1]} < ${fft_size})
            v = ${signal.load_idx}(${idxs[0]} * ${step} + ${idxs[1]}) * ${window.load_idx}(${idxs[1]});
        ${output.store_same}(COMPLEX_CTR(${ctype})(v, 0));
        """,
        render_kwds=dict(rtype=dtypes.ctype(sig.dtype), ctype=dtypes.ctype(cdtype), fft_size=fft_size, step=step),
        connectors=["output"],
    )


def half_spectrum_magnitude(spectrum_t) -> Transformation:
    """Transformation (connector ``spectrum``) from the complex (rows, n)
    output of a row FFT to ``|X|`` of its columns 0 .. n // 2, a real
    (rows, n // 2 + 1) array.  Parameters: ``magnitude``, ``spectrum``."""
    spec = _meta(spectrum_t)
    if len(spec.shape) != 2 or not dtypes.is_complex(spec.dtype):
        raise ValueError("the spectrum must be a complex two-dimensional array")
    half = spec.shape[1] // 2 + 1
    rdtype = dtypes.real_for(spec.dtype)
    out = ArrayMetadata((spec.shape[0], half), rdtype)
    return Transformation(
        [Parameter("magnitude", Annotation(out, "o")), Parameter("spectrum", Annotation(spec, "i"))],
        """
        if (${idxs[1]} < ${half}) {
            const ${ctype} x = ${spectrum.load_same};
            ${magnitude.store_idx}(${idxs[0]}, ${idxs[1]}, hypot(x.x, x.y));
        }
        """,
        render_kwds=dict(half=half, ctype=dtypes.ctype(spec.dtype)),
        connectors=["spectrum"],
    )


class Spectrogram(Computation):
    """``(output, input)``: output[k, f] = |sum_p w[p] x[f * step + p]
    exp(-2 pi i k p / pad_to)| for the non-negative frequencies k of frame f.

    :param x: the real 1-D signal (array or metadata).
    :param fft_size: samples per frame.
    :param noverlap: samples shared by consecutive frames.
    :param pad_to: transform length (zero padding after the frame); default fft_size.
    :param window: ``fft_size`` window values, or a function of the length
        returning them (default :func:`hann_window`).
    """

    def __init__(
        self,
        x,
        fft_size: int = 256,
        noverlap: int = 128,
        pad_to: int | None = None,
        window: Callable[[int], numpy.ndarray] | numpy.ndarray | None = None,
    ):
        sig = _meta(x)
        framing = framed_signal(sig, fft_size, noverlap, pad_to)
        win = numpy.asarray(window(fft_size) if callable(window) else
                            (hann_window(fft_size) if window is None else window), numpy.float64)
        if win.shape != (fft_size,):
            raise ValueError(f"the window must hold fft_size = {fft_size} values, got shape {win.shape}")
        self._window = win.astype(sig.dtype)

        fft = FFT(framing.output, axes=(1,))
        fft.parameter.input.connect(framing, framing.output, signal=framing.signal, window=framing.window)
        mag = half_spectrum_magnitude(fft.parameter.output)
        fft.parameter.output.connect(mag, mag.spectrum, magnitude=mag.magnitude)
        self._fft = fft
        self._transpose = Transpose(fft.parameter.magnitude)
        Computation.__init__(
            self,
            [
                Parameter("output", Annotation(self._transpose.parameter.output, "o")),
                Parameter("input", Annotation(fft.parameter.signal, "i")),
            ],
        )

    def _build_plan(
        self, plan_factory: Callable[[], ComputationPlan], _device_params: DeviceParameters, args: KernelArguments
    ) -> ComputationPlan:
        plan = plan_factory()
        window = plan.persistent_array(self._window)
        magnitude = plan.temp_array_like(self._fft.parameter.magnitude)
        plan.computation_call(self._fft, magnitude, args.input, window)
        plan.computation_call(self._transpose, args.output, magnitude)
        return plan
