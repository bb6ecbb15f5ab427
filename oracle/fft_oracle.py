"""FP64 CPU restatement of the reference's FFT algorithm and oracle.

TEST INFRASTRUCTURE ONLY (see oracle/__init__.py).  Citations are to the
reference tree (reikna 0.10.0.dev).
"""

from __future__ import annotations

import math
from collections.abc import Sequence

import numpy


# ---------------------------------------------------------------------------
# the reference's oracle: numpy.fft in FP64 (tests/test_fft/test_fft.py:163,169)
# ---------------------------------------------------------------------------
def fftn(x: numpy.ndarray, axes: Sequence[int] | None = None, workers: int | None = None) -> numpy.ndarray:
    """Forward FFT over `axes` in complex128 (numpy.fft.fftn).  With
    `workers`, the same pocketfft transform through scipy.fft on that many
    threads (for checks of arrays of 10^8+ elements; pinned against the
    numpy path in tests/test_oracle.py)."""
    x = numpy.asarray(x, numpy.complex128)
    if workers:
        import scipy.fft

        return scipy.fft.fftn(x, axes=axes, workers=workers)
    return numpy.fft.fftn(x, axes=axes)


def ifftn(x: numpy.ndarray, axes: Sequence[int] | None = None, workers: int | None = None) -> numpy.ndarray:
    """Inverse FFT over `axes` in complex128, normalised by 1/prod(N) (numpy.fft.ifftn)."""
    x = numpy.asarray(x, numpy.complex128)
    if workers:
        import scipy.fft

        return scipy.fft.ifftn(x, axes=axes, workers=workers)
    return numpy.fft.ifftn(x, axes=axes)


def dft_direct(x: numpy.ndarray, inverse: bool = False) -> numpy.ndarray:
    """O(N^2) DFT along the last axis straight from the definition, in FP64:
    X[k] = sum_n x[n] exp(-+2 pi i n k / N) (inverse: conjugate kernel and 1/N).
    The phase index n*k is reduced modulo N exactly in integers first."""
    x = numpy.asarray(x, numpy.complex128)
    n = x.shape[-1]
    k = numpy.arange(n)
    e = numpy.outer(k, k) % n
    sign = 1.0 if inverse else -1.0
    w = numpy.exp(sign * 2j * numpy.pi * e / n)
    y = x @ w.T
    return y / n if inverse else y


# ---------------------------------------------------------------------------
# Bluestein (chirp-z) as the reference computes it
# ---------------------------------------------------------------------------
def chirp(n: int, sign: int = -1) -> numpy.ndarray:
    """exp(sign * i * pi * (m^2 mod 2n) / n) for m in [0, n): the reference's
    `xweight` (fft.mako:196-202) computes the phase index m*m mod 2N in exact
    integers before the trigonometric call; here the call is FP64."""
    m = numpy.arange(n, dtype=numpy.int64)
    ph = (m * m) % (2 * n)
    return numpy.exp(sign * 1j * numpy.pi * ph.astype(numpy.float64) / n)


def kweights(size_real: int, size_bound: int) -> numpy.ndarray:
    """The Bluestein kernel spectra of reikna/fft/fft.py:216-236: row 0 is
    FFT_M(exp(+i pi n^2 / N)) and row 1 is IFFT_M(exp(-i pi n^2 / N)) * M / N over
    the wrapped index n = [0 .. M-N, N-1 .. 1]."""
    nv = numpy.concatenate([numpy.arange(size_bound - size_real + 1), numpy.arange(size_real - 1, 0, -1)])
    arg = 1j * numpy.pi / size_real * nv.astype(numpy.float64) ** 2
    return numpy.vstack(
        [numpy.fft.fft(numpy.exp(arg)), numpy.fft.ifft(numpy.exp(-arg)) * size_bound / size_real]
    )


def bounding_power_of_2(n: int) -> int:
    """reikna/helpers.py:110-118."""
    return 1 if n <= 1 else 1 << (n - 1).bit_length()


def bluestein_fft(x: numpy.ndarray, inverse: bool = False) -> numpy.ndarray:
    """reikna's chirp-z transform along the last axis (fft.py:551-582), in FP64.

    The reference pads to M = 2 * bounding_power_of_2(N) (:559), runs a forward
    M-point FFT of x * chirp (pad_in, fft.mako:214-259), multiplies by the
    kernel spectrum, runs the inverse M-point FFT and multiplies by the chirp
    again, truncating to N (unpad_out, fft.mako:296-303).  The inverse
    transform conjugates the chirp direction and divides by N."""
    x = numpy.asarray(x, numpy.complex128)
    n = x.shape[-1]
    m = 2 * bounding_power_of_2(n)
    sign = 1 if inverse else -1
    c = chirp(n, sign)
    a = numpy.zeros(x.shape[:-1] + (m,), numpy.complex128)
    a[..., :n] = x * c
    b = numpy.zeros(m, numpy.complex128)
    b[:n] = numpy.conj(c)
    b[m - n + 1 :] = numpy.conj(c[1:][::-1])
    y = numpy.fft.ifft(numpy.fft.fft(a, axis=-1) * numpy.fft.fft(b), axis=-1)[..., :n] * c
    return y / n if inverse else y


# ---------------------------------------------------------------------------
# the reference's radix planners (for plan-count comparisons only)
# ---------------------------------------------------------------------------
_LOCAL_TABLE = {
    2: [2], 4: [4], 8: [8], 16: [8, 2], 32: [8, 4], 64: [8, 8], 128: [8, 4, 4],
    256: [4, 4, 4, 4], 512: [8, 8, 8], 1024: [16, 16, 4], 2048: [8, 8, 8, 4],
}


def radix_array(size: int, use_max_radix: bool = False) -> list[int]:
    """reikna/fft/fft.py:24-86: radices of the single-work-group kernel."""
    if size & (size - 1):
        raise ValueError("size must be a power of 2")
    if use_max_radix:
        radix = min(size, 16)
        out = []
        while size > radix:
            out.append(radix)
            size //= radix
        out.append(size)
        return out
    if size in _LOCAL_TABLE:
        return list(_LOCAL_TABLE[size])
    lsize = size.bit_length() - 1
    nblocks = (lsize - 1) // 4 + 1
    return [16] * (nblocks - 1) + [16 if lsize % 4 == 0 else 2 ** (lsize % 4)]


def global_radix_info(size: int) -> tuple[list[int], list[int], list[int]]:
    """reikna/fft/fft.py:89-146: the pass radices of the multi-kernel chain."""
    if size & (size - 1):
        raise ValueError("size must be a power of 2")
    base = min(size, 128)
    n = 0
    while size > base:
        size //= base
        n += 1
    radices = [base] * n + [size]
    r1s, r2s = [], []
    for r in radices:
        if r <= 8:
            r1s.append(r)
            r2s.append(1)
        else:
            r1 = 2
            r2 = r // r1
            while r2 > r1:
                r1 *= 2
                r2 = r // r1
            r1s.append(r1)
            r2s.append(r2)
    return radices, r1s, r2s


# ---------------------------------------------------------------------------
# measurement conventions
# ---------------------------------------------------------------------------
def fft_flops(shape: Sequence[int], axes: Sequence[int]) -> float:
    """5 * prod(shape) * sum_a log2(shape[a]), fractional log2 for non-powers of
    two (tests/test_fft/test_fft.py:234)."""
    return 5.0 * math.prod(shape) * sum(math.log2(shape[a]) for a in axes)


def rel_l2(x: numpy.ndarray, ref: numpy.ndarray) -> float:
    """||x - ref||_2 / ||ref||_2 in FP64."""
    x = numpy.asarray(x, numpy.complex128)
    ref = numpy.asarray(ref, numpy.complex128)
    den = numpy.linalg.norm(ref.ravel())
    return float(numpy.linalg.norm((x - ref).ravel()) / (den if den else 1.0))


def tolerance(shape: Sequence[int], axes: Sequence[int] | None = None) -> float:
    """The stated complex64 bound: rel-L2 <= 2e-7 * log2(N), N = transform size."""
    axes = range(len(shape)) if axes is None else axes
    n = math.prod(shape[a] for a in axes)
    return 2e-7 * max(1.0, math.log2(n))


def test_array(shape: Sequence[int], seed: int, dist: str = "uniform") -> numpy.ndarray:
    """Seeded complex64 input: re and im i.i.d. U[-1, 1) ("uniform") or N(0, 1) ("normal")."""
    rng = numpy.random.default_rng(seed)
    if dist == "uniform":
        re = rng.uniform(-1.0, 1.0, shape)
        im = rng.uniform(-1.0, 1.0, shape)
    elif dist == "normal":
        re = rng.standard_normal(shape)
        im = rng.standard_normal(shape)
    else:
        raise ValueError(dist)
    return (re + 1j * im).astype(numpy.complex64)


test_array.__test__ = False   # not a pytest test
