"""GPU parity of complex128 transforms against numpy.fft in FP64.

The complex128 kernels are the complex64 ones instantiated for double2
elements (FP64 butterflies, FP64 twiddle and Bluestein tables, run-time
compiled).  The bound is written here: rel-L2 <= 5e-16 * log2 N, about four
units of FP64 roundoff per radix-2 level (measured: at most 7.4e-17 * log2 N
over these cases; complex64 twiddles or constants in the complex128 path
would give ~1e-8).
"""

import math

import numpy
import pytest

import oracle
from reikna_amd import Array
from reikna_amd.fft import FFT
from reikna_amd.transformations import div_const, mul_param

pytestmark = pytest.mark.gpu


def bound(shape, axes):
    return 5e-16 * max(1.0, math.log2(math.prod(shape[a] for a in axes)))


def data128(shape, seed):
    rng = numpy.random.default_rng(seed)
    return rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape)


def run(queue, data, axes, inverse=False, inplace=False):
    arr = Array.from_host(queue, data)
    fft = FFT(arr, axes=axes)
    c = fft.compile(queue.device)
    if inplace:
        c(queue, arr, arr, int(inverse))
        return arr.get(queue), fft
    out = Array.empty_like(queue.device, arr)
    c(queue, out, arr, int(inverse))
    return out.get(queue), fft


CASES = [
    ((16, 1024), (1,)),          # single pass
    ((8, 1000), (1,)),           # smooth, not a power of two
    ((4, 7 * 2**9), (1,)),       # radix 7
    ((8, 2**16), (1,)),          # two passes
    ((4, 997), (1,)),            # single-kernel Bluestein (M = 2048)
    ((4, 2003), (1,)),           # Bluestein chains (M = 4096)
    ((16, 32, 48), (0, 1, 2)),   # three axes
    ((64, 96), (0,)),            # strided axis
]


@pytest.mark.parametrize("shape,axes", CASES, ids=[f"{s}-{a}" for s, a in CASES])
def test_complex128_forward_inverse(queue, shape, axes):
    data = data128(shape, seed=sum(shape))
    res, fft = run(queue, data, axes)
    assert fft.native_plan.dtype == numpy.complex128
    err = oracle.rel_l2(res, numpy.fft.fftn(data, axes=axes))
    assert err <= bound(shape, axes), (shape, axes, err)
    back, _ = run(queue, data, axes, inverse=True, inplace=True)
    err = oracle.rel_l2(back, numpy.fft.ifftn(data, axes=axes))
    assert err <= bound(shape, axes), (shape, axes, "inverse", err)


def test_complex128_chunked_chain(queue):
    # 16 x 2^20 complex128 (256 MiB): chained passes run in chunks through the
    # Infinity Cache, with nontemporal hints on the caller's arrays
    shape = (16, 2**20)
    data = data128(shape, seed=91)
    res, fft = run(queue, data, (1,))
    assert fft.native_plan.chunk_rows > 0
    for r in (0, 15):
        err = oracle.rel_l2(res[r], numpy.fft.fft(data[r]))
        assert err <= bound(shape, (1,)), (r, err)


def test_complex128_fused_and_strided(queue):
    shape = (32, 4096)
    data = data128(shape, seed=92)
    arr = Array.from_host(queue, data)
    fft = FFT(arr, axes=(1,))
    scale = mul_param(arr, numpy.float64)
    fft.parameter.input.connect(scale, scale.output, input_prime=scale.input, param=scale.param)
    div = div_const(arr, 4.0)
    fft.parameter.output.connect(div, div.input, output_prime=div.output)
    out = Array.empty_like(queue.device, arr)
    fft.compile(queue.device)(queue, out, arr, 3.0, 0)
    ref = numpy.fft.fft(data * 3.0, axis=1) / 4.0
    assert oracle.rel_l2(out.get(queue), ref) <= bound(shape, (1,))
    # a strided view (every other element of each row of a larger array),
    # transformed in place
    from reikna_amd import ArrayMetadata

    base = data128((32, 8192), seed=93)
    meta = ArrayMetadata((32, 4096), numpy.complex128, strides=(8192 * 16, 32))
    view = Array(queue.device, meta)
    view.set(queue, base[:, ::2])
    FFT(meta, axes=(1,)).compile(queue.device)(queue, view, view, 0)
    ref = numpy.fft.fft(base[:, ::2], axis=1)
    assert oracle.rel_l2(view.get(queue), ref) <= bound(shape, (1,))


def test_complex128_trivial_copy(queue):
    data = data128((4, 1, 1), seed=94)
    res, _ = run(queue, data, (1, 2))
    assert numpy.array_equal(res, data)


@pytest.mark.parametrize("n", [1024, 3000, 2 * 997])
def test_float64_real_transforms(queue, n):
    """RFFT of float64 and IRFFT of complex128 run their half-length complex
    FFT in complex128."""
    from reikna_amd.fft.rfft import IRFFT, RFFT

    rng = numpy.random.default_rng(n)
    a = rng.normal(size=(4, n))
    a_dev = Array.from_host(queue, a)
    rf = RFFT(a_dev)
    out = Array.empty_like(queue.device, rf.parameter.output)
    rf.compile(queue.device)(queue, out, a_dev)
    assert out.dtype == numpy.complex128
    assert oracle.rel_l2(out.get(queue), numpy.fft.rfft(a, axis=1)) <= bound((n,), (0,))
    spec = numpy.fft.rfft(a, axis=1)
    s_dev = Array.from_host(queue, spec)
    irf = IRFFT(s_dev)
    back = Array.empty_like(queue.device, irf.parameter.output)
    irf.compile(queue.device)(queue, back, s_dev)
    assert back.dtype == numpy.float64
    ref = numpy.fft.irfft(spec, n=n, axis=1)
    assert numpy.linalg.norm(back.get(queue) - ref) / numpy.linalg.norm(ref) <= bound((n,), (0,))
