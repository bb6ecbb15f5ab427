"""Seeded random FFT configurations against the FP64 oracle: 1-3 dimensional
arrays of smooth, prime and mixed lengths, random axis subsets, forward and
inverse, in and out of place, strided input views.  Every case at the
complex64 bound rel-L2 <= 2e-7 * log2 N."""

import math

import numpy
import pytest

import oracle
from reikna_amd import Array
from reikna_amd.fft import FFT

pytestmark = pytest.mark.gpu

LENGTHS = [1, 2, 3, 5, 7, 8, 12, 16, 17, 30, 31, 64, 97, 100, 128, 243, 256, 343, 500, 512, 625, 1000, 1024,
           1031, 2048, 3000, 4096, 4099, 6144, 8192, 10007, 16384, 32768, 65536]


def random_case(rng):
    ndim = int(rng.integers(1, 4))
    budget = 2**22
    shape = []
    for _ in range(ndim):
        cap = max(1, budget // max(1, math.prod(shape)))
        choices = [n for n in LENGTHS if n <= cap] or [1]
        shape.append(int(rng.choice(choices)))
    k = int(rng.integers(1, ndim + 1))
    axes = tuple(sorted(int(a) for a in rng.choice(ndim, size=k, replace=False)))
    return tuple(shape), axes


@pytest.mark.parametrize("seed", range(40))
def test_random_configuration(queue, seed):
    rng = numpy.random.default_rng(1000 + seed)
    shape, axes = random_case(rng)
    data = oracle.test_array(shape, seed=seed)
    inverse = bool(rng.integers(2))
    inplace = bool(rng.integers(2))
    arr = Array.from_host(queue, data)
    fft = FFT(arr, axes=axes).compile(queue.device)
    out = arr if inplace else Array.empty_like(queue.device, arr)
    fft(queue, out, arr, int(inverse))
    res = out.get(queue)
    ref = oracle.ifftn(data, axes) if inverse else oracle.fftn(data, axes)
    if all(shape[a] == 1 for a in axes):
        assert numpy.array_equal(res, data)
        return
    err = oracle.rel_l2(res, ref)
    tol = oracle.tolerance(shape, axes)
    assert err <= tol, f"{shape} axes {axes} inverse={inverse} inplace={inplace}: rel-L2 {err:.3g} > {tol:.3g}"


@pytest.mark.parametrize("seed", range(8))
def test_random_strided_views(queue, seed):
    """Input and output with the layout of a view taking every other element
    of the last axis of a larger array (the computation's signature gives
    both the same type, so the output array has that layout too)."""
    from reikna_amd import ArrayMetadata

    rng = numpy.random.default_rng(2000 + seed)
    shape, axes = random_case(rng)
    if math.prod(shape) > 2**20:
        shape = tuple(min(n, 64) for n in shape)
    base = oracle.test_array(shape[:-1] + (2 * shape[-1],), seed=seed)
    view = base[..., ::2]
    meta = ArrayMetadata(view.shape, view.dtype, strides=view.strides, buffer_size=base.nbytes)
    inp = Array(queue.device, meta)
    inp.set(queue, view)
    out = Array(queue.device, meta)
    fft = FFT(meta, axes=axes).compile(queue.device)
    fft(queue, out, inp, 0)
    res = out.get(queue)
    if all(shape[a] == 1 for a in axes):
        assert numpy.array_equal(res, view)
        return
    ref = oracle.fftn(numpy.ascontiguousarray(view), axes)
    assert oracle.rel_l2(res, ref) <= oracle.tolerance(shape, axes), (shape, axes)


def test_mismatched_output_array_refused(queue):
    """A contiguous output for a computation compiled with a strided output
    layout spans less memory than the kernels would write: refused."""
    from reikna_amd import ArrayMetadata

    base = oracle.test_array((2 * 512,), seed=1)
    view = base[::2]
    meta = ArrayMetadata(view.shape, view.dtype, strides=view.strides, buffer_size=base.nbytes)
    inp = Array(queue.device, meta)
    inp.set(queue, view)
    fft = FFT(meta, axes=(0,)).compile(queue.device)
    small = Array.empty(queue.device, view.shape, numpy.complex64)
    with pytest.raises(ValueError, match="strides"):
        fft(queue, small, inp, 0)


@pytest.mark.parametrize("seed", range(16))
def test_random_configuration_complex128(queue, seed):
    """The same random configurations in complex128 (run-time compiled FP64
    kernels), at rel-L2 <= 5e-16 * log2 N against numpy.fft."""
    rng = numpy.random.default_rng(3000 + seed)
    shape, axes = random_case(rng)
    if math.prod(shape) > 2**20:
        shape = tuple(min(n, 1024) for n in shape)
    data = oracle.test_array(shape, seed=seed).astype(numpy.complex128)
    inverse = bool(rng.integers(2))
    arr = Array.from_host(queue, data)
    out = Array.empty_like(queue.device, arr)
    FFT(arr, axes=axes).compile(queue.device)(queue, out, arr, int(inverse))
    res = out.get(queue)
    if all(shape[a] == 1 for a in axes):
        assert numpy.array_equal(res, data)
        return
    ref = numpy.fft.ifftn(data, axes=axes) if inverse else numpy.fft.fftn(data, axes=axes)
    tol = 5e-16 * max(1.0, math.log2(math.prod(shape[a] for a in axes)))
    err = oracle.rel_l2(res, ref)
    assert err <= tol, f"{shape} axes {axes} inverse={inverse}: rel-L2 {err:.3g} > {tol:.3g}"
