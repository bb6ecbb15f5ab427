"""FFTShift (reikna/fft/fftshift.py) and the shift as an output
transformation of an FFT (examples/demo_fftshift_transformation.py), against
numpy.fft.fftshift / ifftshift, on the shapes of the reference's
tests/test_fft/test_fftshift.py:14-174 (1-D to 3-D, even and odd lengths,
partial axes, in place)."""

import numpy
import pytest

import oracle
from reikna_amd import Array, ArrayMetadata
from reikna_amd.fft import FFT
from reikna_amd.fft.fftshift import FFTShift, fftshift_transformation
from reikna_amd.runtime import VirtualDevice

CASES = [
    ((16,), None), ((17,), None), ((32, 16), None), ((31, 17), None), ((32, 17), (0,)), ((31, 16), (1,)),
    ((8, 16, 32), None), ((7, 16, 33), (0, 2)), ((1, 64), None), ((6, 1, 5), (1,)), ((128, 128), (0,)),
]


@pytest.mark.parametrize("shape,axes", CASES[:5])
def test_kernels_compile(shape, axes):
    FFTShift(ArrayMetadata(shape, numpy.complex64), axes=axes).compile(VirtualDevice())


def test_transformation_fused_into_fft_compiles():
    arr = ArrayMetadata((64, 96), numpy.complex64)
    fft = FFT(arr)
    sh = fftshift_transformation(arr)
    fft.parameter.output.connect(sh, sh.input, shifted=sh.output)
    assert list(fft.signature.parameters) == ["shifted", "input", "inverse"]
    fft.compile(VirtualDevice())


@pytest.mark.gpu
@pytest.mark.parametrize("inplace", [False, True])
@pytest.mark.parametrize("inverse", [0, 1])
@pytest.mark.parametrize("shape,axes", CASES, ids=[f"{s}-{a}" for s, a in CASES])
def test_fftshift_gpu(queue, shape, axes, inverse, inplace):
    a = oracle.test_array(shape, seed=len(shape) * 100 + shape[0])
    a_dev = Array.from_host(queue, a)
    sh = FFTShift(a_dev, axes=axes).compile(queue.device)
    out = a_dev if inplace else Array.empty_like(queue.device, a_dev)
    sh(queue, out, a_dev, inverse)
    ref = numpy.fft.ifftshift(a, axes=axes) if inverse else numpy.fft.fftshift(a, axes=axes)
    assert numpy.array_equal(out.get(queue), ref)


@pytest.mark.gpu
@pytest.mark.parametrize("shape,axes", [((64, 96), None), ((33, 250), (1,)), ((16, 24, 40), (0, 2))])
def test_fft_with_shift_transformation(queue, shape, axes):
    a = oracle.test_array(shape, seed=7)
    a_dev = Array.from_host(queue, a)
    fft = FFT(a_dev, axes=axes)
    sh = fftshift_transformation(a_dev, axes=axes)
    fft.parameter.output.connect(sh, sh.input, shifted=sh.output)
    out = Array.empty_like(queue.device, a_dev)
    fft.compile(queue.device)(queue, out, a_dev, 0)
    ref = numpy.fft.fftshift(oracle.fftn(a, axes if axes is not None else tuple(range(len(shape)))), axes=axes)
    assert oracle.rel_l2(out.get(queue), ref) <= oracle.tolerance(shape, axes)
