"""Transpose (reikna/algorithms/transpose.py): the axis-group decomposition on
the CPU, generated kernels compiled for gfx950 without a GPU, and GPU results
against numpy.transpose on the reference's own cases
(tests/test_algorithms/test_transpose.py:11-45)."""

import itertools

import numpy
import pytest

from reikna_amd import Array, ArrayMetadata
from reikna_amd.algorithms import Transpose
from reikna_amd.algorithms.transpose import _swap, transpose_operations
from reikna_amd.runtime import VirtualDevice
from reikna_amd.transformations import mul_const

SHAPES = {
    2: [(13, 11), (511, 134), (77, 2049)],
    3: [(13, 33, 1029), (77, 55, 33)],
    4: [(11, 13, 19, 31), (35, 4, 57, 8)],
}
AXES = [None, (1, 0), (1, 0, 2), (2, 0, 1), (2, 3, 0, 1), (0, 3, 1, 2)]


def cases():
    out = []
    for dim, shapes in SHAPES.items():
        for shape, axes in itertools.product(shapes, [a for a in AXES if a is None or len(a) == dim]):
            out.append((shape, axes))
    return out


@pytest.mark.parametrize("n", [2, 3, 4, 5])
def test_operations_reach_every_permutation(n):
    for target in itertools.permutations(range(n)):
        ops = transpose_operations(n, target)
        axes = tuple(range(n))
        for b, c in ops:
            axes = _swap(axes, b, c)
        assert axes == target
        # a single group swap when the permutation is one
        one = any(_swap(tuple(range(n)), b, c) == target for b in range(n - 1) for c in range(b + 1, n))
        if one and target != tuple(range(n)):
            assert len(ops) == 1


def test_errors():
    arr = ArrayMetadata((4, 5, 6), numpy.int32)
    with pytest.raises(ValueError, match="distinct"):
        Transpose(arr, axes=(0, 0, 1))
    with pytest.raises(ValueError, match="Expected output array shape"):
        Transpose(arr, output_arr_t=ArrayMetadata((4, 5, 6), numpy.int32), axes=(2, 1, 0))
    with pytest.raises(ValueError, match="dtype"):
        Transpose(arr, output_arr_t=ArrayMetadata((6, 5, 4), numpy.float32))


@pytest.mark.parametrize("shape,axes", cases()[:6])
def test_kernels_compile(shape, axes):
    tr = Transpose(ArrayMetadata(shape, numpy.complex64), axes=axes)
    tr.compile(VirtualDevice())


@pytest.mark.gpu
@pytest.mark.parametrize("shape,axes", cases(), ids=[f"{s}-{a}" for s, a in cases()])
@pytest.mark.parametrize("dtype", [numpy.int32, numpy.complex64])
def test_transpose_gpu(queue, shape, axes, dtype):
    rng = numpy.random.default_rng(abs(hash((shape, axes))) % 2**32)
    a = (rng.integers(-1000, 1000, shape) + (1j * rng.normal(size=shape) if dtype == numpy.complex64 else 0)).astype(dtype)
    a_dev = Array.from_host(queue, a)
    tr = Transpose(a_dev, axes=axes)
    res = Array.empty_like(queue.device, tr.parameter.output)
    tr.compile(queue.device)(queue, res, a_dev)
    assert numpy.array_equal(res.get(queue), numpy.transpose(a, axes))


@pytest.mark.gpu
def test_transpose_with_transformation(queue):
    a = numpy.arange(64 * 130, dtype=numpy.float32).reshape(64, 130)
    a_dev = Array.from_host(queue, a)
    tr = Transpose(a_dev)
    scale = mul_const(tr.parameter.output, 3.0)
    tr.parameter.output.connect(scale, scale.input, scaled=scale.output)
    res = Array.empty(queue.device, (130, 64), numpy.float32)
    tr.compile(queue.device)(queue, res, a_dev)
    assert numpy.array_equal(res.get(queue), a.T * 3)
