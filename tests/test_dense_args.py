"""Entry points that take bare data pointers (BatchFFT, the request server)
refuse strided, broadcast and short arrays before any launch (CPU only: the
arrays wrap a dummy pointer that nothing dereferences)."""

import numpy
import pytest

from reikna_amd.fft.serve import RequestSpec, _check_spec
from reikna_amd.runtime import Array, ArrayMetadata, _ForeignBuffer, require_dense


def fake_array(shape, dtype=numpy.complex64, strides=None, offset=0, buffer_size=None):
    meta = ArrayMetadata(shape, dtype, strides=strides, first_element_offset=offset, buffer_size=buffer_size)
    size = meta.buffer_size if buffer_size is None else buffer_size
    return Array(None, meta, _ForeignBuffer(None, 0x1000, size, None))


def test_dense_array_accepted():
    require_dense("x", fake_array((4, 256)), 8 * 1024)
    # dimensions of length one may carry any stride
    require_dense("x", fake_array((1, 1024), strides=(12345, 8)), 8 * 1024)


def test_strided_view_refused():
    # every other element of a 2048-element buffer
    arr = fake_array((1024,), strides=(16,), buffer_size=8 * 2048)
    with pytest.raises(ValueError, match="not C-contiguous"):
        require_dense("input", arr, 8 * 1024)
    # a column slice of a matrix
    arr = fake_array((256, 4), strides=(64, 8))
    with pytest.raises(ValueError, match="not C-contiguous"):
        require_dense("input", arr, 8 * 1024)


def test_zero_stride_view_refused():
    # a broadcast (expand) of one element: the buffer holds 8 bytes
    arr = fake_array((1024,), strides=(0,), buffer_size=8)
    with pytest.raises(ValueError, match="not C-contiguous"):
        require_dense("input", arr, 8 * 1024)


def test_short_buffer_refused():
    # an offset view whose buffer ends before first element + need
    arr = fake_array((1024,), offset=64, buffer_size=8 * 1024)
    with pytest.raises(ValueError, match="needed"):
        require_dense("output", arr, 8 * 1024)
    with pytest.raises(ValueError, match="needed"):
        require_dense("output", fake_array((100,)), 8 * 101)


def test_request_spec_checks_layout():
    good = RequestSpec((16, 16), 2, fake_array((512,)))
    _check_spec(0, good)
    strided = RequestSpec((256,), 1, fake_array((256,), strides=(16,), buffer_size=8 * 512))
    with pytest.raises(ValueError, match="not C-contiguous"):
        _check_spec(1, strided)
    broadcast = RequestSpec((256,), 1, fake_array((256,), strides=(0,), buffer_size=8))
    with pytest.raises(ValueError, match="not C-contiguous"):
        _check_spec(2, broadcast)
