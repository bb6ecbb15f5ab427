"""rk_fft_execute_scratch (the entry point that allocates nothing: the
library's own device scratch array is the workspace) against rk_fft_execute
with a caller workspace, through the C-ABI: bit-identical outputs on two
calls in a row, the second with a smaller input, both checked against the FP64
oracle; a plan whose workspace exceeds the array is refused before any
launch."""

import ctypes

import numpy
import pytest

import oracle
from reikna_amd import Array, _lib
from reikna_amd.fft.native import NativeFftPlan
from reikna_amd.runtime import ArrayMetadata, Buffer

pytestmark = pytest.mark.gpu


def _plan(shape):
    meta = ArrayMetadata(shape, numpy.complex64)
    return NativeFftPlan(meta, meta, (len(shape) - 1,))


def _run_both(queue, shape, seed):
    plan = _plan(shape)
    assert 0 < plan.workspace_bytes <= _lib.RK_SCRATCH_BYTES, plan.workspace_bytes
    data = oracle.test_array(shape, seed=seed)
    inp = Array.from_host(queue, data)
    out_ws = Array.empty(queue.device, shape, numpy.complex64)
    out_scratch = Array.empty(queue.device, shape, numpy.complex64)
    ws = Buffer(queue.device, plan.workspace_bytes)
    h = ctypes.c_void_p(plan.handle)
    lib = _lib.load()
    for inverse in (0, 1):
        _lib.check(lib.rk_fft_execute(h, ctypes.c_void_p(out_ws.data_ptr), ctypes.c_void_p(inp.data_ptr), inverse,
                                      ctypes.c_void_p(ws.ptr), ctypes.c_void_p(queue.stream)), "rk_fft_execute")
        _lib.check(lib.rk_fft_execute_scratch(h, ctypes.c_void_p(out_scratch.data_ptr),
                                              ctypes.c_void_p(inp.data_ptr), inverse,
                                              ctypes.c_void_p(queue.stream)), "rk_fft_execute_scratch")
        a, b = out_ws.get(queue), out_scratch.get(queue)
        assert numpy.array_equal(a, b), f"{shape} inverse={inverse}: scratch and workspace results differ"
        ref = oracle.ifftn(data, (len(shape) - 1,)) if inverse else oracle.fftn(data, (len(shape) - 1,))
        assert oracle.rel_l2(b, ref) <= oracle.tolerance(shape, (len(shape) - 1,))


def test_scratch_matches_workspace_execute(queue):
    # a three-pass chain (two workspace buffers, 32 MiB), then a smaller
    # two-pass one in the same array
    _run_both(queue, (16, 2**17), seed=51)
    _run_both(queue, (8, 8192), seed=52)


def test_scratch_refuses_large_workspace(queue):
    plan = _plan((64, 2**20))
    assert plan.workspace_bytes > _lib.RK_SCRATCH_BYTES
    buf = Buffer(queue.device, 16)
    rc = _lib.load().rk_fft_execute_scratch(ctypes.c_void_p(plan.handle), ctypes.c_void_p(buf.ptr),
                                            ctypes.c_void_p(buf.ptr), 0, ctypes.c_void_p(queue.stream))
    assert rc == _lib.RK_ERR_INVALID
    assert "scratch array" in _lib.last_error()
