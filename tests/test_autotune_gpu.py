"""GPU tests of planner knobs and the plan autotuner: every tuned plan stays
within the complex64 bound (rel-L2 <= 2e-7 * log2 N) of numpy.fft in FP64,
whichever pass split, tile width, chunking or kernel twin it runs."""

import numpy
import pytest

import oracle
from reikna_amd import Array
from reikna_amd.fft import FFT, PlanTuning

pytestmark = pytest.mark.gpu


def tuned_result(queue, data, axes, tuning):
    arr = Array.from_host(queue, data)
    out = Array.empty_like(queue.device, arr)
    fft = FFT(arr, axes=axes, tuning=tuning)
    fft.compile(queue.device)(queue, out, arr, 0)
    return out.get(queue), fft.native_plan


@pytest.mark.parametrize("knobs", [
    {"first_factor": 64}, {"factor_max": 1024}, {"factor_max": 64}, {"col_tile": 64}, {"col_tile": 16},
    {"chunk_bytes": 0}, {"chunk_bytes": 8 << 20}, {"fixed_modes": 0},
], ids=lambda k: "-".join(f"{a}={b}" for a, b in k.items()))
def test_tuned_plans_match_oracle(queue, knobs):
    # 256 MiB: large enough for chained passes to run in chunks
    shape, axes = (128, 2**18), (1,)
    data = oracle.test_array(shape, seed=71)
    res, plan = tuned_result(queue, data, axes, PlanTuning(**knobs))
    for r in (0, 63, 127):
        err = oracle.rel_l2(res[r], oracle.fftn(data[r], (0,)))
        assert err <= oracle.tolerance(shape[1:], (0,)), (knobs, r, err, [(p.L, p.C) for p in plan.passes])
    if "chunk_bytes" in knobs:
        assert (plan.chunk_rows > 0) == (knobs["chunk_bytes"] > 0)


def test_autotune_fft_picks_a_correct_plan_and_caches(queue, tmp_path, monkeypatch):
    from reikna_amd.fft.autotune import autotune_fft

    monkeypatch.setenv("RK_CACHE_DIR", str(tmp_path))
    shape, axes = (8, 3 * 2**16), (1,)
    best, results = autotune_fft(queue, shape, axes, reps=3, warmup=1)
    assert results and results[0][0] == PlanTuning()
    assert all(ms > 0 for _, ms in results)
    times = dict(results)
    assert best in times and times[best] <= times[PlanTuning()]
    data = oracle.test_array(shape, seed=72)
    res, _ = tuned_result(queue, data, axes, best)
    assert oracle.rel_l2(res, oracle.fftn(data, (1,), workers=16)) <= oracle.tolerance(shape, axes)
    # the second call answers from the cache without timing anything
    again, measured = autotune_fft(queue, shape, axes)
    assert again == best and measured == []


def test_autotune_fft_complex128(queue, tmp_path, monkeypatch):
    from reikna_amd.fft.autotune import autotune_fft

    monkeypatch.setenv("RK_CACHE_DIR", str(tmp_path))
    shape, axes = (16, 2**16), (1,)
    best, results = autotune_fft(queue, shape, axes, reps=3, warmup=1, dtype="complex128")
    assert results and best in dict(results)
    data = oracle.test_array(shape, seed=73).astype(numpy.complex128)
    res, plan = tuned_result(queue, data, axes, best)
    assert plan.dtype == numpy.complex128
    bound = 5e-16 * 16
    assert oracle.rel_l2(res, numpy.fft.fft(data, axis=1)) <= bound


@pytest.mark.parametrize(
    "shape,dtype,factor_max,lengths",
    [
        ((8, 2**20), numpy.complex128, 1024, [1024, 1024]),   # complex128 passes beyond the default 256
        ((4, 2**22), numpy.complex64, 2048, [2048, 2048]),    # 2048-point column tiles (8 instances)
        ((2, 2**22), numpy.complex128, 2048, [2048, 2048]),   # the same in FP64 (4 instances)
    ],
    ids=["c128-1024", "c64-2048", "c128-2048"],
)
def test_tuned_long_chain_passes(queue, shape, dtype, factor_max, lengths):
    """Chain passes longer than the planner's own limit, as a tuned plan may
    choose them (kernels from the run-time compiler): forward and inverse
    against numpy.fft in FP64 at the element type's bound."""
    data = oracle.test_array(shape, seed=81).astype(dtype)
    res, plan = tuned_result(queue, data, (1,), PlanTuning(factor_max=factor_max))
    assert [p.L for p in plan.passes] == lengths
    n = shape[1]
    bound = (5e-16 if dtype == numpy.complex128 else 2e-7) * numpy.log2(n)
    assert oracle.rel_l2(res, numpy.fft.fft(data.astype(numpy.complex128), axis=1)) <= bound
    arr = Array.from_host(queue, res.astype(dtype))
    fft = FFT(arr, axes=(1,), tuning=PlanTuning(factor_max=factor_max))
    fft.compile(queue.device)(queue, arr, arr, 1)
    assert oracle.rel_l2(arr.get(queue), data) <= 2 * bound
