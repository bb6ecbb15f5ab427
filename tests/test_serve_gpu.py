"""GPU tests of the request server (continuous batching of FFT requests).

Every completed request is checked against numpy.fft in FP64 at the
complex64 bound; a request computed alone and inside a mixed stream gives
bit-identical results (its pass chain depends only on its shape); the HBM
budget bounds what is admitted at any time.
"""

import math

import numpy
import pytest

import oracle
from reikna_amd import _lib
from reikna_amd._lib import BudgetError
from reikna_amd.fft.serve import FFTServer, RequestSpec, make_specs, run_stream, sample_shapes

pytestmark = pytest.mark.gpu

GiB = 2**30


def reference(spec: RequestSpec, host_in: numpy.ndarray) -> numpy.ndarray:
    x = host_in[: spec.elements].reshape(spec.count, *spec.shape)
    axes = tuple(range(1, 1 + len(spec.shape)))
    return oracle.ifftn(x, axes) if spec.inverse else oracle.fftn(x, axes)


def check_output(server, queue, c, spec, cache):
    key = id(spec.input)
    if key not in cache:
        cache[key] = spec.input.get(queue)
    out = server.output(c, spec).get(queue)
    ref = reference(spec, cache[key])
    err = oracle.rel_l2(out, ref)
    tol = oracle.tolerance(spec.shape)
    assert err <= tol, f"request {c.id} {spec.shape} x {spec.count}: rel-L2 {err:.3g} > {tol:.3g}"


def drain(server, specs, queue=None, check=True, keep=()):
    server.submit([(i, s, 0) for i, s in enumerate(specs)])
    done, cache, outs = 0, {}, {}
    while done < len(specs):
        for c in server.progress():
            done += 1
            if check:
                check_output(server, queue, c, specs[c.id], cache)
            if c.id in keep:
                outs[c.id] = server.output(c, specs[c.id]).get(queue)
            server.release(c.id)
    return outs


@pytest.mark.parametrize("mode", ["batched", "plain"])
def test_mixed_requests_match_oracle(queue, mode):
    shapes = sample_shapes(40, seed=17, one_d=(8, 16), two_d=(7, 9), max_count=4)
    shapes += [((997,), 2), ((65536,), 1), ((96, 160), 1), ((1000,), 3)]
    specs = make_specs(queue, shapes, seed=18)
    server = FFTServer(queue.device, budget_bytes=1 * GiB, slots=3, mode=mode)
    drain(server, specs, queue)
    st = server.stats
    assert st["completed"] == len(specs) and st["pending"] == 0 and st["budget_used"] == 0
    if mode == "batched":
        assert st["launches"] < len(specs)   # requests shared launches
    server.close()


def test_inverse_requests(queue):
    shapes = [((1024,), 2), ((128, 128), 1), ((5003,), 1)]
    specs = make_specs(queue, shapes, seed=19, inverse=True)
    server = FFTServer(queue.device, budget_bytes=256 * 2**20)
    drain(server, specs, queue)
    server.close()


def test_alone_and_in_mix_identical(queue):
    shapes = sample_shapes(48, seed=20, one_d=(8, 17), two_d=(7, 10), max_count=8)
    specs = make_specs(queue, shapes, seed=21)
    keep = set(range(0, len(specs), 5))
    server = FFTServer(queue.device, budget_bytes=2 * GiB, slots=4)
    res = run_stream(server, specs, rate=20000.0, seed=22, keep=keep,
                     on_complete=lambda c, s: None)
    mixed = {}
    for c in res.completions:
        if c.id in keep:
            mixed[c.id] = server.output(c, specs[c.id]).get(queue)
            server.release(c.id)
    server.close()
    for rid in sorted(keep):
        alone = FFTServer(queue.device, budget_bytes=2 * GiB, slots=1)
        out = drain(alone, [specs[rid]], queue, check=True, keep={0})[0]
        alone.close()
        assert numpy.array_equal(out, mixed[rid]), f"request {rid} {specs[rid].shape} x {specs[rid].count}"


def test_budget_bounds_admission(queue):
    shapes = [((4096,), 16)] * 12          # 512 KiB output each
    specs = make_specs(queue, shapes, seed=23)
    budget = 3 * 512 * 1024 + 4096         # three requests' outputs at a time
    server = FFTServer(queue.device, budget_bytes=budget, slots=4)
    server.submit([(i, s, 0) for i, s in enumerate(specs)])
    done, peak, cache = 0, 0, {}
    while done < len(specs):
        comps = server.progress()
        st = server.stats
        peak = max(peak, st["budget_used"])
        assert st["budget_used"] <= budget
        for c in comps:
            check_output(server, queue, c, specs[c.id], cache)
            server.release(c.id)
            done += 1
    assert peak > 0
    with pytest.raises(BudgetError):
        server.submit([(99, make_specs(queue, [((65536,), 8)], seed=24)[0], 0)])
    server.close()


def test_large_requests_run_plain(queue):
    shapes = [((2**18,), 4), ((512,), 3), ((256, 256), 2), ((2**20,), 1)]
    specs = make_specs(queue, shapes, seed=25)
    server = FFTServer(queue.device, budget_bytes=1 * GiB, slots=2, plain_min_elements=2**20)
    drain(server, specs, queue)
    server.close()


def test_stream_reports_latency(queue):
    shapes = sample_shapes(60, seed=26, one_d=(8, 14), two_d=(7, 8), max_count=4)
    specs = make_specs(queue, shapes, seed=27)
    server = FFTServer(queue.device, budget_bytes=1 * GiB)
    res = run_stream(server, specs, rate=5000.0, seed=28)
    assert res.completed == len(specs)
    assert 0 < res.p50_ms <= res.p99_ms
    assert res.throughput_rps > 0 and math.isfinite(res.gflops)
    server.close()


def test_ids_in_use_and_busy_replay_refused(queue):
    specs = make_specs(queue, [((1024,), 2), ((256,), 1)], seed=29)
    server = FFTServer(queue.device, budget_bytes=64 * 2**20)
    server.submit([(0, specs[0], 0)])
    with pytest.raises(Exception, match="still in use"):
        server.submit([(0, specs[1], 0)])
    # a native replay needs an idle server (its completions are all its own)
    with pytest.raises(Exception, match="in flight"):
        run_stream(server, specs, rate=1000.0)
    drain_ids = 0
    while drain_ids < 1:
        for c in server.progress():
            check_output(server, queue, c, specs[0], {})
            server.release(c.id)
            drain_ids += 1
    server.submit([(0, specs[1], 0)])   # id free again after release
    while not server.progress():
        pass
    server.release(0)
    res = run_stream(server, specs, rate=1000.0)
    assert res.completed == 2
    server.close()


def test_failed_waits_are_reported(queue):
    """A launch whose dependency waits give up (forced with a poll limit of 1:
    every wait is bounded, so every kernel still ends) is reported on its
    completions and in the stats, and the server keeps serving."""
    specs = make_specs(queue, [((65536,), 16), ((65536,), 16)], seed=40)
    server = FFTServer(queue.device, budget_bytes=256 * 2**20, slots=1, spin_limit=1)
    server.submit([(i, s, 0) for i, s in enumerate(specs)])
    comps = []
    while len(comps) < len(specs):
        comps += server.progress()
    assert any(c.status == 1 for c in comps)
    assert server.stats["failed_launches"] >= 1
    for c in comps:
        server.release(c.id)
    # a single-pass request needs no waits: served correctly after the failure
    small = make_specs(queue, [((1024,), 4)], seed=41)
    server.submit([(10, small[0], 0)])
    done = []
    while not done:
        done = server.progress()
    assert done[0].status == 0
    check_output(server, queue, done[0], small[0], {})
    server.release(10)
    # the native replay raises on a failed launch
    with pytest.raises(RuntimeError, match="gave up"):
        run_stream(server, specs, rate=1000.0)
    server.close()


def test_failed_slot_allocation_is_reported(queue):
    """A slot whose buffer cannot grow (an allocation failure injected from
    the host, before any launch) fails its launch cleanly: progress() raises,
    the admitted requests complete with status 2, their memory is released,
    and the server serves the next requests correctly."""
    specs = make_specs(queue, [((4096,), 4), ((2048,), 8)], seed=42)
    server = FFTServer(queue.device, budget_bytes=64 * 2**20, slots=1,
                       fault_inject=_lib.RK_SERVER_FAULT_SLOT_ALLOC)
    server.submit([(i, s, 0) for i, s in enumerate(specs)])
    with pytest.raises(Exception, match="injected"):
        server.progress()
    comps = server.progress()
    assert sorted(c.id for c in comps) == [0, 1] and all(c.status == 2 for c in comps)
    assert server.stats["failed_launches"] == 1
    for c in comps:
        server.release(c.id)
    assert server.stats["budget_used"] == 0
    server.submit([(i + 10, s, 0) for i, s in enumerate(specs)])
    done = []
    while len(done) < 2:
        done += server.progress()
    for c in done:
        assert c.status == 0
        check_output(server, queue, c, specs[c.id - 10], {})
        server.release(c.id)
    server.close()


def test_submit_is_all_or_nothing(queue):
    """A batch with a bad request (an id given twice, or one still in use) is
    refused whole: none of its requests is queued."""
    specs = make_specs(queue, [((1024,), 2), ((256,), 1)], seed=43)
    server = FFTServer(queue.device, budget_bytes=64 * 2**20)
    with pytest.raises(Exception, match="twice"):
        server.submit([(0, specs[0], 0), (1, specs[1], 0), (0, specs[1], 0)])
    assert server.stats["pending"] == 0
    server.submit([(5, specs[0], 0)])
    with pytest.raises(Exception, match="still in use"):
        server.submit([(6, specs[1], 0), (5, specs[1], 0)])
    assert server.stats["pending"] + server.stats["in_flight_requests"] + server.stats["completed"] == 1
    # ids 0, 1, 6 were never queued: they are free
    server.submit([(0, specs[0], 0), (1, specs[1], 0), (6, specs[1], 0)])
    done = []
    while len(done) < 4:
        done += server.progress()
    for c in done:
        spec = specs[0] if c.id in (0, 5) else specs[1]
        check_output(server, queue, c, spec, {})
        server.release(c.id)
    server.close()


@pytest.mark.timeout(600)
def test_stated_serving_mix(queue):
    """The serving workload as stated: 1-D 256..2^20 and 2-D 128^2..2048^2
    transforms, batch 1..64 (no cap: the largest request is 64 x 2048^2),
    N(0, 1) inputs, arrivals at one Poisson rate; every completion checked
    against numpy.fft in FP64 (threaded pocketfft for these sizes)."""
    import os

    shapes = sample_shapes(24, seed=7)
    assert max(c * math.prod(s) for s, c in shapes) > 2**25   # requests beyond the old cap
    specs = make_specs(queue, shapes, seed=8)
    server = FFTServer(queue.device, budget_bytes=16 * GiB, slots=4, plain_min_elements=2**22)
    keep = set(range(len(specs)))
    res = run_stream(server, specs, rate=2000.0, seed=7, keep=keep)
    assert res.completed == len(specs) and 0 < res.p50_ms <= res.p99_ms
    workers = min(16, os.cpu_count() or 1)
    hosts = {}
    for c in sorted(res.completions, key=lambda c: c.id):
        spec = specs[c.id]
        assert c.status == 0
        key = id(spec.input)
        if key not in hosts:
            hosts[key] = spec.input.get(queue)
        x = hosts[key][: spec.elements].reshape(spec.count, *spec.shape)
        ref = oracle.fftn(x, tuple(range(1, 1 + len(spec.shape))), workers=workers)
        out = server.output(c, spec).get(queue)
        server.release(c.id)
        err = oracle.rel_l2(out, ref)
        tol = oracle.tolerance(spec.shape)
        assert err <= tol, f"request {c.id} {spec.shape} x {spec.count}: rel-L2 {err:.3g} > {tol:.3g}"
        del ref, out
    assert server.stats["budget_used"] == 0
    server.close()
