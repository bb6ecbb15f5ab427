"""Serving FFT requests with continuous batching.

No reference counterpart (reikna has no serving layer).  :class:`FFTServer`
wraps the native request server (``rk_server_*`` in include/reikna_amd.h):
requests are 1-D or 2-D complex64 transforms with their input in device
memory; the server keeps a pass-chain template and its tables resident per
request shape, admits queued requests while their output and workspace fit
an HBM budget, and runs all admitted requests of a step in one launch of the
batch kernel, with several launches in flight.  ``mode="plain"`` runs the
same stream one request per plan execution for comparison.

:func:`run_stream` drives a server with a Poisson arrival process and
reports latency percentiles and throughput.
"""

from __future__ import annotations

import ctypes
import math
from collections.abc import Sequence
from dataclasses import dataclass, field

import numpy

from .. import _lib
from ..runtime import Array, Device, _ForeignBuffer, require_dense
from ..runtime import ArrayMetadata

__all__ = ["FFTServer", "RequestSpec", "make_specs", "run_stream", "sample_shapes"]


@dataclass(frozen=True)
class RequestSpec:
    """``count`` transforms of ``shape`` ((n,) or (n0, n1)) reading `input`."""

    shape: tuple
    count: int
    input: Array
    inverse: bool = False

    @property
    def elements(self) -> int:
        return self.count * math.prod(self.shape)

    @property
    def flops(self) -> float:
        return 5.0 * self.elements * math.log2(math.prod(self.shape))


@dataclass
class Completion:
    id: int
    out_ptr: int
    t_arrival_ns: int
    t_admitted_ns: int
    t_done_ns: int
    # 1: a dependency wait of its launch gave up; 2: its launch could not be
    # enqueued (the progress() call that tried raised); the output is not
    # valid in either case
    status: int = 0

    @property
    def latency_ms(self) -> float:
        return (self.t_done_ns - self.t_arrival_ns) / 1e6


def _check_spec(rid: int, spec: RequestSpec) -> None:
    """The kernels read count * prod(shape) complex64 values from the input."""
    need = int(spec.count) * math.prod(spec.shape)
    if spec.input.dtype != numpy.complex64 or spec.input.size < need:
        raise ValueError(f"request {rid}: input must be complex64 with at least {need} elements "
                         f"(got {spec.input.dtype} x {spec.input.size})")
    require_dense(f"request {rid} input", spec.input, 8 * need)


class FFTServer:
    def __init__(
        self,
        device: Device,
        budget_bytes: int,
        slots: int = 4,
        mode: str = "batched",
        max_requests: int = 1024,
        plain_min_elements: int = 0,
        spin_limit: int = 0,
        fault_inject: int = 0,
    ):
        if mode not in ("batched", "plain"):
            raise ValueError(f"unknown server mode {mode!r}")
        device.activate()
        self.device = device
        self.mode = mode
        cfg = _lib.ServerConfig()
        cfg.budget_bytes = int(budget_bytes)
        cfg.slots = int(slots)
        cfg.max_requests = int(max_requests)
        cfg.mode = _lib.RK_SERVER_BATCHED if mode == "batched" else _lib.RK_SERVER_PLAIN
        cfg.plain_min_elements = int(plain_min_elements)
        cfg.spin_limit = int(spin_limit)
        cfg.fault_inject = int(fault_inject)
        h = ctypes.c_void_p()
        _lib.call("rk_server_create", ctypes.byref(cfg), ctypes.byref(h))
        self._h = h.value
        self._done = (_lib.Completion * 4096)()

    @staticmethod
    def now_ns() -> int:
        return int(_lib.load().rk_server_now_ns())

    def submit(self, reqs: Sequence[tuple[int, RequestSpec, int]]) -> None:
        """Queue (id, spec, arrival time in now_ns() nanoseconds) triples.
        Inputs must already be in device memory: the server's streams do not
        wait for the stream that wrote them (synchronise it first)."""
        if not reqs:
            return
        arr = (_lib.Request * len(reqs))()
        for k, (rid, spec, t_arr) in enumerate(reqs):
            _check_spec(rid, spec)
            r = arr[k]
            r.id = int(rid)
            r.ndim = len(spec.shape)
            r.inverse = int(bool(spec.inverse))
            r.shape[0] = int(spec.shape[0])
            r.shape[1] = int(spec.shape[1]) if len(spec.shape) == 2 else 1
            r.count = int(spec.count)
            r.input = spec.input.data_ptr
            r.t_arrival_ns = int(t_arr)
        _lib.call("rk_server_submit", self._h, arr, len(reqs))

    def progress(self) -> list[Completion]:
        n = ctypes.c_int(0)
        _lib.call("rk_server_progress", self._h, self._done, len(self._done), ctypes.byref(n))
        return [
            Completion(c.id, int(c.out or 0), c.t_arrival_ns, c.t_admitted_ns, c.t_done_ns, c.status)
            for c in self._done[: n.value]
        ]

    def output(self, completion: Completion, spec: RequestSpec) -> Array:
        """The device output of a completed request (valid until release)."""
        meta = ArrayMetadata((spec.count, *spec.shape), numpy.complex64)
        return Array(self.device, meta, _ForeignBuffer(self.device, completion.out_ptr, meta.buffer_size, self))

    def release(self, rid: int) -> None:
        _lib.call("rk_server_release", self._h, ctypes.c_uint64(int(rid)))

    @property
    def stats(self) -> dict:
        st = _lib.ServerStats()
        _lib.call("rk_server_stats", self._h, ctypes.byref(st))
        return {f: int(getattr(st, f)) for f, _ in _lib.ServerStats._fields_}

    def close(self) -> None:
        if getattr(self, "_h", None):
            _lib.load().rk_server_destroy(self._h)
            self._h = None

    def __del__(self) -> None:
        try:
            self.close()
        except Exception:  # pragma: no cover - interpreter shutdown
            pass


def sample_shapes(
    n: int,
    seed: int = 7,
    one_d: tuple[int, int] = (8, 20),
    two_d: tuple[int, int] = (7, 11),
    max_count: int = 64,
    max_elements: int | None = None,
) -> list[tuple[tuple, int]]:
    """(shape, count) of `n` requests: half 1-D (n = 2^k, k uniform in
    one_d), half 2-D (n x n, n = 2^k, k uniform in two_d); count log-uniform
    in [1, max_count].  With `max_elements`, the count is capped so a request
    holds at most that many elements (no cap by default: the largest request
    of the default mix is 64 x 2048^2 = 2^28 elements, 2 GiB)."""
    rng = numpy.random.default_rng(seed)
    out = []
    for _ in range(n):
        if rng.random() < 0.5:
            shape = (2 ** int(rng.integers(one_d[0], one_d[1] + 1)),)
        else:
            m = 2 ** int(rng.integers(two_d[0], two_d[1] + 1))
            shape = (m, m)
        count = int(numpy.exp(rng.uniform(0.0, numpy.log(max_count + 1))))
        count = max(1, min(max_count, count))
        if max_elements is not None:
            count = max(1, min(count, max_elements // math.prod(shape)))
        out.append((shape, count))
    return out


def make_specs(queue, shapes: Sequence[tuple[tuple, int]], seed: int = 8, inverse: bool = False) -> list[RequestSpec]:
    """Requests over shared device inputs: one N(0, 1) complex64 array per
    distinct shape, as long as its largest request (read-only, so requests
    of one shape share it)."""
    need: dict = {}
    for shape, count in shapes:
        need[shape] = max(need.get(shape, 0), count)
    rng = numpy.random.default_rng(seed)
    inputs = {}
    for shape, count in sorted(need.items(), key=lambda kv: (len(kv[0]), kv[0])):
        n = count * math.prod(shape)
        host = numpy.empty(n, numpy.complex64)
        host.real = rng.standard_normal(n, dtype=numpy.float32)
        host.imag = rng.standard_normal(n, dtype=numpy.float32)
        inputs[shape] = Array.from_host(queue, host)
    # the server's streams do not wait for `queue`: the inputs must be in place
    queue.synchronize()
    return [RequestSpec(shape, count, inputs[shape], inverse) for shape, count in shapes]


@dataclass
class StreamResult:
    rate: float
    offered: int
    completed: int
    duration_s: float
    throughput_rps: float
    gflops: float
    p50_ms: float
    p99_ms: float
    mean_ms: float
    launches: int
    completions: list = field(default_factory=list, repr=False)


def run_stream(
    server: FFTServer,
    specs: Sequence[RequestSpec],
    rate: float,
    seed: int = 7,
    keep: set | None = None,
    on_complete=None,
) -> StreamResult:
    """Offer `specs` as a Poisson stream of `rate` requests/s; wait for all.
    Latency = completion time - scheduled arrival time (queueing included).
    Outputs are released right after completion, except ids in `keep`
    (``on_complete(completion, spec)`` runs first for every request).  Without
    `keep` and `on_complete` the stream is replayed by the native loop
    (rk_server_replay), which can offer far higher rates than Python."""
    rng = numpy.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, len(specs))
    t0 = server.now_ns() + 2_000_000
    arrivals = t0 + numpy.cumsum(gaps * 1e9).astype(numpy.int64)
    launches0 = server.stats["launches"]
    if keep is None and on_complete is None:
        return _replay_native(server, specs, rate, t0, arrivals, launches0)
    nxt = 0
    done: list[Completion] = []
    n = len(specs)
    while len(done) < n:
        now = server.now_ns()
        batch = []
        while nxt < n and arrivals[nxt] <= now:
            batch.append((nxt, specs[nxt], int(arrivals[nxt])))
            nxt += 1
        if batch:
            server.submit(batch)
        for c in server.progress():
            done.append(c)
            if on_complete is not None:
                on_complete(c, specs[c.id])
            if keep is None or c.id not in keep:
                server.release(c.id)
    t_end = max(c.t_done_ns for c in done)
    duration = (t_end - t0) / 1e9
    lat = numpy.array([c.latency_ms for c in done])
    flops = sum(s.flops for s in specs)
    return StreamResult(
        rate=rate,
        offered=n,
        completed=len(done),
        duration_s=duration,
        throughput_rps=n / duration,
        gflops=flops / duration / 1e9,
        p50_ms=float(numpy.percentile(lat, 50)),
        p99_ms=float(numpy.percentile(lat, 99)),
        mean_ms=float(lat.mean()),
        launches=server.stats["launches"] - launches0,
        completions=done,
    )


def _replay_native(server: FFTServer, specs, rate, t0, arrivals, launches0) -> StreamResult:
    n = len(specs)
    arr = (_lib.Request * max(1, n))()
    for k, spec in enumerate(specs):
        _check_spec(k, spec)
        r = arr[k]
        r.id = k
        r.ndim = len(spec.shape)
        r.inverse = int(bool(spec.inverse))
        r.shape[0] = int(spec.shape[0])
        r.shape[1] = int(spec.shape[1]) if len(spec.shape) == 2 else 1
        r.count = int(spec.count)
        r.input = spec.input.data_ptr
        r.t_arrival_ns = int(arrivals[k])
    done = (_lib.Completion * max(1, n))()
    _lib.call("rk_server_replay", server._h, arr, n, done, 0, 0)
    comps = [Completion(c.id, 0, c.t_arrival_ns, c.t_admitted_ns, c.t_done_ns, c.status) for c in done[:n]]
    if any(c.status for c in comps):
        raise RuntimeError("request server: a batch-kernel dependency wait gave up; results are not valid")
    t_end = max(c.t_done_ns for c in comps)
    duration = (t_end - t0) / 1e9
    lat = numpy.array([c.latency_ms for c in comps])
    flops = sum(s.flops for s in specs)
    return StreamResult(
        rate=rate,
        offered=n,
        completed=n,
        duration_s=duration,
        throughput_rps=n / duration,
        gflops=flops / duration / 1e9,
        p50_ms=float(numpy.percentile(lat, 50)),
        p99_ms=float(numpy.percentile(lat, 99)),
        mean_ms=float(lat.mean()),
        launches=server.stats["launches"] - launches0,
        completions=comps,
    )
