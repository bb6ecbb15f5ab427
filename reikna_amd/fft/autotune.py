"""Measured choice of batch planner knobs.

For a batch (its items), every candidate :class:`BatchOptions` is planned,
run on the device and timed with HIP events (median of ``reps`` launches after
``warmup``); the fastest wins.  Results are cached per device name and item
list in a JSON file under the run-time cache directory (``RK_CACHE_DIR`` or
``~/.cache/reikna_amd``), so a batch shape is tuned once per machine.

No reference counterpart: reikna picks local sizes by halving a limit until a
kernel compiles (reikna/fft/fft.py:723-756), not by timing.
"""

from __future__ import annotations

import hashlib
import json
import os
import statistics
from collections.abc import Iterable, Sequence
from dataclasses import asdict
from pathlib import Path

from ..runtime import Array, Device, Event, Queue
from .batch import BatchFFT, BatchItem, BatchOptions

__all__ = ["default_candidates", "autotune_batch"]


def default_candidates() -> list[BatchOptions]:
    out = []
    for single in (1024, 2048, 4096):
        for factor in (64, 128, 256):
            for order in (0, 2):
                out.append(BatchOptions(single_max=single, factor_max=factor, order=order))
    return out


def _cache_file() -> Path:
    root = Path(os.environ.get("RK_CACHE_DIR", str(Path.home() / ".cache" / "reikna_amd")))
    return root / "batch_autotune.json"


def _key(device: Device, items: Sequence[BatchItem]) -> str:
    h = hashlib.sha1()
    h.update(device.name.encode())
    h.update(json.dumps([asdict(it) for it in items]).encode())
    return h.hexdigest()


def time_batch(batch: BatchFFT, queue: Queue, out: Array, inp: Array, reps: int = 10, warmup: int = 2) -> float:
    """Median device time (ms) of one execute."""
    for _ in range(warmup):
        batch(queue, out, inp)
    times = []
    e0, e1 = Event(), Event()
    for _ in range(reps):
        e0.record(queue)
        batch(queue, out, inp)
        e1.record(queue)
        e1.synchronize()
        times.append(e0.elapsed_ms(e1))
    return statistics.median(times)


def autotune_batch(
    device: Device,
    items: Sequence[BatchItem],
    queue: Queue | None = None,
    candidates: Iterable[BatchOptions] | None = None,
    reps: int = 10,
    use_cache: bool = True,
) -> tuple[BatchOptions, list[tuple[BatchOptions, float]]]:
    """The fastest candidate for `items` and every (candidate, ms) measured
    (empty when the answer came from the cache)."""
    import numpy

    items = list(items)
    key = _key(device, items)
    path = _cache_file()
    cache: dict = {}
    if use_cache and path.exists():
        try:
            cache = json.loads(path.read_text())
        except (OSError, ValueError):
            cache = {}
        if key in cache:
            return BatchOptions(**cache[key]["options"]), []
    queue = queue or Queue(device)
    n_in = max((it.in_offset + (it.count - 1) * it.in_dist + it.length for it in items if it.count), default=1)
    n_out = max((it.out_offset + (it.count - 1) * it.out_dist + it.length for it in items if it.count), default=1)
    rng = numpy.random.default_rng(0)
    host = (rng.uniform(-1, 1, n_in) + 1j * rng.uniform(-1, 1, n_in)).astype(numpy.complex64)
    inp = Array.from_host(queue, host)
    out = Array.empty(device, (n_out,), numpy.complex64)
    results: list[tuple[BatchOptions, float]] = []
    for opt in candidates if candidates is not None else default_candidates():
        batch = BatchFFT(device, items, opt)
        results.append((opt, time_batch(batch, queue, out, inp, reps=reps)))
        del batch
    best = min(results, key=lambda r: r[1])[0]
    if use_cache:
        cache[key] = {"options": asdict(best), "ms": min(r[1] for r in results)}
        try:
            path.parent.mkdir(parents=True, exist_ok=True)
            tmp = path.with_suffix(".tmp")
            tmp.write_text(json.dumps(cache, indent=1))
            os.replace(tmp, path)
        except OSError:
            pass
    return best, results
