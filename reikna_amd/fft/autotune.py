"""Measured choice of planner knobs.

Two searches, both timing candidates on the device with HIP events (median
of ``reps`` executions after ``warmup``) and caching the winner per device
and problem in a JSON file under the run-time cache directory
(``RK_CACHE_DIR`` or ``~/.cache/reikna_amd``), so a problem is tuned once
per machine:

* :func:`autotune_batch` — the batch planner's knobs (single-pass limit,
  chain pass limit, queue order) for a list of uneven transforms;
* :func:`autotune_fft` — one transform's plan: the longest chain pass and
  the first pass length (its radix split), column / row tile widths, the
  chunk size of chained passes and compiled-in load/store modes
  (:class:`PlanTuning`).  A candidate whose result departs from the default
  plan's by more than twice the complex64 error bound is rejected.

No reference counterpart: reikna picks local sizes by halving a limit until a
kernel compiles (reikna/fft/fft.py:723-756), not by timing.
"""

from __future__ import annotations

import hashlib
import json
import os
import statistics
from collections.abc import Iterable, Sequence
from typing import Any
from dataclasses import asdict
from pathlib import Path

from .. import _lib
from ..runtime import Array, ArrayMetadata, Device, Event, Queue
from .batch import BatchFFT, BatchItem, BatchOptions
from .native import NativeFftPlan, PlanTuning

__all__ = ["default_candidates", "autotune_batch", "autotune_fft", "PlanTuning"]


def default_candidates() -> list[BatchOptions]:
    out = []
    for single in (1024, 2048, 4096):
        for factor in (64, 128, 256, 512, 1024):
            for order in (0, 2):
                out.append(BatchOptions(single_max=single, factor_max=factor, order=order))
    return out


def _cache_file(name: str = "batch_autotune.json") -> Path:
    root = Path(os.environ.get("RK_CACHE_DIR", str(Path.home() / ".cache" / "reikna_amd")))
    return root / name


def _read_cache(path: Path) -> dict:
    try:
        return json.loads(path.read_text()) if path.exists() else {}
    except (OSError, ValueError):
        return {}


def _write_cache(path: Path, cache: dict) -> None:
    try:
        path.parent.mkdir(parents=True, exist_ok=True)
        tmp = path.with_suffix(".tmp")
        tmp.write_text(json.dumps(cache, indent=1))
        os.replace(tmp, path)
    except OSError:
        pass


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
    cache = _read_cache(path) if use_cache else {}
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
        _write_cache(path, cache)
    return best, results


# ---------------------------------------------------------------------------
# single-transform plans
# ---------------------------------------------------------------------------

def _knob_values(native: NativeFftPlan, nbytes: int) -> dict[str, list]:
    """Values worth trying for each knob of a plan, from its default pass list."""
    passes = native.passes
    by_axis: dict[int, list[int]] = {}
    for p in passes:
        by_axis.setdefault(p.axis, []).append(p.L)
    multi = {a: ls for a, ls in by_axis.items() if len(ls) > 1}
    knobs: dict[str, list] = {}
    if multi:
        # the planner's limit is 1024 points for complex64 chains, 256 for complex128
        knobs["factor_max"] = [0, 128, 256, 512, 1024]
        firsts = set()
        for ls in multi.values():
            n = 1
            for f in ls:
                n *= f
            firsts.update(f for f in set(ls) if f != ls[0])
            firsts.update(f for f in (32, 64, 128, 256) if n % f == 0 and f != ls[0] and n // f > 1)
        knobs["first_factor"] = [0] + sorted(firsts)[:4]
    # column-mapped passes have tiles of 16+ instances, row-mapped ones fewer
    cols = {p.C for p in passes if p.C >= 16}
    rows = {p.C for p in passes if p.C < 16}
    if cols:
        knobs["col_tile"] = [0] + [c for c in (16, 32, 64) if c not in cols]
    if rows:
        knobs["row_tile"] = [0] + [c for c in (1, 2, 4, 8) if c not in rows]
    if native.chunk_rows or 2 * nbytes > (256 << 20):
        knobs["chunk_bytes"] = [-1, 0, 64 << 20, 128 << 20, 256 << 20]
    knobs["fixed_modes"] = [-1, 0]
    return knobs


def _plan_key(device: Device, shape: Sequence[int], axes: Sequence[int], dtype: Any) -> str:
    return f"{device.name}|{tuple(shape)}|{tuple(axes)}|{dtype}"


def autotune_fft(
    queue: Queue,
    shape: Sequence[int],
    axes: Sequence[int] | None = None,
    reps: int = 10,
    warmup: int = 10,
    use_cache: bool = True,
    dtype: Any = "complex64",
) -> tuple[PlanTuning, list[tuple[PlanTuning, float]]]:
    """The fastest :class:`PlanTuning` for a C-contiguous complex64 or
    complex128 array of ``shape`` transformed over ``axes``, and every
    (tuning, ms) measured (empty when the answer came from the cache).

    One coordinate sweep from the planner's own plan: each knob in turn takes
    each of its values with the others at the best found so far, and a value
    replaces the current one only when it is more than 2 % faster.  A
    candidate the planner or the run-time compiler refuses is skipped."""
    import math

    import numpy

    shape = tuple(int(n) for n in shape)
    axes = tuple(range(len(shape))) if axes is None else tuple(int(a) for a in axes)
    dtype = numpy.dtype(dtype)
    if dtype not in (numpy.complex64, numpy.complex128):
        raise ValueError(f"plans are tuned for complex64 or complex128 arrays, not {dtype}")
    key = _plan_key(queue.device, shape, axes, dtype)
    path = _cache_file("fft_autotune.json")
    cache = _read_cache(path) if use_cache else {}
    if key in cache:
        return PlanTuning(**cache[key]["tuning"]), []

    meta = ArrayMetadata(shape, dtype)
    rng = numpy.random.default_rng(0)
    host = (rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape)).astype(dtype)
    inp = Array.from_host(queue, host)
    out = Array.empty(queue.device, shape, dtype)
    ref_out: numpy.ndarray | None = None
    # twice the stated bound of the element type (two plans, each within it)
    unit = 2e-7 if dtype == numpy.complex64 else 5e-16
    bound = 2 * unit * max(1.0, math.log2(math.prod(shape[a] for a in axes)))

    def measure(t: PlanTuning) -> float | None:
        nonlocal ref_out
        try:
            plan = NativeFftPlan(meta, meta, axes, 0, None if t == PlanTuning() else t)
        except (ValueError, RuntimeError):
            return None
        ws = Array.empty(queue.device, (plan.workspace_bytes,), numpy.uint8) if plan.workspace_bytes else None
        run = lambda: plan.execute(out.ptr, inp.ptr, 0, ws.ptr if ws is not None else None, queue.stream)  # noqa: E731
        for _ in range(warmup):
            run()
        e0, e1 = Event(), Event()
        times = []
        for _ in range(reps):
            e0.record(queue)
            run()
            e1.record(queue)
            e1.synchronize()
            times.append(e0.elapsed_ms(e1))
        res = out.get(queue)
        if ref_out is None:
            ref_out = res
        else:
            den = float(numpy.linalg.norm(ref_out.astype(numpy.complex128)))
            err = float(numpy.linalg.norm((res - ref_out).astype(numpy.complex128))) / max(den, 1e-30)
            if not err <= bound:
                return None
        return statistics.median(times)

    best = PlanTuning()
    best_ms = measure(best)
    if best_ms is None:
        raise ValueError(f"no plan for {shape} over axes {axes}")
    results: list[tuple[PlanTuning, float]] = [(best, best_ms)]
    default_plan = NativeFftPlan(meta, meta, axes, _lib.RK_FFT_DRY_RUN)
    for knob, values in _knob_values(default_plan, host.nbytes).items():
        for v in values:
            cand = PlanTuning(**{**asdict(best), knob: v})
            if cand == best:
                continue
            ms = measure(cand)
            if ms is None:
                continue
            results.append((cand, ms))
            if ms < 0.98 * best_ms:
                best, best_ms = cand, ms
    if use_cache:
        cache[key] = {"tuning": asdict(best), "ms": best_ms}
        _write_cache(path, cache)
    return best, results
