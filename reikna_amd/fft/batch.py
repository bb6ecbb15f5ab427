"""Batches of 1-D complex64 transforms of uneven lengths.

No reference counterpart: reikna builds one computation per array shape
(reikna/fft/fft.py:630-756), so a batch of transforms of different lengths
costs one compiled computation and one chain of launches per length.  Here a
:class:`BatchFFT` plans every transform of the batch as a chain of passes and
runs all of them in one launch of the batch kernel (``mode="queue"``); the
per-item path (``mode="grouped"``: one single-shape plan per item, launched
one after the other) is kept for comparison.

Transforms follow numpy.fft conventions: forward ``X[k] = sum_n x[n]
exp(-2 pi i k n / N)``, inverse with ``+`` and a factor 1/N (or the factor on
the forward transform for items created with ``scale=True``).
"""

from __future__ import annotations

import ctypes
from collections.abc import Sequence
from dataclasses import dataclass

from .. import _lib
from ..runtime import Array, Buffer, Device, Queue, VirtualDevice, require_dense

__all__ = ["BatchFFT", "BatchItem", "BatchOptions"]


@dataclass(frozen=True)
class BatchItem:
    """``count`` transforms of ``length`` complex64 elements; transform j of
    the item reads ``in[in_offset + j * in_dist + n]`` and writes
    ``out[out_offset + j * out_dist + k]`` (offsets and distances in
    complex64 elements).

    With ``r2c=True`` the item is numpy.fft.rfft of ``length`` (even, >= 4)
    float32 values: transform j reads them from complex64 element
    ``in_offset + j * in_dist`` on (as length/2 complex values) and writes
    the length/2 + 1 non-negative frequencies; forward only, one launch with
    the rest of the batch (a half-length complex FFT and the combine pass of
    examples/demo_specialized_fft.py:156-331 inside the batch kernel)."""

    length: int
    count: int = 1
    in_offset: int = 0
    out_offset: int = 0
    in_dist: int = 0
    out_dist: int = 0
    scale: bool = False   # forward result multiplied by 1/N (inverse unscaled)
    r2c: bool = False

    @property
    def in_length(self) -> int:
        """complex64 elements one transform reads"""
        return self.length // 2 if self.r2c else self.length

    @property
    def out_length(self) -> int:
        """complex64 elements one transform writes"""
        return self.length // 2 + 1 if self.r2c else self.length


@dataclass(frozen=True)
class BatchOptions:
    """Planner knobs (the autotuner searches over them; 0 = library default)."""

    mode: str = "queue"      # "queue" (one launch) or "grouped" (one plan per item)
    single_max: int = 0      # longest single-pass length (<= 4096)
    factor_max: int = 0      # longest pass of a multi-pass chain (<= 1024; 0: 512)
    blue_max: int = 0        # reserved (Bluestein runs as two chains in the batch kernel)
    order: int = 0           # queue order: 0 critical path, 1 levels in item order, 2 levels, largest tiles first

    def to_c(self, dry: bool = False) -> _lib.BatchOptions:
        if self.mode not in ("queue", "grouped"):
            raise ValueError(f"unknown batch mode {self.mode!r}")
        o = _lib.BatchOptions()
        o.mode = _lib.RK_BATCH_MODE_QUEUE if self.mode == "queue" else _lib.RK_BATCH_MODE_GROUPED
        o.flags = _lib.RK_BATCH_DRY_RUN if dry else 0
        o.single_max = int(self.single_max)
        o.factor_max = int(self.factor_max)
        o.blue_max = int(self.blue_max)
        o.order = int(self.order)
        return o


def packed_items(
    lengths: Sequence[int],
    counts: Sequence[int] | None = None,
    scale: bool = False,
    r2c: Sequence[bool] | bool = False,
) -> list[BatchItem]:
    """Items laid out back to back in one input and one output buffer
    (``r2c``: per item or for all, real-to-complex items)."""
    counts = list(counts) if counts is not None else [1] * len(lengths)
    if len(counts) != len(lengths):
        raise ValueError("lengths and counts differ in size")
    kinds = [bool(r2c)] * len(lengths) if isinstance(r2c, bool) else [bool(x) for x in r2c]
    if len(kinds) != len(lengths):
        raise ValueError("lengths and r2c differ in size")
    items, ioff, ooff = [], 0, 0
    for n, c, real in zip(lengths, counts, kinds):
        n, c = int(n), int(c)
        ni, no = (n // 2, n // 2 + 1) if real else (n, n)
        items.append(BatchItem(n, c, ioff, ooff, ni, no, scale, real))
        ioff += ni * c
        ooff += no * c
    return items


class BatchFFT:
    """A compiled batch of uneven 1-D complex64 transforms on one device."""

    def __init__(self, device: Device | VirtualDevice, items: Sequence[BatchItem], options: BatchOptions | None = None):
        self.device = device
        self.items = list(items)
        self.options = options or BatchOptions()
        self.virtual = isinstance(device, VirtualDevice)
        arr = (_lib.BatchItem * max(1, len(self.items)))()
        for k, it in enumerate(self.items):
            arr[k].length = int(it.length)
            arr[k].count = int(it.count)
            arr[k].in_offset = int(it.in_offset)
            arr[k].out_offset = int(it.out_offset)
            arr[k].in_dist = int(it.in_dist)
            arr[k].out_dist = int(it.out_dist)
            arr[k].flags = (_lib.RK_BATCH_SCALE if it.scale else 0) | (_lib.RK_BATCH_R2C if it.r2c else 0)
        if not self.virtual:
            device.activate()
        opts = self.options.to_c(dry=self.virtual)
        h = ctypes.c_void_p()
        _lib.call("rk_batch_plan_create_ex", arr, len(self.items), ctypes.byref(opts), ctypes.byref(h))
        self._h = h.value
        sz = ctypes.c_size_t(0)
        _lib.call("rk_batch_plan_workspace_size", self._h, ctypes.byref(sz))
        self.workspace_bytes = int(sz.value)
        self._ws: Buffer | None = None
        # elements each side must hold
        self.in_elements = max(
            (it.in_offset + (it.count - 1) * it.in_dist + it.in_length for it in self.items if it.count), default=0
        )
        self.out_elements = max(
            (it.out_offset + (it.count - 1) * it.out_dist + it.out_length for it in self.items if it.count), default=0
        )

    @classmethod
    def packed(
        cls,
        device: Device | VirtualDevice,
        lengths: Sequence[int],
        counts: Sequence[int] | None = None,
        scale: bool = False,
        options: BatchOptions | None = None,
    ) -> BatchFFT:
        return cls(device, packed_items(lengths, counts, scale), options)

    # -- introspection ----------------------------------------------------
    @property
    def info(self) -> dict:
        i = _lib.BatchInfo()
        _lib.call("rk_batch_plan_info", self._h, ctypes.byref(i))
        return {f: int(getattr(i, f)) for f, _ in _lib.BatchInfo._fields_}

    def check(self) -> None:
        """Raises RuntimeError if a dependency wait of an earlier call gave up
        (synchronises the device)."""
        hit = ctypes.c_int(0)
        _lib.call("rk_batch_plan_check", self._h, ctypes.byref(hit))
        if hit.value:
            raise RuntimeError("batch kernel: a dependency wait hit its poll limit; results are not valid")

    # -- execution ----------------------------------------------------------
    def _workspace(self) -> int:
        if not self.workspace_bytes:
            return 0
        if self._ws is None:
            self._ws = Buffer(self.device, self.workspace_bytes)
        return self._ws.ptr

    def __call__(self, queue: Queue, output: Array, input: Array, inverse: int = 0) -> None:
        """Runs the batch.  ``output`` is complex64; ``input`` is complex64, or
        float32 (two values per complex64 element of the item offsets: the
        natural buffer of a batch of real-to-complex items)."""
        if self.virtual:
            raise ValueError("a batch planned for a VirtualDevice cannot execute")
        if output.dtype.itemsize != 8 or output.dtype.kind != "c":
            raise ValueError("output must be complex64")
        if output.size < self.out_elements:
            raise ValueError(f"output holds {output.size} elements, the batch needs {self.out_elements}")
        real_in = input.dtype.kind == "f" and input.dtype.itemsize == 4
        if not real_in and (input.dtype.itemsize != 8 or input.dtype.kind != "c"):
            raise ValueError("input must be complex64 or float32")
        have = input.size // 2 if real_in else input.size
        if have < self.in_elements:
            raise ValueError(f"input holds {have} complex64 elements' worth, the batch needs {self.in_elements}")
        # the kernel walks both buffers as dense runs from their data pointers
        require_dense("output", output, 8 * self.out_elements)
        require_dense("input", input, 8 * self.in_elements)
        _lib.call(
            "rk_batch_execute", self._h, ctypes.c_void_p(output.data_ptr), ctypes.c_void_p(input.data_ptr),
            int(bool(inverse)), ctypes.c_void_p(self._workspace() or None), queue.stream,
        )

    def __del__(self) -> None:
        h = getattr(self, "_h", None)
        if h:
            try:
                _lib.load().rk_batch_plan_destroy(h)
            except Exception:  # pragma: no cover - interpreter shutdown
                pass
            self._h = None


def _is_prime(n: int) -> bool:
    if n < 2:
        return False
    if n % 2 == 0:
        return n == 2
    f = 3
    while f * f <= n:
        if n % f == 0:
            return False
        f += 2
    return True


def sample_lengths(count: int, lo: int = 64, hi: int = 65536, seed: int = 5) -> list[int]:
    """Transform lengths log-uniform in [lo, hi], each snapped to one of four
    families drawn with equal probability: 2^k, 3*2^k, 5*2^k (nearest in log
    scale) or the next prime (the mixed-length batch workload)."""
    import numpy

    rng = numpy.random.default_rng(seed)
    out: list[int] = []
    for _ in range(count):
        x = float(numpy.exp(rng.uniform(numpy.log(lo), numpy.log(hi))))
        fam = int(rng.integers(4))
        if fam < 3:
            base = (1, 3, 5)[fam]
            k = max(0, round(numpy.log2(x / base)))
            n = base * 2**k
            while n > hi and k > 0:
                k -= 1
                n = base * 2**k
            while n < lo:
                k += 1
                n = base * 2**k
        else:
            n = max(lo, int(x))
            while not _is_prime(n):
                n += 1
            if n > hi:
                n = int(x)
                while not _is_prime(n):
                    n -= 1
        out.append(int(n))
    return out
