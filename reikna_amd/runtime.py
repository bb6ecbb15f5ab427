"""Device runtime: array metadata, devices, queues and device arrays.

This is the part of grunnur (reikna's external backend) that the FFT path
uses: ``ArrayMetadata`` / ``AsArrayMetadata`` (reikna/core/signature.py:37-111),
``BoundDevice`` with ``DeviceParameters`` (reikna/fft/fft.py:239-250, 295-335),
``Queue`` and ``Array`` (``Array.from_host``, ``Array.empty_like``, ``get``:
tests/test_fft/test_fft.py:156-170).  It sits on the C-ABI of libreikna_amd
(device memory, HIP streams, events), so it needs neither PyTorch nor
PyOpenCL; torch tensors can still be wrapped without a copy
(``Array.from_torch``).
"""

from __future__ import annotations

import ctypes
import math
import weakref
from collections.abc import Sequence
from typing import Any

import numpy

from . import _lib


# ---------------------------------------------------------------------------
# array metadata
# ---------------------------------------------------------------------------
def _c_strides(shape: tuple[int, ...], itemsize: int) -> tuple[int, ...]:
    strides = []
    s = itemsize
    for n in reversed(shape):
        strides.append(s)
        s *= max(n, 1)
    return tuple(reversed(strides))


class AsArrayMetadata:
    """Anything that can describe an array (shape, dtype, strides, offset)."""

    def as_array_metadata(self) -> ArrayMetadata:  # pragma: no cover - interface
        raise NotImplementedError


class ArrayMetadata(AsArrayMetadata):
    """
    Shape, dtype, byte strides, byte offset of the first element and buffer size
    of an array (the grunnur ``ArrayMetadata`` reikna's ``Type`` wraps).
    """

    def __init__(
        self,
        shape: int | Sequence[int],
        dtype: Any,
        strides: Sequence[int] | None = None,
        first_element_offset: int = 0,
        buffer_size: int | None = None,
    ):
        if isinstance(shape, (int, numpy.integer)):
            shape = (int(shape),)
        self.shape = tuple(int(n) for n in shape)
        self.dtype = numpy.dtype(dtype)
        self.strides = (
            _c_strides(self.shape, self.dtype.itemsize) if strides is None else tuple(int(s) for s in strides)
        )
        if len(self.strides) != len(self.shape):
            raise ValueError("strides and shape must have the same length")
        self.first_element_offset = int(first_element_offset)
        if buffer_size is None:
            if any(n == 0 for n in self.shape):
                buffer_size = self.first_element_offset
            else:
                last = sum((n - 1) * s for n, s in zip(self.shape, self.strides) if s > 0)
                buffer_size = self.first_element_offset + last + self.dtype.itemsize
        self.buffer_size = int(buffer_size)

    @classmethod
    def from_arraylike(cls, arr: Any) -> ArrayMetadata:
        if isinstance(arr, ArrayMetadata):
            return arr
        if isinstance(arr, AsArrayMetadata):
            return arr.as_array_metadata()
        if isinstance(arr, numpy.ndarray):
            return cls(arr.shape, arr.dtype, strides=arr.strides)
        return cls(
            arr.shape,
            arr.dtype,
            strides=getattr(arr, "strides", None),
            first_element_offset=getattr(arr, "offset", 0),
        )

    def as_array_metadata(self) -> ArrayMetadata:
        return self

    @property
    def size(self) -> int:
        return math.prod(self.shape)

    @property
    def nbytes(self) -> int:
        return self.size * self.dtype.itemsize

    @property
    def ndim(self) -> int:
        return len(self.shape)

    def is_contiguous(self) -> bool:
        return self.strides == _c_strides(self.shape, self.dtype.itemsize)

    def with_(self, **kwds: Any) -> ArrayMetadata:
        d = dict(
            shape=self.shape,
            dtype=self.dtype,
            strides=self.strides,
            first_element_offset=self.first_element_offset,
            buffer_size=self.buffer_size,
        )
        if "dtype" in kwds and "strides" not in kwds:
            # strides in bytes follow the new item size when the array was contiguous
            if self.is_contiguous():
                d["strides"] = None
                d["buffer_size"] = None
        if "shape" in kwds and "strides" not in kwds:
            d["strides"] = None
            d["buffer_size"] = None
        d.update(kwds)
        return ArrayMetadata(**d)

    def _key(self) -> tuple:
        return (self.shape, self.dtype, self.strides, self.first_element_offset, self.buffer_size)

    def __eq__(self, other: object) -> bool:
        return isinstance(other, ArrayMetadata) and self._key() == other._key()

    def __hash__(self) -> int:
        return hash(self._key())

    def __repr__(self) -> str:
        extra = "" if self.is_contiguous() else f", strides={self.strides}"
        if self.first_element_offset:
            extra += f", offset={self.first_element_offset}"
        return f"ArrayMetadata({self.shape}, {self.dtype}{extra})"


# ---------------------------------------------------------------------------
# devices
# ---------------------------------------------------------------------------
class DeviceParameters:
    """
    Device limits a planner needs (grunnur's DeviceParameters: the reference
    reads max_total_local_size, local_mem_size, local_mem_banks and
    align_words in reikna/fft/fft.py:239-250, 295-335).
    """

    def __init__(
        self,
        max_total_local_size: int = 1024,
        local_mem_size: int = 160 * 1024,
        local_mem_banks: int = 64,
        warp_size: int = 64,
        compute_units: int = 256,
        global_mem_size: int = 288 * 2**30,
        l2_cache_size: int = 4 * 2**20,
        arch: str = "gfx950",
        name: str = "AMD Instinct MI355X",
    ):
        self.max_total_local_size = max_total_local_size
        self.max_local_sizes = (max_total_local_size, max_total_local_size, max_total_local_size)
        self.local_mem_size = local_mem_size
        self.local_mem_banks = local_mem_banks
        self.warp_size = warp_size
        self.compute_units = compute_units
        self.global_mem_size = global_mem_size
        self.l2_cache_size = l2_cache_size
        self.arch = arch
        self.name = name
        self.type = "gpu"

    def align_words(self, word_size: int) -> int:
        # coalescing width in words: a 64-lane wave over 128-byte lines
        return max(1, 128 // word_size)

    def __repr__(self) -> str:
        return f"DeviceParameters({self.name}, {self.arch}, CUs={self.compute_units})"


def device_count() -> int:
    n = ctypes.c_int(0)
    _lib.call("rk_device_count", ctypes.byref(n))
    return n.value


class Device:
    """A GPU the computations are compiled for (grunnur's BoundDevice)."""

    _cache: dict[int, Device] = {}

    def __new__(cls, index: int = 0) -> Device:
        if index in cls._cache:
            return cls._cache[index]
        self = super().__new__(cls)
        self._init(index)
        cls._cache[index] = self
        return self

    def _init(self, index: int) -> None:
        n = device_count()
        if n == 0:
            raise _lib.RkError(_lib.RK_ERR_HIP, "no GPU device is visible")
        if not 0 <= index < n:
            raise ValueError(f"device index {index} out of range ({n} devices)")
        self.index = index
        info = _lib.DeviceInfo()
        _lib.call("rk_get_device_info", index, ctypes.byref(info))
        self.name = info.name.decode()
        self.arch = info.arch.decode().split(":")[0]
        self.params = DeviceParameters(
            max_total_local_size=info.max_threads_per_block,
            local_mem_size=int(info.lds_per_block),
            warp_size=info.wavefront_size,
            compute_units=info.compute_units,
            global_mem_size=int(info.total_global_mem),
            l2_cache_size=int(info.l2_cache_bytes),
            arch=self.arch,
            name=self.name,
        )
        self.memory_clock_khz = info.memory_clock_khz
        self.memory_bus_width = info.memory_bus_width

    @property
    def device(self) -> Device:
        return self

    def activate(self) -> None:
        _lib.call("rk_set_device", self.index)

    def synchronize(self) -> None:
        self.activate()
        _lib.call("rk_device_synchronize")

    def mem_info(self) -> tuple[int, int]:
        self.activate()
        free, total = ctypes.c_size_t(0), ctypes.c_size_t(0)
        _lib.call("rk_mem_info", ctypes.byref(free), ctypes.byref(total))
        return free.value, total.value

    def __repr__(self) -> str:
        return f"Device({self.index}: {self.name}, {self.arch})"


class VirtualDevice:
    """
    A stand-in device for planning without a GPU: ``compile()`` against it runs
    the planners (native plans in dry-run mode) and compiles every generated
    kernel for gfx950 with hiprtc, but allocates nothing and cannot execute.
    Used by the CPU test-suite to cover code generation.
    """

    virtual = True

    def __init__(self, params: DeviceParameters | None = None):
        self.index = -1
        self.params = params or DeviceParameters()
        self.name = self.params.name
        self.arch = self.params.arch

    @property
    def device(self) -> VirtualDevice:
        return self

    def activate(self) -> None:
        pass

    def __repr__(self) -> str:
        return f"VirtualDevice({self.arch})"


def default_device() -> Device:
    return Device(0)


def _as_device(obj: Any) -> Device:
    if obj is None:
        return default_device()
    if isinstance(obj, Device):
        return obj
    if isinstance(obj, Queue):
        return obj.device
    if hasattr(obj, "device") and isinstance(obj.device, Device):
        return obj.device
    raise TypeError(f"expected a Device or Queue, got {type(obj)}")


class Queue:
    """A HIP stream on one device (grunnur's Queue)."""

    def __init__(self, device: Device | None = None, *, stream: int | None = None):
        self.device = _as_device(device)
        self.device.activate()
        if stream is None:
            h = ctypes.c_void_p()
            _lib.call("rk_stream_create", ctypes.byref(h))
            self._stream = h.value
            self._owned = True
            weakref.finalize(self, _destroy_stream, self._stream)
        else:
            self._stream = stream
            self._owned = False

    @property
    def stream(self) -> int | None:
        return self._stream

    def synchronize(self) -> None:
        _lib.call("rk_stream_synchronize", self._stream)

    def capture(self) -> _Capture:
        """Record the work enqueued inside a ``with`` block into a HIP graph
        instead of running it::

            with queue.capture() as cap:
                fft(queue, out, inp)
            cap.graph.launch(queue)    # replays the recorded launches
        """
        return _Capture(self)

    def __repr__(self) -> str:
        return f"Queue({self.device}, stream=0x{(self._stream or 0):x})"


class Graph:
    """An instantiated HIP graph (the launches recorded by Queue.capture)."""

    def __init__(self, handle: int):
        self._h = handle
        weakref.finalize(self, _destroy_graph, handle)

    def launch(self, queue: Queue) -> None:
        _lib.call("rk_graph_launch", self._h, queue.stream)


def _destroy_graph(h: int) -> None:
    try:
        _lib.load().rk_graph_destroy(h)
    except Exception:  # pragma: no cover - interpreter shutdown
        pass


class _Capture:
    def __init__(self, queue: Queue):
        self.queue = queue
        self.graph: Graph | None = None

    def __enter__(self) -> _Capture:
        if self.queue.stream is None:
            raise ValueError("the default stream cannot be captured; use a created Queue")
        _lib.call("rk_graph_begin", self.queue.stream)
        return self

    def __exit__(self, exc_type: Any, exc: Any, tb: Any) -> None:
        h = ctypes.c_void_p()
        status = _lib.load().rk_graph_end(self.queue.stream, ctypes.byref(h))
        if exc_type is None:
            _lib.check(status, "rk_graph_end")
            self.graph = Graph(h.value)
        elif h.value:
            _destroy_graph(h.value)


def _destroy_stream(stream: int) -> None:
    try:
        _lib.load().rk_stream_destroy(stream)
    except Exception:  # pragma: no cover - interpreter shutdown
        pass


class Event:
    """A HIP event for device-side timing."""

    def __init__(self) -> None:
        h = ctypes.c_void_p()
        _lib.call("rk_event_create", ctypes.byref(h))
        self._h = h.value
        weakref.finalize(self, _destroy_event, self._h)

    def record(self, queue: Queue) -> Event:
        _lib.call("rk_event_record", self._h, queue.stream)
        return self

    def synchronize(self) -> None:
        _lib.call("rk_event_synchronize", self._h)

    def query(self) -> bool:
        done = ctypes.c_int(0)
        _lib.call("rk_event_query", self._h, ctypes.byref(done))
        return bool(done.value)

    def elapsed_ms(self, end: Event) -> float:
        ms = ctypes.c_float(0)
        _lib.call("rk_event_elapsed_ms", self._h, end._h, ctypes.byref(ms))
        return float(ms.value)


def _destroy_event(h: int) -> None:
    try:
        _lib.load().rk_event_destroy(h)
    except Exception:  # pragma: no cover
        pass


# ---------------------------------------------------------------------------
# device memory
# ---------------------------------------------------------------------------
class Buffer:
    """An owned device allocation."""

    def __init__(self, device: Device, size: int):
        self.device = device
        self.size = int(size)
        device.activate()
        p = ctypes.c_void_p()
        _lib.call("rk_malloc", ctypes.byref(p), ctypes.c_size_t(max(self.size, 1)))
        self.ptr = int(p.value or 0)
        weakref.finalize(self, _free_ptr, self.ptr)

    @classmethod
    def allocate(cls, device: Device, size: int) -> Buffer:
        return cls(device, size)


def _free_ptr(ptr: int) -> None:
    try:
        _lib.load().rk_free(ptr)
    except Exception:  # pragma: no cover
        pass


class _ForeignBuffer:
    """Device memory owned by someone else (a torch tensor); kept alive by `owner`."""

    def __init__(self, device: Device, ptr: int, size: int, owner: Any):
        self.device = device
        self.ptr = int(ptr)
        self.size = int(size)
        self.owner = owner


def require_dense(name: str, arr: Any, need_bytes: int) -> None:
    """Native entry points that take a bare data pointer (the batch kernel,
    the request server) read and write `need_bytes` consecutive bytes from it:
    the array must be C-contiguous over its dimensions longer than one, and
    its buffer must hold that many bytes from the first element on.  Strided
    or broadcast views are refused here, before any launch."""
    if not isinstance(arr, Array):
        raise TypeError(f"{name}: expected a device Array, got {type(arr)}")
    meta = arr.metadata
    dense = _c_strides(meta.shape, meta.dtype.itemsize)
    if any(n > 1 and s != d for n, s, d in zip(meta.shape, meta.strides, dense)):
        raise ValueError(f"{name}: array strides {meta.strides} are not C-contiguous for shape {meta.shape}")
    avail = getattr(arr.buffer, "size", None)
    if avail is None:
        avail = meta.buffer_size
    avail -= meta.first_element_offset
    if avail < need_bytes or meta.nbytes < need_bytes:
        raise ValueError(f"{name}: {min(avail, meta.nbytes)} bytes from the first element, {need_bytes} needed")


class Array(AsArrayMetadata):
    """A device array: a buffer plus ArrayMetadata (grunnur's Array)."""

    def __init__(self, device: Device, metadata: ArrayMetadata, buffer: Buffer | _ForeignBuffer | None = None):
        self.device = device
        self.metadata = metadata
        if buffer is None:
            buffer = Buffer(device, max(metadata.buffer_size, 1))
        self.buffer = buffer

    # -- construction ------------------------------------------------------
    @classmethod
    def empty(
        cls,
        device: Any,
        shape: int | Sequence[int],
        dtype: Any,
        strides: Sequence[int] | None = None,
        first_element_offset: int = 0,
        allocator: Any = None,
    ) -> Array:
        meta = ArrayMetadata(shape, dtype, strides=strides, first_element_offset=first_element_offset)
        dev = _as_device(device)
        buf = allocator(dev, max(meta.buffer_size, 1)) if allocator is not None else None
        return cls(dev, meta, buf)

    @classmethod
    def empty_like(cls, device: Any, arr: Any, allocator: Any = None) -> Array:
        meta = ArrayMetadata.from_arraylike(arr)
        dev = _as_device(device)
        buf = allocator(dev, max(meta.buffer_size, 1)) if allocator is not None else None
        return cls(dev, meta, buf)

    @classmethod
    def from_host(cls, queue_or_device: Any, arr: numpy.ndarray) -> Array:
        arr = numpy.asarray(arr)
        dev = _as_device(queue_or_device)
        queue = queue_or_device if isinstance(queue_or_device, Queue) else None
        if not (arr.flags.c_contiguous or arr.flags.f_contiguous) or any(s < 0 for s in arr.strides):
            arr = numpy.ascontiguousarray(arr)
        # keep the host layout (C or Fortran order) on the device
        meta = ArrayMetadata(arr.shape, arr.dtype, strides=arr.strides)
        out = cls(dev, meta)
        out.set(queue, arr)
        return out

    @classmethod
    def from_torch(cls, tensor: Any) -> Array:
        """Wrap a CUDA/ROCm torch tensor without copying (the tensor is kept alive)."""
        import torch

        if not tensor.is_cuda:
            raise ValueError("from_torch needs a device tensor")
        dtype = numpy.dtype(str(tensor.dtype).replace("torch.", ""))
        itemsize = dtype.itemsize
        base = tensor.untyped_storage().data_ptr()
        meta = ArrayMetadata(
            tuple(tensor.shape),
            dtype,
            strides=tuple(s * itemsize for s in tensor.stride()),
            first_element_offset=tensor.storage_offset() * itemsize,
            buffer_size=tensor.untyped_storage().nbytes(),
        )
        dev = Device(tensor.device.index or 0)
        del torch
        return cls(dev, meta, _ForeignBuffer(dev, base, meta.buffer_size, tensor))

    # -- metadata ----------------------------------------------------------
    def as_array_metadata(self) -> ArrayMetadata:
        return self.metadata

    @property
    def shape(self) -> tuple[int, ...]:
        return self.metadata.shape

    @property
    def dtype(self) -> numpy.dtype:
        return self.metadata.dtype

    @property
    def strides(self) -> tuple[int, ...]:
        return self.metadata.strides

    @property
    def offset(self) -> int:
        return self.metadata.first_element_offset

    @property
    def nbytes(self) -> int:
        return self.metadata.nbytes

    @property
    def size(self) -> int:
        return self.metadata.size

    @property
    def ptr(self) -> int:
        """Base pointer of the buffer (element 0 sits at ptr + offset)."""
        return self.buffer.ptr

    @property
    def data_ptr(self) -> int:
        return self.buffer.ptr + self.metadata.first_element_offset

    # -- transfers ---------------------------------------------------------
    def _span(self) -> tuple[int, int]:
        return 0, self.metadata.buffer_size

    def set(self, queue: Queue | None, arr: numpy.ndarray) -> None:
        arr = numpy.asarray(arr, dtype=self.dtype)
        if arr.shape != self.shape:
            raise ValueError(f"shape mismatch: {arr.shape} vs {self.shape}")
        stream = queue.stream if queue is not None else None
        self.device.activate()
        host = self._host_image(arr)
        _lib.call(
            "rk_memcpy_htod", ctypes.c_void_p(self.buffer.ptr), host.ctypes.data_as(ctypes.c_void_p),
            ctypes.c_size_t(host.nbytes), ctypes.c_void_p(stream),
        )

    def _host_image(self, arr: numpy.ndarray) -> numpy.ndarray:
        """The byte image of the whole buffer with `arr` placed per the metadata."""
        meta = self.metadata
        if meta.first_element_offset == 0 and meta.strides == arr.strides and meta.buffer_size == arr.nbytes:
            return numpy.ascontiguousarray(arr) if arr.flags.c_contiguous or arr.flags.f_contiguous else arr
        raw = numpy.zeros(meta.buffer_size, numpy.uint8)
        view = numpy.lib.stride_tricks.as_strided(
            raw[meta.first_element_offset :].view(self.dtype) if meta.buffer_size > meta.first_element_offset else raw.view(self.dtype),
            shape=meta.shape,
            strides=meta.strides,
        )
        view[...] = arr
        return raw

    def get(self, queue: Queue | None = None) -> numpy.ndarray:
        meta = self.metadata
        stream = queue.stream if queue is not None else None
        self.device.activate()
        raw = numpy.empty(max(meta.buffer_size, 1), numpy.uint8)
        _lib.call(
            "rk_memcpy_dtoh", raw.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(self.buffer.ptr),
            ctypes.c_size_t(meta.buffer_size), ctypes.c_void_p(stream),
        )
        if meta.size == 0:
            return numpy.empty(meta.shape, meta.dtype)
        view = numpy.lib.stride_tricks.as_strided(
            raw[meta.first_element_offset :].view(self.dtype), shape=meta.shape, strides=meta.strides
        )
        return numpy.array(view)

    def copy_from(self, queue: Queue | None, other: Array) -> None:
        if other.metadata.buffer_size != self.metadata.buffer_size:
            raise ValueError("buffer sizes differ")
        stream = queue.stream if queue is not None else None
        _lib.call(
            "rk_memcpy_dtod", ctypes.c_void_p(self.buffer.ptr), ctypes.c_void_p(other.buffer.ptr),
            ctypes.c_size_t(self.metadata.buffer_size), ctypes.c_void_p(stream),
        )

    def __repr__(self) -> str:
        return f"Array({self.metadata}, device={self.device.index})"
