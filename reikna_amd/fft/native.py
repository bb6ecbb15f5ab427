"""Python handle of a native FFT plan (rk_fft_plan_create / rk_fft_execute)."""

from __future__ import annotations

import ctypes
import weakref
from collections.abc import Sequence
from dataclasses import dataclass
from typing import Any

import numpy

from .. import _lib
from ..runtime import ArrayMetadata

LOAD_MODES = {0: "direct", 1: "column", 2: "flat", 3: "row"}
BUFFERS = {0: "in", 1: "out", 2: "ws0", 3: "ws1"}


@dataclass
class PassInfo:
    L: int
    C: int
    threads: int
    kind: int
    load_mode: str
    store_mode: str
    src: str
    dst: str
    axis: int
    pre_ops: int
    post_ops: int
    n_inst: int
    grid: int
    axis_len: int
    Ns: int
    lds_bytes: int
    jit: bool
    variant: int
    stream: bool = False   # nontemporal loads/stores (transform larger than the Infinity Cache)
    fixed_modes: bool = False   # the kernel has the pass's load/store modes compiled in
    nt_load: bool = False       # nontemporal loads
    nt_store: bool = False      # nontemporal stores
    wave: bool = False          # whole instances per wave, wave-local stage exchanges
    pipe: bool = False          # persistent kernel that loads its next tile while computing one
    occ: bool = False           # launch bounds for as many resident work-groups as the LDS allows


@dataclass(frozen=True)
class PlanTuning:
    """Planner knobs of one plan (rk_fft_tuning); the defaults leave every
    choice to the planner.  ``first_factor`` is the length of the first pass of
    every multi-pass axis, ``col_tile`` / ``row_tile`` the instances per tile of
    column- / row-mapped passes, ``chunk_bytes`` the chunk size of chained passes
    (0: no chunks), ``fixed_modes`` whether kernels have their load/store modes
    compiled in.  A knob the plan cannot honour is ignored."""

    factor_max: int = 0
    first_factor: int = 0
    col_tile: int = 0
    row_tile: int = 0
    chunk_bytes: int = -1
    fixed_modes: int = -1

    def to_c(self) -> _lib.FftTuning:
        t = _lib.FftTuning()
        t.factor_max = self.factor_max
        t.first_factor = self.first_factor
        t.col_tile = self.col_tile
        t.row_tile = self.row_tile
        t.chunk_bytes = self.chunk_bytes
        t.fixed_modes = self.fixed_modes
        t.reserved = 0
        return t


def _element_layout(meta: ArrayMetadata) -> tuple[list[int], int]:
    item = meta.dtype.itemsize
    if meta.first_element_offset % item or any(s % item for s in meta.strides):
        raise ValueError("strides and offset must be multiples of the item size")
    return [s // item for s in meta.strides], meta.first_element_offset // item


class NativeFftPlan:
    """A compiled pass chain for complex64 data of a given shape/layout/axes."""

    def __init__(
        self,
        in_meta: ArrayMetadata,
        out_meta: ArrayMetadata,
        axes: Sequence[int] | None,
        flags: int = 0,
        tuning: PlanTuning | None = None,
    ):
        if in_meta.shape != out_meta.shape:
            raise ValueError("input and output shapes differ")
        if in_meta.dtype != out_meta.dtype or in_meta.dtype not in (numpy.complex64, numpy.complex128):
            raise NotImplementedError(
                f"the gfx950 FFT kernels compute complex64 or complex128; got {in_meta.dtype} -> {out_meta.dtype}"
            )
        ndim = len(in_meta.shape)
        if ndim > _lib.RK_MAX_DIMS:
            raise ValueError(f"at most {_lib.RK_MAX_DIMS} dimensions are supported")
        desc = _lib.FftDesc()
        desc.dtype = _lib.RK_C128 if in_meta.dtype == numpy.complex128 else _lib.RK_C64
        self.dtype = in_meta.dtype
        desc.ndim = ndim
        in_str, in_off = _element_layout(in_meta)
        out_str, out_off = _element_layout(out_meta)
        for k in range(ndim):
            desc.shape[k] = in_meta.shape[k]
            desc.in_strides[k] = in_str[k]
            desc.out_strides[k] = out_str[k]
        desc.in_offset = in_off
        desc.out_offset = out_off
        axes = list(range(ndim)) if axes is None else list(axes)
        desc.naxes = len(axes)
        for k, a in enumerate(axes):
            desc.axes[k] = a
        desc.flags = flags
        self.shape = tuple(in_meta.shape)
        self.axes = tuple(axes)
        self.tuning = tuning
        h = ctypes.c_void_p()
        if tuning is None:
            _lib.call("rk_fft_plan_create", ctypes.byref(desc), ctypes.byref(h))
        else:
            tc = tuning.to_c()
            _lib.call("rk_fft_plan_create_tuned", ctypes.byref(desc), ctypes.byref(tc), ctypes.byref(h))
        self.handle = h.value
        self._finalizer = weakref.finalize(self, _destroy, self.handle)
        ws = ctypes.c_size_t(0)
        _lib.call("rk_fft_plan_workspace_size", ctypes.c_void_p(self.handle), ctypes.byref(ws))
        self.workspace_bytes = int(ws.value)

    @property
    def chunk_rows(self) -> int:
        """Outer-batch indices per chunk of a chunked plan (0: not chunked)."""
        c, n = ctypes.c_int64(0), ctypes.c_int64(0)
        _lib.call("rk_fft_plan_chunking", ctypes.c_void_p(self.handle), ctypes.byref(c), ctypes.byref(n))
        return int(c.value)

    @property
    def passes(self) -> list[PassInfo]:
        n = ctypes.c_int(0)
        _lib.call("rk_fft_plan_num_passes", ctypes.c_void_p(self.handle), ctypes.byref(n))
        out = []
        for p in range(n.value):
            info = _lib.FftPassInfo()
            _lib.call("rk_fft_plan_pass_info", ctypes.c_void_p(self.handle), p, ctypes.byref(info))
            out.append(
                PassInfo(
                    L=info.L, C=info.C, threads=info.threads, kind=info.kind,
                    load_mode=LOAD_MODES[info.load_mode], store_mode=LOAD_MODES[info.store_mode],
                    src=BUFFERS[info.src], dst=BUFFERS[info.dst], axis=info.axis, pre_ops=info.pre_ops,
                    post_ops=info.post_ops, n_inst=info.n_inst, grid=info.grid, axis_len=info.axis_len,
                    Ns=info.Ns, lds_bytes=info.lds_bytes, jit=bool(info.jit), variant=info.variant,
                    stream=bool(info.stream),
                    fixed_modes=bool(info.fixed_modes),
                    nt_load=bool(info.nt_load), nt_store=bool(info.nt_store), wave=bool(info.wave),
                    pipe=bool(info.pipe), occ=bool(info.occ),
                )
            )
        return out

    def execute(self, out_ptr: int, in_ptr: int, inverse: int, ws_ptr: int | None, stream: int | None,
                load_blob: bytes | None = None, store_blob: bytes | None = None) -> None:
        h = ctypes.c_void_p(self.handle)
        if load_blob is None and store_blob is None:
            rc = _lib.load().rk_fft_execute(h, ctypes.c_void_p(out_ptr), ctypes.c_void_p(in_ptr), int(inverse),
                                            ctypes.c_void_p(ws_ptr), ctypes.c_void_p(stream))
        else:
            lb = load_blob or b""
            sb = store_blob or b""
            rc = _lib.load().rk_fft_execute_fused(
                h, ctypes.c_void_p(out_ptr), ctypes.c_void_p(in_ptr), int(inverse), ctypes.c_void_p(ws_ptr),
                ctypes.c_void_p(stream), lb, len(lb), sb, len(sb),
            )
        _lib.check(rc, "rk_fft_execute")

    def execute_checked(self, out_ptr: int, out_bytes: int, in_ptr: int, in_bytes: int, inverse: int,
                        ws_ptr: int | None, ws_bytes: int, stream: int | None) -> None:
        """rk_fft_execute_checked: the buffers' sizes are checked against the
        extents the plan's layouts reach before anything is launched."""
        rc = _lib.load().rk_fft_execute_checked(
            ctypes.c_void_p(self.handle), ctypes.c_void_p(out_ptr), int(out_bytes), ctypes.c_void_p(in_ptr),
            int(in_bytes), int(inverse), ctypes.c_void_p(ws_ptr), int(ws_bytes), ctypes.c_void_p(stream),
        )
        _lib.check(rc, "rk_fft_execute_checked")

    def set_pass_kernel(self, pass_index: int, function: int) -> None:
        _lib.call("rk_fft_plan_set_pass_kernel", ctypes.c_void_p(self.handle), pass_index, ctypes.c_void_p(function))

    def set_timing(self, enable: bool) -> None:
        _lib.call("rk_fft_plan_set_timing", ctypes.c_void_p(self.handle), int(bool(enable)))

    def pass_times_ms(self) -> list[float]:
        out = []
        for p in range(len(self.passes)):
            ms = ctypes.c_float(-1)
            _lib.call("rk_fft_plan_pass_time_ms", ctypes.c_void_p(self.handle), p, ctypes.byref(ms))
            out.append(float(ms.value))
        return out


def _destroy(handle: Any) -> None:
    try:
        _lib.load().rk_fft_plan_destroy(ctypes.c_void_p(handle))
    except Exception:  # pragma: no cover - interpreter shutdown
        pass
