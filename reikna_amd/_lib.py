"""ctypes binding of libreikna_amd.so (the C-ABI declared in include/reikna_amd.h).

Loading is lazy: importing the package never touches the GPU.  Every wrapped
call checks the returned status and raises: RK_ERR_INVALID becomes
ValueError (reikna raises ValueError for bad requests, e.g.
reikna/fft/fft.py:633-634), everything else RuntimeError.  There is no CPU
fallback: without the library, or without a GPU, the first device call fails.
"""

from __future__ import annotations

import ctypes
import os
import threading
from pathlib import Path

_LIB_PATH = Path(__file__).resolve().parent / "libreikna_amd.so"

RK_OK = 0
RK_ERR_INVALID = -1
RK_ERR_HIP = -2
RK_ERR_NO_KERNEL = -3
RK_ERR_UNSUPPORTED = -4
RK_ERR_BUDGET = -5

RK_MAX_DIMS = 8
RK_SCRATCH_BYTES = 64 << 20   # rk_fft_execute_scratch: the library's device scratch array
RK_FUSED_BLOB_BYTES = 512

RK_FFT_FORCE_BLUESTEIN = 0x1
RK_FFT_FORCE_MULTIPASS = 0x2
RK_FFT_NO_JIT = 0x4
RK_FFT_NORM_FORWARD = 0x8
RK_FFT_OUT_LAST_ONLY = 0x10
RK_FFT_DRY_RUN = 0x20
RK_C64 = 0
RK_C128 = 1

class ServerConfig(ctypes.Structure):
    _fields_ = [
        ("budget_bytes", ctypes.c_size_t),
        ("slots", ctypes.c_int32),
        ("max_requests", ctypes.c_int32),
        ("mode", ctypes.c_uint32),
        ("plain_min_elements", ctypes.c_int64),
        ("spin_limit", ctypes.c_uint32),
        ("fault_inject", ctypes.c_uint32),
    ]


class Request(ctypes.Structure):
    _fields_ = [
        ("id", ctypes.c_uint64),
        ("ndim", ctypes.c_int32),
        ("inverse", ctypes.c_int32),
        ("shape", ctypes.c_int64 * 2),
        ("count", ctypes.c_int64),
        ("input", ctypes.c_void_p),
        ("t_arrival_ns", ctypes.c_uint64),
    ]


class Completion(ctypes.Structure):
    _fields_ = [
        ("id", ctypes.c_uint64),
        ("out", ctypes.c_void_p),
        ("t_arrival_ns", ctypes.c_uint64),
        ("t_admitted_ns", ctypes.c_uint64),
        ("t_done_ns", ctypes.c_uint64),
        ("status", ctypes.c_int32),
        ("_reserved", ctypes.c_int32),
    ]


class ServerStats(ctypes.Structure):
    _fields_ = [
        ("pending", ctypes.c_int64),
        ("in_flight_requests", ctypes.c_int64),
        ("in_flight_launches", ctypes.c_int64),
        ("completed", ctypes.c_int64),
        ("launches", ctypes.c_int64),
        ("templates", ctypes.c_int64),
        ("budget_used", ctypes.c_size_t),
        ("budget_bytes", ctypes.c_size_t),
        ("failed_launches", ctypes.c_int64),
    ]


RK_SERVER_BATCHED = 0
RK_SERVER_PLAIN = 1
RK_SERVER_FAULT_SLOT_ALLOC = 1


class RkError(RuntimeError):
    """A failed call into libreikna_amd (HIP error, missing kernel, ...)."""

    def __init__(self, code: int, message: str):
        super().__init__(message)
        self.code = code


class BudgetError(RkError):
    """A request exceeded an HBM budget."""


class DeviceInfo(ctypes.Structure):
    _fields_ = [
        ("name", ctypes.c_char * 256),
        ("arch", ctypes.c_char * 64),
        ("compute_units", ctypes.c_int),
        ("wavefront_size", ctypes.c_int),
        ("max_threads_per_block", ctypes.c_int),
        ("lds_per_block", ctypes.c_size_t),
        ("total_global_mem", ctypes.c_size_t),
        ("l2_cache_bytes", ctypes.c_int),
        ("clock_khz", ctypes.c_int),
        ("memory_clock_khz", ctypes.c_int),
        ("memory_bus_width", ctypes.c_int),
    ]


class FftDesc(ctypes.Structure):
    _fields_ = [
        ("dtype", ctypes.c_int),
        ("ndim", ctypes.c_int),
        ("shape", ctypes.c_int64 * RK_MAX_DIMS),
        ("in_strides", ctypes.c_int64 * RK_MAX_DIMS),
        ("out_strides", ctypes.c_int64 * RK_MAX_DIMS),
        ("in_offset", ctypes.c_int64),
        ("out_offset", ctypes.c_int64),
        ("naxes", ctypes.c_int),
        ("axes", ctypes.c_int * RK_MAX_DIMS),
        ("flags", ctypes.c_uint32),
    ]


class FftTuning(ctypes.Structure):
    """rk_fft_tuning: planner knobs of one plan (0 / -1: the planner's choice)."""

    _fields_ = [
        ("factor_max", ctypes.c_int32),
        ("first_factor", ctypes.c_int32),
        ("col_tile", ctypes.c_int32),
        ("row_tile", ctypes.c_int32),
        ("chunk_bytes", ctypes.c_int64),
        ("fixed_modes", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class FftPassInfo(ctypes.Structure):
    _fields_ = [
        ("L", ctypes.c_int),
        ("C", ctypes.c_int),
        ("threads", ctypes.c_int),
        ("kind", ctypes.c_int),
        ("load_mode", ctypes.c_int),
        ("store_mode", ctypes.c_int),
        ("src", ctypes.c_int),
        ("dst", ctypes.c_int),
        ("axis", ctypes.c_int),
        ("pre_ops", ctypes.c_uint32),
        ("post_ops", ctypes.c_uint32),
        ("n_inst", ctypes.c_uint64),
        ("grid", ctypes.c_uint64),
        ("axis_len", ctypes.c_uint64),
        ("Ns", ctypes.c_uint64),
        ("lds_bytes", ctypes.c_int),
        ("jit", ctypes.c_int),
        ("variant", ctypes.c_int),
        ("in_dims", ctypes.c_int64 * 3),
        ("stream", ctypes.c_int),
        ("fixed_modes", ctypes.c_int),
        ("nt_load", ctypes.c_int),
        ("nt_store", ctypes.c_int),
        ("wave", ctypes.c_int),
        ("pipe", ctypes.c_int),
        ("occ", ctypes.c_int),
    ]


class BatchItem(ctypes.Structure):
    _fields_ = [
        ("length", ctypes.c_int64),
        ("count", ctypes.c_int64),
        ("in_offset", ctypes.c_int64),
        ("out_offset", ctypes.c_int64),
        ("in_dist", ctypes.c_int64),
        ("out_dist", ctypes.c_int64),
        ("flags", ctypes.c_uint32),
    ]


class BatchOptions(ctypes.Structure):
    _fields_ = [
        ("mode", ctypes.c_uint32),
        ("flags", ctypes.c_uint32),
        ("single_max", ctypes.c_int32),
        ("factor_max", ctypes.c_int32),
        ("blue_max", ctypes.c_int32),
        ("order", ctypes.c_int32),
    ]


class BatchInfo(ctypes.Structure):
    _fields_ = [
        ("items", ctypes.c_int64),
        ("passes", ctypes.c_int64),
        ("tiles", ctypes.c_int64),
        ("workspace_bytes", ctypes.c_int64),
        ("levels", ctypes.c_int32),
        ("launches", ctypes.c_int32),
        ("mode", ctypes.c_int32),
        ("lds_bytes", ctypes.c_int32),
        ("flavour", ctypes.c_int32),
        ("_reserved", ctypes.c_int32),
    ]


RK_BATCH_SCALE = 0x1
RK_BATCH_R2C = 0x2
RK_BATCH_MODE_QUEUE = 0
RK_BATCH_MODE_GROUPED = 1
RK_BATCH_DRY_RUN = 0x1

_P = ctypes.c_void_p
_PP = ctypes.POINTER(ctypes.c_void_p)
_I = ctypes.c_int
_SZ = ctypes.c_size_t

# name -> (restype, argtypes)
SIGNATURES: dict[str, tuple[object, list[object]]] = {
    "rk_last_error": (ctypes.c_char_p, []),
    "rk_version": (ctypes.c_char_p, []),
    "rk_device_count": (_I, [ctypes.POINTER(_I)]),
    "rk_set_device": (_I, [_I]),
    "rk_get_device_info": (_I, [_I, ctypes.POINTER(DeviceInfo)]),
    "rk_malloc": (_I, [_PP, _SZ]),
    "rk_free": (_I, [_P]),
    "rk_memcpy_htod": (_I, [_P, _P, _SZ, _P]),
    "rk_memcpy_dtoh": (_I, [_P, _P, _SZ, _P]),
    "rk_memcpy_dtod": (_I, [_P, _P, _SZ, _P]),
    "rk_memset": (_I, [_P, _I, _SZ, _P]),
    "rk_mem_info": (_I, [ctypes.POINTER(_SZ), ctypes.POINTER(_SZ)]),
    "rk_stream_create": (_I, [_PP]),
    "rk_stream_destroy": (_I, [_P]),
    "rk_stream_synchronize": (_I, [_P]),
    "rk_device_synchronize": (_I, []),
    "rk_event_create": (_I, [_PP]),
    "rk_event_destroy": (_I, [_P]),
    "rk_event_record": (_I, [_P, _P]),
    "rk_event_synchronize": (_I, [_P]),
    "rk_event_query": (_I, [_P, ctypes.POINTER(_I)]),
    "rk_event_elapsed_ms": (_I, [_P, _P, ctypes.POINTER(ctypes.c_float)]),
    "rk_stream_wait_event": (_I, [_P, _P]),
    "rk_module_compile": (_I, [ctypes.c_char_p, ctypes.c_char_p, _PP, ctypes.c_char_p, _SZ]),
    "rk_module_check": (_I, [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_char_p, _SZ]),
    "rk_module_get_function": (_I, [_P, ctypes.c_char_p, _PP]),
    "rk_module_unload": (_I, [_P]),
    "rk_launch": (
        _I,
        [_P, ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32), ctypes.c_uint32, _P, _P, _SZ],
    ),
    "rk_fft_plan_create": (_I, [ctypes.POINTER(FftDesc), _PP]),
    "rk_fft_plan_create_tuned": (_I, [ctypes.POINTER(FftDesc), ctypes.POINTER(FftTuning), _PP]),
    "rk_fft_plan_workspace_size": (_I, [_P, ctypes.POINTER(_SZ)]),
    "rk_fft_execute": (_I, [_P, _P, _P, _I, _P, _P]),
    "rk_fft_execute_checked": (_I, [_P, _P, _SZ, _P, _SZ, _I, _P, _SZ, _P]),
    "rk_fft_execute_scratch": (_I, [_P, _P, _P, _I, _P]),
    "rk_fft_execute_fused": (_I, [_P, _P, _P, _I, _P, _P, _P, _SZ, _P, _SZ]),
    "rk_fft_plan_destroy": (_I, [_P]),
    "rk_fft_plan_num_passes": (_I, [_P, ctypes.POINTER(_I)]),
    "rk_fft_plan_chunking": (_I, [_P, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]),
    "rk_fft_plan_pass_info": (_I, [_P, _I, ctypes.POINTER(FftPassInfo)]),
    "rk_fft_plan_set_pass_kernel": (_I, [_P, _I, _P]),
    "rk_fft_plan_set_timing": (_I, [_P, _I]),
    "rk_fft_plan_pass_time_ms": (_I, [_P, _I, ctypes.POINTER(ctypes.c_float)]),
    "rk_batch_plan_create": (_I, [ctypes.POINTER(BatchItem), _I, _PP]),
    "rk_batch_plan_workspace_size": (_I, [_P, ctypes.POINTER(_SZ)]),
    "rk_batch_execute": (_I, [_P, _P, _P, _I, _P, _P]),
    "rk_batch_plan_destroy": (_I, [_P]),
    "rk_batch_plan_stats": (_I, [_P, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(_I)]),
    "rk_batch_plan_create_ex": (_I, [ctypes.POINTER(BatchItem), _I, ctypes.POINTER(BatchOptions), _PP]),
    "rk_batch_plan_info": (_I, [_P, ctypes.POINTER(BatchInfo)]),
    "rk_batch_plan_check": (_I, [_P, ctypes.POINTER(_I)]),
    "rk_server_create": (_I, [ctypes.POINTER(ServerConfig), _PP]),
    "rk_server_submit": (_I, [_P, ctypes.POINTER(Request), _I]),
    "rk_server_progress": (_I, [_P, ctypes.POINTER(Completion), _I, ctypes.POINTER(_I)]),
    "rk_server_release": (_I, [_P, ctypes.c_uint64]),
    "rk_server_stats": (_I, [_P, ctypes.POINTER(ServerStats)]),
    "rk_server_now_ns": (ctypes.c_uint64, []),
    "rk_server_destroy": (_I, [_P]),
    "rk_server_replay": (_I, [_P, ctypes.POINTER(Request), _I, ctypes.POINTER(Completion), _I, _I]),
    "rk_graph_begin": (_I, [_P]),
    "rk_graph_end": (_I, [_P, _PP]),
    "rk_graph_launch": (_I, [_P, _P]),
    "rk_graph_destroy": (_I, [_P]),
}

_lock = threading.Lock()
_lib: ctypes.CDLL | None = None


def lib_path() -> Path:
    return Path(os.environ.get("REIKNA_AMD_LIB", str(_LIB_PATH)))


def load() -> ctypes.CDLL:
    """Load (once) and return the native library; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is None:
            path = lib_path()
            if not path.exists():
                raise RkError(
                    RK_ERR_NO_KERNEL,
                    f"{path} is not built: run `python -m reikna_amd.csrc.build` (or __graft_entry__.build())",
                )
            lib = ctypes.CDLL(str(path))
            for name, (restype, argtypes) in SIGNATURES.items():
                fn = getattr(lib, name)
                fn.restype = restype
                fn.argtypes = argtypes
            _lib = lib
    return _lib


def last_error() -> str:
    msg = load().rk_last_error()
    return msg.decode(errors="replace") if msg else ""


def check(code: int, what: str = "") -> None:
    if code == RK_OK:
        return
    msg = last_error()
    text = f"{what}: {msg}" if what else msg
    if code == RK_ERR_INVALID:
        raise ValueError(text)
    if code == RK_ERR_BUDGET:
        raise BudgetError(code, text)
    raise RkError(code, text)


def call(name: str, *args: object) -> None:
    check(getattr(load(), name)(*args), name)
