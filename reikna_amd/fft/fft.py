"""The FFT computation (reikna/fft/fft.py:595-756), on gfx950 kernels.

``FFT(arr_t, axes=None)`` keeps the reference's signature
``(output:o, input:i, inverse:int32=0)``: ``output`` and ``input`` may be the
same array, ``inverse`` is a run-time scalar, the inverse is normalised so
that IFFT(FFT(x)) = x, and a non-complex dtype raises ValueError with the
reference's message (:630-645).

Compiling builds a native plan (include/reikna_amd.h, rk_fft_plan_create):
radix factorisation, pass chain, buffer routing and device tables.  Without
attached transformations every pass runs an ahead-of-time gfx950 kernel.
With transformations on ``input`` and/or ``output`` (connected through
``Computation.connect``), the first pass's load and the last pass's store are
replaced by kernels generated at compile time that inline the transformation
chain (the reference splices them into fft.mako through
``load_combined_idx`` / ``store_combined_idx``, fft.mako:220-221, 871-872,
1000-1002), so fused transformations cost no extra pass over memory.
If every transformed axis has length 1, the plan is a copy through
PureParallel so that transformations still run (:647-662).
"""

from __future__ import annotations

import ctypes
import os
from collections.abc import Callable, Iterable
from typing import Any

import numpy

from .. import _lib, dtypes, helpers
from ..algorithms.pureparallel import PureParallel
from ..core import Annotation, Computation, ComputationPlan, KernelArguments, Parameter
from ..core.computation import PRELUDE, compile_kernel, pack_arguments
from ..core.transformation import leaf_name, param_cname
from ..runtime import Array, ArrayMetadata, DeviceParameters, Queue
from ..template import render_with_modules
from ..transformations import copy
from .native import NativeFftPlan, PlanTuning


def _meta(arr_t: Any) -> ArrayMetadata:
    if hasattr(arr_t, "as_array_metadata"):
        return arr_t.as_array_metadata()
    return ArrayMetadata.from_arraylike(arr_t)


# RK_MODE_* codes of fft_args.h by the pass-info names of native.LOAD_MODES
_MODE_CODES = {"direct": 0, "column": 1, "flat": 2, "row": 3}


def _fixed_modes() -> bool:
    """Whether kernels get their pass's load/store modes compiled in (the
    library's RK_FIXED_MODES switch: 0 turns it off)."""
    v = os.environ.get("RK_FIXED_MODES", "")
    return not (v.strip().lstrip("-").isdigit() and int(v) == 0)


def _ptr(val: Any) -> int:
    if isinstance(val, Array):
        return val.ptr
    if hasattr(val, "is_cuda") and hasattr(val, "data_ptr"):
        return Array.from_torch(val).ptr
    raise TypeError(f"expected a device Array, got {type(val)}")


class FFT(Computation):
    """
    Fast Fourier transform over ``axes`` of a complex array (numpy.fft.fftn
    conventions; the inverse is normalised).

    .. py:method:: compiled_signature(output:o, input:i, inverse:s)

        ``output`` and ``input`` may be the same array; ``inverse`` = 0 gives the
        forward transform, 1 the inverse one.
    """

    def __init__(self, arr_t: Any, axes: Iterable[int] | None = None, *, flags: int = 0,
                 tuning: PlanTuning | None = None):
        arr = _meta(arr_t)
        if not dtypes.is_complex(arr.dtype):
            raise ValueError("FFT computation requires array of a complex dtype")
        if arr.dtype not in (numpy.complex64, numpy.complex128):
            # the kernels compute complex64 (FP32 butterflies, FP64-derived
            # tables) and complex128 (FP64 throughout)
            raise NotImplementedError(f"FFT of {arr.dtype}: this implementation computes complex64 and complex128 transforms")
        Computation.__init__(
            self,
            [
                Parameter("output", Annotation(arr, "o")),
                Parameter("input", Annotation(arr, "i")),
                Parameter("inverse", Annotation(numpy.int32), default=0),
            ],
        )
        ndim = len(arr.shape)
        self._axes = tuple(range(ndim)) if axes is None else helpers.normalize_axes(ndim, tuple(axes))
        self._flags = flags
        self._tuning = tuning
        self.native_plan: NativeFftPlan | None = None

    def _build_trivial_plan(self, plan_factory: Callable[[], ComputationPlan], args: KernelArguments) -> ComputationPlan:
        plan = plan_factory()
        copy_trf = copy(args.input, out_arr_t=args.output)
        plan.computation_call(PureParallel.from_trf(copy_trf, copy_trf.input), args.output, args.input)
        return plan

    def _build_plan(
        self, plan_factory: Callable[[], ComputationPlan], device_params: DeviceParameters, args: KernelArguments
    ) -> ComputationPlan:
        if helpers.product(args.input.shape[a] for a in self._axes) == 1:
            return self._build_trivial_plan(plan_factory, args)
        plan = plan_factory()
        output, input_, inverse = args.output, args.input, args.inverse
        tree = plan._tr_tree
        fused_in = tree.nodes[input_.name].input_ntr is not None
        fused_out = tree.nodes[output.name].output_ntr is not None
        in_meta = _meta(input_)
        out_meta = _meta(output)
        if fused_in:   # the transformation addresses the user arrays; the plan sees a contiguous view
            in_meta = ArrayMetadata(in_meta.shape, in_meta.dtype)
        if fused_out:
            out_meta = ArrayMetadata(out_meta.shape, out_meta.dtype)
        flags = self._flags | (_lib.RK_FFT_OUT_LAST_ONLY if fused_out else 0)
        if plan.virtual:
            flags |= _lib.RK_FFT_DRY_RUN
        native = NativeFftPlan(in_meta, out_meta, self._axes, flags, self._tuning)
        self.native_plan = native
        ws = plan.temp_array((native.workspace_bytes,), numpy.uint8) if native.workspace_bytes else None

        if not fused_in and not fused_out:
            in_name, out_name, inv_name = input_.name, output.name, inverse.name
            ws_name = ws.name if ws is not None else None

            def run(queue: Queue, values: dict[str, Any]) -> None:
                native.execute(
                    _ptr(values[out_name]), _ptr(values[in_name]), int(values[inv_name]),
                    _ptr(values[ws_name]) if ws_name else None, queue.stream,
                )

            names = [out_name, in_name, inv_name] + ([ws_name] if ws_name else [])
            plan.native_call(run, names, keep=native)
            return plan

        fused = _FusedFft(plan, native, output, input_, inverse, ws, fused_in, fused_out)
        plan.native_call(fused, fused.argnames, keep=(native, fused))
        return plan


class _FusedFft:
    """Generates and launches the transformation-carrying first/last pass kernels."""

    def __init__(self, plan: ComputationPlan, native: NativeFftPlan, output: Any, input_: Any, inverse: Any,
                 ws: Any, fused_in: bool, fused_out: bool):
        self.native = native
        tree = plan._tr_tree
        out_ann = plan._get_annotation(output.name)
        in_ann = plan._get_annotation(input_.name)
        subtree = tree.get_subtree([Parameter(output.name, out_ann), Parameter(input_.name, in_ann)])
        argobjs = {a.name: a for a in subtree.get_kernel_argobjects()}
        self.load_leaves = subtree.get_leaf_parameters([input_.name]) if fused_in else []
        self.store_leaves = subtree.get_leaf_parameters([output.name]) if fused_out else []
        self.in_name = None if fused_in else input_.name
        self.out_name = None if fused_out else output.name
        self.inv_name = inverse.name
        self.ws_name = ws.name if ws is not None else None
        self.argnames = [p.name for p in self.load_leaves + self.store_leaves] + [inverse.name]
        if self.in_name:
            self.argnames.append(self.in_name)
        if self.out_name:
            self.argnames.append(self.out_name)
        if self.ws_name:
            self.argnames.append(self.ws_name)
        passes = native.passes
        ndim = len(native.shape)
        targets = {}
        if fused_in:
            targets.setdefault(0, {})["load"] = passes[0]
        if fused_out:
            targets.setdefault(len(passes) - 1, {})["store"] = passes[-1]
        self.sources = {}
        options = []
        for index, sides in targets.items():
            info = passes[index]
            a = info.axis
            slices = (a, 1, ndim - a - 1)
            src = self._source(info, slices, argobjs.get(input_.name) if "load" in sides else None,
                               argobjs.get(output.name) if "store" in sides else None)
            self.sources[index] = src
            fn = compile_kernel(src, "rk_fused_pass", options, check_only=plan.virtual)
            if not plan.virtual:
                native.set_pass_kernel(index, fn)

    def _source(self, info: Any, slices: tuple[int, int, int], in_obj: Any, out_obj: Any) -> str:
        def struct(name: str, leaves: list[Parameter]) -> str:
            fields = [f"    {param_cname(p, qualified=True)};" for p in leaves] or ["    char _unused;"]
            return f"struct {name} {{\n" + "\n".join(fields) + "\n};\n"

        def bind(src: str, leaves: list[Parameter]) -> str:
            return "".join(f"        auto {leaf_name(p.name)} = {src}.{leaf_name(p.name)};\n" for p in leaves)

        xtype = "double2" if self.native.dtype == numpy.complex128 else "float2"

        def render() -> str:
            parts = [struct("RkLoadArgs", self.load_leaves), struct("RkStoreArgs", self.store_leaves)]
            load_body = "        return reinterpret_cast<const value_type*>(a.in)[addr];\n"
            store_body = "        reinterpret_cast<value_type*>(a.out)[addr] = v;\n"
            if in_obj is not None:
                mod = in_obj.load_combined_idx(slices)
                load_body = bind("la", self.load_leaves) + f"        return {mod}((long long)o, (long long)m, (long long)i);\n"
            if out_obj is not None:
                mod = out_obj.store_combined_idx(slices)
                store_body = bind("sa", self.store_leaves) + f"        {mod}((long long)o, (long long)k, (long long)i, v);\n"
            parts.append(
                "struct RkFusedIO {\n"
                f"    using value_type = {xtype};\n"
                "    static constexpr bool kBranchFree = true;\n"

                "    const RkLoadArgs& la;\n"
                "    const RkStoreArgs& sa;\n"
                f"    __device__ __forceinline__ {xtype} load(const rk_pass_args& a, rk_u64 o, rk_u64 m, rk_u64 i,"
                " rk_i64 addr) const {\n" + load_body + "    }\n"
                "    __device__ __forceinline__ void store(const rk_pass_args& a, rk_u64 o, rk_u64 k, rk_u64 i,"
                f" rk_i64 addr, {xtype} v) const {{\n" + store_body + "    }\n"
                "};\n"
            )
            # the pass's load/store modes compiled in (fft_device.hpp RK_MODES)
            modes = (f"RK_MODES({_MODE_CODES[info.load_mode]}, {_MODE_CODES[info.store_mode]})"
                     if _fixed_modes() else "0")
            # launch bounds of the plan's own kernel for the residency the
            # tile's LDS allows (fft_device.hpp RK_VAR_OCC, RK_VAR_F64)
            bounds = f"{info.threads}"
            if info.occ:
                v = info.variant | 65536 | (131072 if xtype == "double2" else 0)
                bounds += f", (rkfft::pass_min_waves<{info.L}, {info.C}, {v}>())"
            parts.append(
                f'extern "C" __global__ void __launch_bounds__({bounds}) '
                "rk_fused_pass(rk_pass_args a, rk_user_blob lb, rk_user_blob sb) {\n"
                "    const RkLoadArgs& la = *reinterpret_cast<const RkLoadArgs*>(lb.b);\n"
                "    const RkStoreArgs& sa = *reinterpret_cast<const RkStoreArgs*>(sb.b);\n"
                f"    rkfft::fft_pass_body<{info.L}, {info.C}, {info.variant}, RkFusedIO, {modes}, "
                f"{'true' if info.wave else 'false'}>(a, RkFusedIO{{la, sa}});\n"
                "}\n"
            )
            return "".join(parts)

        body = render_with_modules(render)
        return _device_headers() + PRELUDE + body

    def _blob(self, leaves: list[Parameter], values: dict[str, Any]) -> bytes | None:
        if not leaves:
            return None
        packed = []
        for p in leaves:
            v = values[p.name]
            packed.append((p.annotation, _ptr(v) if p.annotation.is_array() else v))
        blob = pack_arguments(packed)
        if len(blob) > _lib.RK_FUSED_BLOB_BYTES:
            raise ValueError("too many transformation arguments for one fused FFT kernel")
        return blob

    def __call__(self, queue: Queue, values: dict[str, Any]) -> None:
        in_ptr = _ptr(values[self.in_name]) if self.in_name else 1   # unused when the load is fused
        out_ptr = _ptr(values[self.out_name]) if self.out_name else 1
        ws_ptr = _ptr(values[self.ws_name]) if self.ws_name else None
        self.native.execute(
            out_ptr, in_ptr, int(values[self.inv_name]), ws_ptr, queue.stream,
            load_blob=self._blob(self.load_leaves, values) or b"\0",
            store_blob=self._blob(self.store_leaves, values) or b"\0",
        )


_HEADERS: str | None = None


def _device_headers() -> str:
    """fft_args.h + fft_device.hpp inlined for hiprtc (no include paths needed)."""
    global _HEADERS
    if _HEADERS is None:
        from pathlib import Path

        here = Path(__file__).resolve().parent.parent / "csrc"
        args_h = (here / "fft_args.h").read_text()
        dev = (here / "fft_device.hpp").read_text().replace('#include "fft_args.h"', "")
        _HEADERS = args_h + "\n" + dev + "\n"
    return _HEADERS
