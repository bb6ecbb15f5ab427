"""Computations: prepare (connect transformations), compile (build a plan),
call (enqueue).

Contract of reikna/core/computation.py:
  * ``Computation(root_parameters)`` with ``signature``/``parameter`` and
    ``connect`` (:115-222) — the signature rewriting is the
    TransformationTree's;
  * ``compile(device, virtual_manager=None, compiler_options=[], *,
    fast_math=False, keep=False)`` calls the subclass's ``_build_plan`` with a
    plan factory, the device parameters and ``KernelArguments`` (:227-308);
  * ``ComputationPlan`` records kernel calls, nested computation calls, temp /
    persistent / constant arrays (:346-703);
  * ``ComputationCallable(queue, *args, **kwds)`` binds the arguments
    (casting scalars, filling defaults) and enqueues every step (:732-801).

Kernels given to ``kernel_call`` are templates (reikna_amd.template) that
render HIP C++; they are compiled by hiprtc (cached on disk) and launched on
the queue's stream.  Native steps (the FFT passes) go through the C-ABI.
"""

from __future__ import annotations

import ctypes
import hashlib
import math
import weakref
from collections.abc import Callable, Iterable, Mapping, Sequence
from typing import Any

import numpy

from .. import _lib, dtypes
from ..helpers import Graph, bounding_power_of_2, product
from ..runtime import Array, ArrayMetadata, Device, DeviceParameters, Queue
from ..template import DefTemplate, render_with_modules
from .signature import Annotation, Parameter, Signature, Type
from .transformation import TransformationParameter, TransformationTree, leaf_name


class ComputationParameter(Type):
    """A typed computation parameter; usable where array metadata is expected."""

    def __init__(self, computation: Any, name: str, type_: Type):
        super().__init__(type_)
        self._computation = weakref.ref(computation)
        self._name = name

    def belongs_to(self, comp: Any) -> bool:
        return self._computation() is comp

    def connect(self, _trf: Any, _tr_connector: Any, **tr_from_comp: Any) -> Computation:
        return self._computation().connect(self._name, _trf, _tr_connector, **tr_from_comp)

    def __str__(self) -> str:
        return self._name


class Translator:
    """Maps known names to new ones and prefixes the rest (nested computations)."""

    def __init__(self, known_old: Sequence[str], known_new: Sequence[str], prefix: str):
        self._mapping = dict(zip(known_old, known_new))
        self._prefix = prefix

    def __call__(self, name: str) -> str:
        if name in self._mapping:
            return self._mapping[name]
        return (self._prefix + "_" if self._prefix else "") + name

    def get_nested(self, known_old: Sequence[str], known_new: Sequence[str], prefix: str) -> Translator:
        return Translator(known_old, known_new, prefix + self._prefix)

    @classmethod
    def identity(cls) -> Translator:
        return cls([], [], "")


def check_external_parameter_name(name: str) -> None:
    if name.startswith("_"):
        raise ValueError("External parameter name cannot start with the underscore.")


class ParameterContainer:
    def __init__(self, parent: Any, parameters: Iterable[Parameter]):
        self._param_objs = {p.name: ComputationParameter(parent, p.name, p.annotation) for p in parameters}

    def __getattr__(self, name: str) -> ComputationParameter:
        try:
            return self.__dict__["_param_objs"][name]
        except KeyError:
            raise AttributeError(name) from None


class Computation:
    """Base class of computations; subclasses implement ``_build_plan``."""

    def __init__(self, root_parameters: Iterable[Parameter]):
        root_parameters = list(root_parameters)
        for p in root_parameters:
            check_external_parameter_name(p.name)
        self._original_parameters = root_parameters
        self._tr_tree = TransformationTree(root_parameters)
        self._update_attributes()

    def _update_attributes(self) -> None:
        leaves = self._tr_tree.get_leaf_parameters()
        self.signature = Signature(leaves)
        self.parameter = ParameterContainer(self, leaves)

    def connect(self, _comp_connector: Any, _trf: Any, _tr_connector: Any, **tr_from_comp: Any) -> Computation:
        """Attach transformation ``_trf`` to parameter ``_comp_connector`` (see reikna docs)."""
        if isinstance(_comp_connector, ComputationParameter) and not _comp_connector.belongs_to(self):
            raise ValueError("The connection target must belong to this computation.")
        name = str(_comp_connector)
        if name in tr_from_comp:
            raise ValueError(
                "Parameter '" + name + "' cannot be supplied both as the main connector and one of the child connections"
            )
        tr_from_comp[name] = _tr_connector
        comp_from_tr = {}
        for comp_name, tr_conn in tr_from_comp.items():
            check_external_parameter_name(comp_name)
            if isinstance(tr_conn, TransformationParameter) and not tr_conn.belongs_to(_trf):
                raise ValueError("The transformation parameter must belong to the provided transformation")
            comp_from_tr[str(tr_conn)] = comp_name
        self._tr_tree.connect(name, _trf, comp_from_tr)
        self._update_attributes()
        return self

    def _translate_tree(self, translator: Translator) -> TransformationTree:
        return self._tr_tree.translate(translator)

    def _get_plan(
        self,
        tr_tree: TransformationTree,
        translator: Translator,
        device: Device,
        compiler_options: Sequence[str],
        *,
        fast_math: bool,
        keep: bool,
    ) -> ComputationPlan:
        def plan_factory() -> ComputationPlan:
            return ComputationPlan(tr_tree, translator, device, compiler_options, fast_math=fast_math, keep=keep)

        args = KernelArguments(
            {
                lp.name: KernelArgument(p.name, p.annotation)
                for lp, p in zip(self._original_parameters, tr_tree.get_root_parameters())
            }
        )
        return self._build_plan(plan_factory, device.params, args)

    def compile(
        self,
        device: Any = None,
        virtual_manager: Any = None,
        compiler_options: Sequence[str] = (),
        *,
        fast_math: bool = False,
        keep: bool = False,
    ) -> ComputationCallable:
        """Build the plan for ``device`` (a Device, a Queue or None for device 0)."""
        dev = device.device if isinstance(device, Queue) else (device if device is not None else Device(0))
        dev.activate()
        plan = self._get_plan(
            self._tr_tree, Translator.identity(), dev, list(compiler_options), fast_math=fast_math, keep=keep
        )
        return plan.finalize()

    def _build_plan(
        self, plan_factory: Callable[[], ComputationPlan], device_params: DeviceParameters, args: KernelArguments
    ) -> ComputationPlan:
        raise NotImplementedError


class KernelArguments:
    def __init__(self, args: Mapping[str, KernelArgument]):
        self._args = dict(args)

    def all(self) -> list[KernelArgument]:
        return list(self._args.values())

    def __getattr__(self, name: str) -> KernelArgument:
        try:
            return self.__dict__["_args"][name]
        except KeyError:
            raise AttributeError(name) from None


class KernelArgument(Type):
    """An argument usable in kernel or nested computation calls of a plan."""

    def __init__(self, name: str, type_: Type):
        super().__init__(type_)
        self.name = name

    def __repr__(self) -> str:
        return "KernelArgument(" + self.name + ")"


class _IdGen:
    def __init__(self, prefix: str):
        self._n = 0
        self._prefix = prefix

    def __call__(self) -> str:
        self._n += 1
        return self._prefix + str(self._n)


# ---------------------------------------------------------------------------
# generated kernels
# ---------------------------------------------------------------------------
PRELUDE = r"""
#define KERNEL extern "C" __global__
#define GLOBAL_MEM
#define LOCAL_MEM __shared__
#define LOCAL_MEM_DYNAMIC extern __shared__
#define LOCAL_MEM_ARG
#define CONSTANT_MEM
#define CONSTANT_MEM_ARG
#define INLINE __forceinline__
#define FUNCTION __device__
#define WITHIN_KERNEL __device__
#define LOCAL_BARRIER __syncthreads()
#define VSIZE_T long long
#define COMPLEX_CTR(T) make_##T
"""


def _hip_align(dtype: numpy.dtype) -> int:
    if dtypes.is_complex(dtype):
        return dtype.itemsize
    return max(1, dtype.alignment)


class StaticSizes:
    """Virtual (row-major, any number of dims) global/local sizes mapped to a 1-D HIP grid."""

    def __init__(self, global_size: Sequence[int], local_size: Sequence[int] | None, max_threads: int = 1024):
        gs = tuple(int(g) for g in global_size) or (1,)
        if local_size is None:
            ls = [1] * len(gs)
            budget = 256
            for d in range(len(gs) - 1, -1, -1):
                ls[d] = min(bounding_power_of_2(gs[d]), budget)
                budget //= ls[d]
                if budget <= 1:
                    break
            local_size = ls
        ls = tuple(int(x) for x in local_size)
        if len(ls) != len(gs):
            raise ValueError("global and local sizes must have the same number of dimensions")
        if product(ls) > max_threads:
            raise ValueError(f"local size {ls} exceeds {max_threads} threads")
        self.global_size = gs
        self.local_size = ls
        self.num_groups = tuple((g + l - 1) // l for g, l in zip(gs, ls))
        self.threads = product(ls)
        self.blocks = product(self.num_groups)
        n = len(gs)
        lines = []
        for name, vals in (("rk_vgs", gs), ("rk_vls", ls), ("rk_vng", self.num_groups)):
            cases = " ".join(f"case {d}: return {v}LL;" for d, v in enumerate(vals))
            lines.append(f"__device__ __forceinline__ long long {name}(int d) {{ switch (d) {{ {cases} }} return 1LL; }}")

        def decompose(fn: str, flat: str, sizes: str) -> str:
            return (
                f"__device__ __forceinline__ long long {fn}(int d) {{\n"
                f"    long long r = {flat};\n"
                f"    for (int k = {n - 1}; k > d; --k) r /= {sizes}(k);\n"
                f"    return r % {sizes}(d);\n}}"
            )

        lines.append(decompose("rk_local_id", "(long long)threadIdx.x", "rk_vls"))
        lines.append(decompose("rk_group_id", "(long long)blockIdx.x + (long long)blockIdx.y * gridDim.x", "rk_vng"))
        lines.append(
            "__device__ __forceinline__ long long rk_global_id(int d) { return rk_group_id(d) * rk_vls(d) + rk_local_id(d); }"
        )
        checks = " || ".join(f"rk_global_id({d}) >= rk_vgs({d})" for d in range(n))
        lines.append(f"__device__ __forceinline__ bool rk_skip() {{ return {checks}; }}")
        lines.append(
            "__device__ __forceinline__ long long rk_global_flat_id() { "
            + "long long f = 0; "
            + " ".join(f"f = f * rk_vgs({d}) + rk_global_id({d});" for d in range(n))
            + " return f; }"
        )
        self.source = "\n".join(lines) + "\n"


class _Static:
    """The ``static`` object kernel templates use (``${static.global_id}(0)`` ...)."""

    skip = "rk_skip"
    global_id = "rk_global_id"
    local_id = "rk_local_id"
    group_id = "rk_group_id"
    global_size = "rk_vgs"
    local_size = "rk_vls"
    num_groups = "rk_vng"
    global_flat_id = "rk_global_flat_id"


_module_cache: dict[str, tuple[int, int]] = {}


def compile_kernel(source: str, name: str, options: Sequence[str] = (), *, check_only: bool = False) -> int:
    """Compile HIP source with hiprtc (cached) and return the function handle
    (with ``check_only``: compile without a device and return 0)."""
    key = hashlib.sha256((source + "\0" + " ".join(options) + "\0" + name).encode()).hexdigest()
    if key in _module_cache:
        return _module_cache[key][1]
    lib = _lib.load()
    mod = ctypes.c_void_p()
    log = ctypes.create_string_buffer(65536)
    opts = " ".join(["-Xclang -target-feature -Xclang -packed-fp32-ops", *options])
    if check_only:
        rc = lib.rk_module_check(source.encode(), opts.encode(), log, len(log))
        if rc != 0:
            raise _lib.RkError(rc, f"kernel {name} failed to compile: {_lib.last_error()}\n--- source ---\n{source}")
        return 0
    rc = lib.rk_module_compile(source.encode(), opts.encode(), ctypes.byref(mod), log, len(log))
    if rc != 0:
        raise _lib.RkError(rc, f"kernel {name} failed to compile: {_lib.last_error()}\n--- source ---\n{source}")
    fn = ctypes.c_void_p()
    _lib.call("rk_module_get_function", mod, name.encode(), ctypes.byref(fn))
    _module_cache[key] = (mod.value, fn.value)
    return fn.value


def pack_arguments(values: Sequence[tuple[Annotation, Any]]) -> bytes:
    """Lay out kernel arguments (arrays as base pointers, scalars by value)."""
    buf = bytearray()
    for ann, val in values:
        if ann.is_array():
            align, data = 8, int(val).to_bytes(8, "little")
        else:
            dt = ann.dtype
            align = _hip_align(dt)
            data = numpy.asarray(val, dtype=dt).tobytes()
        while len(buf) % align:
            buf.append(0)
        buf += data
    while len(buf) % 8:
        buf.append(0)
    return bytes(buf)


def _check_array_argument(name: str, val: Any, meta: ArrayMetadata) -> None:
    """The kernels address an array argument with the layout of its parameter
    (compiled in): an array of another shape, type or layout, or one whose
    buffer is smaller than that layout spans, is refused instead of being
    read or written out of place."""
    if isinstance(val, Array):
        arr = val
    elif hasattr(val, "data_ptr") and hasattr(val, "is_cuda"):   # torch tensor
        arr = Array.from_torch(val)
    else:
        raise TypeError(f"{name}: expected a device Array, got {type(val)}")
    have = arr.metadata
    if have.shape != meta.shape or have.dtype != meta.dtype:
        raise ValueError(f"{name}: array of shape {have.shape} and type {have.dtype}, "
                         f"the computation was compiled for {meta.shape} {meta.dtype}")
    moved = any(n > 1 and a != b for n, a, b in zip(meta.shape, have.strides, meta.strides))
    if moved or have.first_element_offset != meta.first_element_offset:
        raise ValueError(f"{name}: array strides {have.strides} / offset {have.first_element_offset} differ "
                         f"from the compiled {meta.strides} / {meta.first_element_offset}")
    size = getattr(arr.buffer, "size", None)
    if size is not None and size < meta.buffer_size:
        raise ValueError(f"{name}: buffer of {size} bytes, the compiled layout spans {meta.buffer_size}")


def _array_ptr(val: Any) -> int:
    if isinstance(val, Array):
        return val.ptr
    if hasattr(val, "data_ptr") and hasattr(val, "is_cuda"):   # torch tensor
        return Array.from_torch(val).ptr
    raise TypeError(f"expected a device Array, got {type(val)}")


class PlannedKernelCall:
    """A compiled generated kernel and the names of its (leaf) arguments."""

    def __init__(self, fn: int, argnames: Sequence[str], annotations: Sequence[Annotation], sizes: StaticSizes,
                 adhoc: Mapping[str, Any], source: str, name: str):
        self.fn = fn
        self.argnames = list(argnames)
        self.annotations = list(annotations)
        self.sizes = sizes
        self.adhoc = dict(adhoc)
        self.source = source
        self.name = name

    def __call__(self, queue: Queue, values: Mapping[str, Any]) -> None:
        packed = []
        for n, ann in zip(self.argnames, self.annotations):
            v = self.adhoc[n] if n in self.adhoc else values[n]
            packed.append((ann, _array_ptr(v) if ann.is_array() else v))
        blob = pack_arguments(packed)
        blocks = self.sizes.blocks
        gx = min(blocks, 2**31 - 1)
        gy = (blocks + gx - 1) // gx
        grid = (ctypes.c_uint32 * 3)(gx, gy, 1)
        block = (ctypes.c_uint32 * 3)(self.sizes.threads, 1, 1)
        buf = ctypes.create_string_buffer(blob, len(blob))
        _lib.call("rk_launch", ctypes.c_void_p(self.fn), grid, block, 0, ctypes.c_void_p(queue.stream), buf,
                  ctypes.c_size_t(len(blob)))


class NativeStep:
    """A step implemented by native code (e.g. an FFT plan); `fn(queue, values)`."""

    def __init__(self, fn: Callable[[Queue, Mapping[str, Any]], None], argnames: Sequence[str], keep: Any = None):
        self.fn = fn
        self.argnames = list(argnames)
        self.keep = keep

    def __call__(self, queue: Queue, values: Mapping[str, Any]) -> None:
        self.fn(queue, values)


class ComputationPlan:
    """Records the steps of a computation (see the module docstring)."""

    def __init__(
        self,
        tr_tree: TransformationTree,
        translator: Translator,
        device: Device,
        compiler_options: Sequence[str],
        *,
        fast_math: bool,
        keep: bool,
    ):
        self._device = device
        self._tr_tree = tr_tree
        self._translator = translator
        self._compiler_options = list(compiler_options)
        self._fast_math = fast_math
        self._keep = keep
        self._nested_idgen = _IdGen("_nested")
        self._persistent_idgen = _IdGen("_value")
        self._constant_idgen = _IdGen("_constant")
        self._temp_idgen = _IdGen("_temp")
        self._external_annotations = tr_tree.get_root_annotations()
        self._persistent_values: dict[str, Any] = {}
        self._temp_arrays: set[str] = set()
        self._internal_annotations: dict[str, Annotation] = {}
        self._steps: list[Any] = []

    @property
    def device(self) -> Device:
        return self._device

    @property
    def virtual(self) -> bool:
        """True when planning against a VirtualDevice (no GPU: nothing is allocated)."""
        return bool(getattr(self._device, "virtual", False))

    # -- arrays ------------------------------------------------------------
    def _scalar(self, val: Any) -> KernelArgument:
        name = self._translator(self._persistent_idgen())
        ann = Annotation(numpy.asarray(val).dtype)
        self._internal_annotations[name] = ann
        self._persistent_values[name] = ann.cast_scalar(val)
        return KernelArgument(name, ann)

    def persistent_array(self, arr: numpy.ndarray) -> KernelArgument:
        name = self._translator(self._persistent_idgen())
        arr = numpy.ascontiguousarray(arr)
        ann = Annotation(ArrayMetadata.from_arraylike(arr), "i")
        self._internal_annotations[name] = ann
        self._persistent_values[name] = arr if self.virtual else Array.from_host(self._device, arr)
        return KernelArgument(name, ann)

    def constant_array(self, arr: numpy.ndarray) -> KernelArgument:
        name = self._translator(self._constant_idgen())
        arr = numpy.ascontiguousarray(arr)
        ann = Annotation(ArrayMetadata.from_arraylike(arr), constant=True)
        self._internal_annotations[name] = ann
        self._persistent_values[name] = arr if self.virtual else Array.from_host(self._device, arr)
        return KernelArgument(name, ann)

    def temp_array(
        self,
        shape: Sequence[int] | int,
        dtype: Any,
        strides: Sequence[int] | None = None,
        offset: int = 0,
        nbytes: int | None = None,
    ) -> KernelArgument:
        name = self._translator(self._temp_idgen())
        meta = ArrayMetadata(shape, dtype, strides=strides, first_element_offset=offset, buffer_size=nbytes)
        ann = Annotation(meta, "io")
        self._internal_annotations[name] = ann
        self._temp_arrays.add(name)
        return KernelArgument(name, ann)

    def temp_array_like(self, arr: Any) -> KernelArgument:
        meta = ArrayMetadata.from_arraylike(arr)
        return self.temp_array(meta.shape, meta.dtype, strides=meta.strides, offset=meta.first_element_offset,
                               nbytes=meta.buffer_size)

    def _get_annotation(self, name: str) -> Annotation:
        if name in self._external_annotations:
            return self._external_annotations[name]
        return self._internal_annotations[name]

    # -- kernel calls ------------------------------------------------------
    def _process_kernel_arguments(self, args: Iterable[Any]) -> tuple[list[Parameter], dict[str, Any]]:
        processed, adhoc = [], {}
        gen = _IdGen("_adhoc")
        for arg in args:
            if not isinstance(arg, KernelArgument):
                if hasattr(arg, "shape") and hasattr(arg, "dtype"):
                    if len(arg.shape) > 0:
                        raise ValueError("Arrays are not allowed as ad hoc arguments")
                    name = self._translator(gen())
                    adhoc[name] = arg
                    ann = Annotation(arg.dtype)
                else:
                    raise TypeError("Unknown argument type: " + str(type(arg)))
            else:
                ann = self._get_annotation(arg.name)
                name = arg.name
            processed.append(Parameter(name, ann))
        return processed, adhoc

    def kernel_call(
        self,
        template_def: DefTemplate,
        args: Iterable[Any],
        global_size: Sequence[int] | int,
        local_size: Sequence[int] | None = None,
        render_kwds: Mapping[str, Any] | None = None,
        kernel_name: str = "_kernel_func",
    ) -> None:
        """Add a generated kernel: ``template_def`` is rendered with the kernel
        declaration and one KernelParameter per argument, compiled for gfx950
        and launched over ``global_size`` (row-major)."""
        processed, adhoc = self._process_kernel_arguments(args)
        subtree = self._tr_tree.get_subtree(processed)
        decl, leaf_names = subtree.get_kernel_declaration(kernel_name)
        argobjects = subtree.get_kernel_argobjects()
        if isinstance(global_size, int):
            global_size = (global_size,)
        sizes = StaticSizes(global_size, local_size, self._device.params.max_total_local_size)
        kwds = dict(render_kwds or {})
        kwds.setdefault("static", _Static)
        kwds.setdefault("dtypes", dtypes)
        body = render_with_modules(lambda: template_def.render(decl, *argobjects, **kwds))
        source = PRELUDE + sizes.source + body
        opts = list(self._compiler_options) + (["-ffast-math"] if self._fast_math else [])
        fn = compile_kernel(source, kernel_name, opts, check_only=self.virtual)
        leaf_params = {p.name: p for p in subtree.get_leaf_parameters()}
        anns = [leaf_params[n].annotation for n in leaf_names]
        self._steps.append(PlannedKernelCall(fn, leaf_names, anns, sizes, adhoc, source, kernel_name))

    def native_call(self, fn: Callable[[Queue, Mapping[str, Any]], None], argnames: Sequence[str],
                    keep: Any = None) -> None:
        """Add a native step (used by the FFT): ``fn(queue, values_by_name)``."""
        self._steps.append(NativeStep(fn, argnames, keep))

    # -- nested computations -----------------------------------------------
    def _process_computation_arguments(self, signature: Signature, args: Sequence[Any],
                                       kwds: Mapping[str, Any]) -> list[str]:
        bound = signature.bind_with_defaults(args, kwds, cast=False)
        names = []
        for arg, param in zip(bound.args, signature.parameters.values()):
            if not isinstance(arg, KernelArgument):
                if param.annotation.is_array():
                    raise ValueError("Ad hoc arguments are only allowed for scalar parameters")
                karg = self._scalar(param.annotation.cast_scalar(arg))
            else:
                karg = arg
            ann = self._get_annotation(karg.name)
            if not ann.can_be_argument_for(param.annotation):
                raise ValueError(f"Got {ann} for '{param.name}', expected {param.annotation}")
            names.append(karg.name)
        return names

    def computation_call(self, computation: Computation, *args: Any, **kwds: Any) -> None:
        """Add a call of a nested computation with plan arguments or scalars."""
        signature = computation.signature
        argnames = self._process_computation_arguments(signature, args, kwds)
        translator = self._translator.get_nested(list(signature.parameters), argnames, self._nested_idgen())
        new_tree = computation._translate_tree(translator)
        new_tree.reconnect(self._tr_tree)
        nested = computation._get_plan(new_tree, translator, self._device, self._compiler_options,
                                       fast_math=self._fast_math, keep=self._keep)
        self._steps += nested._steps
        self._persistent_values.update(nested._persistent_values)
        self._temp_arrays.update(nested._temp_arrays)
        self._internal_annotations.update(nested._internal_annotations)

    # -- finalisation ------------------------------------------------------
    def finalize(self) -> ComputationCallable:
        usage: dict[str, list[int]] = {}
        for i, step in enumerate(self._steps):
            for n in step.argnames:
                if n in self._temp_arrays:
                    usage.setdefault(n, [i, i])[1] = i
        deps = Graph()
        for n, (a, b) in usage.items():
            for i in range(a, b + 1):
                for other in self._steps[i].argnames:
                    if other in self._temp_arrays and other != n:
                        deps.add_edge(n, other)
        internal: dict[str, Any] = dict(self._persistent_values)
        temps = []
        if self.virtual:
            return ComputationCallable(self._device, self._tr_tree.get_leaf_parameters(), self._steps, internal, temps)
        # temporaries whose live ranges do not overlap share one allocation
        pool: list[tuple[Array, set[str]]] = []
        for name in sorted(self._temp_arrays, key=lambda n: -self._internal_annotations[n].nbytes):
            meta = self._internal_annotations[name].array_metadata
            arr = None
            for base, users in pool:
                if base.metadata.buffer_size >= meta.buffer_size and not (users & deps[name]):
                    arr = Array(self._device, meta, base.buffer)
                    users.add(name)
                    break
            if arr is None:
                arr = Array(self._device, meta)
                pool.append((arr, {name}))
            internal[name] = arr
            temps.append(arr)
        return ComputationCallable(self._device, self._tr_tree.get_leaf_parameters(), self._steps, internal, temps)


class ComputationCallable:
    """A compiled computation: ``callable(queue, *args, **kwds)``."""

    def __init__(self, device: Device, parameters: Sequence[Parameter], steps: Sequence[Any],
                 internal_args: Mapping[str, Any], temp_buffers: Sequence[Array]):
        self.device = device
        self.signature = Signature(parameters)
        self.parameter = ParameterContainer(self, parameters)
        self._steps = list(steps)
        self._internal_args = dict(internal_args)
        self.__tempalloc__ = list(temp_buffers)

    def __call__(self, queue: Queue, *args: Any, **kwds: Any) -> list[Any]:
        bound = self.signature.bind_with_defaults(args, kwds, cast=True)
        for name, val in bound.arguments.items():
            ann = self.signature.parameters[name].annotation
            if ann.is_array():
                _check_array_argument(name, val, ann.array_metadata)
        values = dict(self._internal_args)
        values.update(bound.arguments)
        for step in self._steps:
            step(queue, values)
        return []

    @property
    def kernel_sources(self) -> list[str]:
        """Sources of the generated kernels (the ``keep=True`` analogue)."""
        return [s.source for s in self._steps if isinstance(s, PlannedKernelCall)]
