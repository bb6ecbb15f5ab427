"""Transformations and the transformation tree.

The contract is reikna/core/transformation.py: a ``Transformation`` is a pure
elementwise snippet with array/scalar parameters (no 'io' parameters,
:63-91); connecting it to a computation parameter rewrites the computation's
signature (``TransformationTree.connect`` / ``_connect``, :313-398), and at
kernel-generation time every array argument of a kernel gets load/store
helpers (``load_idx``, ``store_idx``, ``load_same``, ``store_same``,
``load_combined_idx``, ``store_combined_idx``; :489-552) that inline the
connected transformation chain.

The ordering rules of the rewritten signature (depth-first over the
connection graph, with siblings keeping their order, :250-299) are part of
the API (pinned by tests/test_core/test_transformation.py:78-117), so the
traversal below follows the same rules; the code it generates is HIP C++.
"""

from __future__ import annotations

import weakref
from collections.abc import Callable, Iterable, Iterator, Mapping, Sequence
from typing import Any

from ..template import DefTemplate, Module, Snippet
from .signature import Annotation, Parameter, Signature, Type

VALUE_NAME = "_val"
INDEX_NAME = "_idx"


def leaf_name(name: str) -> str:
    return "_leaf_" + name


def index_cnames(shape: Sequence[int]) -> list[str]:
    return [INDEX_NAME + str(i) for i in range(len(shape))]


# ---------------------------------------------------------------------------
# transformations
# ---------------------------------------------------------------------------
class TransformationParameter(Type):
    """A typed parameter of a Transformation (usable wherever array metadata is)."""

    def __init__(self, trf: Transformation, name: str, type_: Type):
        super().__init__(type_)
        self._trf = weakref.ref(trf)
        self._name = name

    def belongs_to(self, trf: Transformation) -> bool:
        return self._trf() is trf

    def __str__(self) -> str:
        return self._name


class _ParamContainer:
    def __init__(self, objs: Mapping[str, Any]):
        self._param_objs = dict(objs)

    def __getattr__(self, name: str) -> Any:
        try:
            return self.__dict__["_param_objs"][name]
        except KeyError:
            raise AttributeError(name) from None

    def __iter__(self) -> Iterator[Any]:
        return iter(self._param_objs.values())


class Transformation:
    """
    A pure elementwise transformation of arrays.

    :param parameters: list of Parameter objects (roles 'i', 'o' or 's').
    :param code: template source; it is rendered as a def with positional
        parameters ``idxs`` followed by the parameter names.
    :param render_kwds: extra names visible in the template.
    :param connectors: parameter names that may be connected (default: arrays).
    """

    def __init__(
        self,
        parameters: Sequence[Parameter],
        code: str,
        render_kwds: Mapping[str, Any] | None = None,
        connectors: Sequence[str] | None = None,
    ):
        for param in parameters:
            if param.annotation.input and param.annotation.output:
                raise ValueError(f"Transformation cannot have 'io' parameters ('{param.name}')")
        self.signature = Signature(parameters)
        objs = {p.name: TransformationParameter(self, p.name, p.annotation) for p in parameters}
        for name, obj in objs.items():
            setattr(self, name, obj)
        self.parameter = _ParamContainer(objs)
        self.connectors = list(connectors) if connectors is not None else [
            p.name for p in parameters if p.annotation.is_array()
        ]
        names = ["idxs"] + [p.name for p in parameters]
        self.snippet = Snippet(DefTemplate.from_string("transformation", names, code), render_kwds or {})


# ---------------------------------------------------------------------------
# the tree
# ---------------------------------------------------------------------------
class NodeTransformation:
    """An edge: `trf` attached to node `connector_node_name`."""

    def __init__(self, connector_node_name: str, trf: Transformation, node_from_tr: Mapping[str, str]):
        self.trf = trf
        self.connector_node_name = connector_node_name
        self.node_from_tr = dict(node_from_tr)
        given, available = set(self.node_from_tr), set(trf.signature.parameters)
        if given != available:
            raise ValueError(
                f"Supplied transformation names {sorted(given)} "
                f"do not fully coincide with the existing ones: {sorted(available)}"
            )
        conn = [t for t, n in self.node_from_tr.items() if n == connector_node_name]
        if len(conn) != 1:
            raise ValueError("There should be only one transformation parameter corresponding to the connector node")
        # output transformation: the connector parameter is an input of the transformation
        self.output = trf.signature.parameters[conn[0]].annotation.input

    def ordered_node_names(self) -> list[str]:
        return [self.node_from_tr[p] for p in self.trf.signature.parameters]

    def get_child_names(self) -> tuple[list[str], list[str]]:
        names = self.ordered_node_names()
        k = names.index(self.connector_node_name)
        return names[:k], names[k + 1 :]

    def translate_node_names(self, translator: Callable[[str], str]) -> NodeTransformation:
        return NodeTransformation(
            translator(self.connector_node_name),
            self.trf,
            {t: translator(n) for t, n in self.node_from_tr.items()},
        )


class Node:
    """A tree node: at most one input and one output transformation."""

    def __init__(self, input_ntr: NodeTransformation | None = None, output_ntr: NodeTransformation | None = None):
        self.input_ntr = input_ntr
        self.output_ntr = output_ntr

    def connect(self, ntr: NodeTransformation) -> Node:
        return Node(self.input_ntr, ntr) if ntr.output else Node(ntr, self.output_ntr)

    def get_child_names(self) -> tuple[list[str], list[str]]:
        o = self.output_ntr.get_child_names() if self.output_ntr else ([], [])
        i = self.input_ntr.get_child_names() if self.input_ntr else ([], [])
        if self.output_ntr is None:
            return i
        if self.input_ntr is None:
            return o
        # both: output side first, then the input side
        return o[0] + o[1], i[0] + i[1]

    def get_connections(self) -> list[NodeTransformation]:
        return [n for n in (self.output_ntr, self.input_ntr) if n is not None]


class TransformationTree:
    """The graph of transformations attached to a computation's parameters."""

    def __init__(self, root_parameters: Iterable[Parameter]):
        self.root_names: list[str] = []
        self.root_parameters: dict[str, Parameter] = {}
        self.nodes: dict[str, Node] = {}
        self.leaf_parameters: dict[str, Parameter] = {}
        for param in root_parameters:
            self.root_names.append(param.name)
            old = self.root_parameters.get(param.name)
            if old is not None and old != param:
                # the same array passed as 'i' and 'o' of a nested computation: merge into 'io'
                if old.annotation.type != param.annotation.type:
                    raise ValueError(f"Type mismatch for two parameters with name `{param.name}`")
                if old.default != param.default:
                    raise ValueError(f"Default value mismatch for two parameters with name `{param.name}`")
                merged = Parameter(param.name, Annotation(param.annotation, "io"), default=param.default)
                self.root_parameters[param.name] = merged
                self.leaf_parameters[param.name] = merged
            elif old is None:
                self.nodes[param.name] = Node()
                self.root_parameters[param.name] = param
                self.leaf_parameters[param.name] = param

    # -- traversal ---------------------------------------------------------
    def _walk(self, names: Sequence[str], ignore: set[str], visited: set[str], leaves_only: bool) -> list[str]:
        result: list[str] = []
        for k, name in enumerate(names):
            if name in ignore or name in visited:
                continue
            visited.add(name)
            later = set(names[k + 1 :])
            before, after = self.nodes[name].get_child_names()
            sub_before = self._walk(before, later, visited, leaves_only)
            sub_after = self._walk(after, later, visited, leaves_only)
            if not leaves_only:
                result.append(name)
            result += sub_before
            if leaves_only and name in self.leaf_parameters:
                result.append(name)
            result += sub_after
        return result

    def get_subtree_names(self, root_names: Sequence[str] | None = None, *, leaves_only: bool = False) -> list[str]:
        return self._walk(self.root_names if root_names is None else list(root_names), set(), set(), leaves_only)

    def get_root_annotations(self) -> dict[str, Annotation]:
        return {n: p.annotation for n, p in self.root_parameters.items()}

    def get_root_parameters(self) -> list[Parameter]:
        return [self.root_parameters[n] for n in self.root_names]

    def get_leaf_parameters(self, root_names: Sequence[str] | None = None) -> list[Parameter]:
        return [self.leaf_parameters[n] for n in self.get_subtree_names(root_names, leaves_only=True)]

    # -- connection --------------------------------------------------------
    def _connect(self, ntr: NodeTransformation) -> None:
        for tr_param in ntr.trf.signature.parameters.values():
            node_name = ntr.node_from_tr[tr_param.name]
            if node_name == ntr.connector_node_name:
                ann = self.leaf_parameters[node_name].annotation
                if ann.input and ann.output:
                    # an 'io' leaf keeps the data path the transformation does not use
                    role = "i" if ntr.output else "o"
                    self.leaf_parameters[node_name] = Parameter(node_name, Annotation(ann, role=role))
                else:
                    del self.leaf_parameters[node_name]
            elif node_name in self.leaf_parameters and self.leaf_parameters[node_name].annotation.is_array():
                ann = self.leaf_parameters[node_name].annotation
                if (ann.input and ntr.output) or (ann.output and not ntr.output):
                    self.leaf_parameters[node_name] = Parameter(node_name, Annotation(ann, role="io"))
            else:
                self.leaf_parameters[node_name] = tr_param.rename(node_name)
            if node_name not in self.nodes:
                self.nodes[node_name] = Node()
        self.nodes[ntr.connector_node_name] = self.nodes[ntr.connector_node_name].connect(ntr)

    def connect(self, comp_connector: str, trf: Transformation, comp_from_tr: Mapping[str, str]) -> None:
        ntr = NodeTransformation(comp_connector, trf, comp_from_tr)
        for tr_name, node_name in comp_from_tr.items():
            if node_name not in self.leaf_parameters:
                if node_name == comp_connector:
                    raise ValueError("Parameter '" + node_name + "' is not a part of the signature")
                if node_name in self.nodes:
                    raise ValueError("Parameter '" + node_name + "' is hidden by transformations")
                continue
            node_ann = self.leaf_parameters[node_name].annotation
            tr_ann = trf.signature.parameters[tr_name].annotation
            if tr_ann.type != node_ann.type:
                raise ValueError(
                    "Incompatible types of the transformation parameter "
                    f"'{tr_name}' ({tr_ann}) and the node '{node_name}' ({node_ann})"
                )
            if not tr_ann.is_array():
                continue
            if node_name == comp_connector:
                if ntr.output and not node_ann.output:
                    raise ValueError("'" + node_name + "' is not an output node")
                if not ntr.output and not node_ann.input:
                    raise ValueError("'" + node_name + "' is not an input node")
            elif ntr.output and node_ann.output:
                raise ValueError(
                    f"Cannot connect transformation parameter '{tr_name}' to an existing output node '{node_name}'"
                )
        self._connect(ntr)

    def reconnect(self, other: TransformationTree, translator: Callable[[str], str] | None = None) -> None:
        for ntr in other.connections():
            t = ntr.translate_node_names(translator) if translator is not None else ntr
            if t.connector_node_name not in self.leaf_parameters:
                continue
            ann = self.leaf_parameters[t.connector_node_name].annotation
            if (t.output and ann.output) or (not t.output and ann.input):
                self._connect(t)

    def connections(self) -> Iterator[NodeTransformation]:
        for name in self.get_subtree_names(leaves_only=False):
            yield from self.nodes[name].get_connections()

    def has_connections(self, names: Sequence[str] | None = None) -> bool:
        names = self.root_names if names is None else names
        return any(self.nodes[n].get_connections() for n in names if n in self.nodes)

    def translate(self, translator: Callable[[str], str]) -> TransformationTree:
        new = TransformationTree([p.rename(translator(p.name)) for p in self.get_root_parameters()])
        new.reconnect(self, translator=translator)
        return new

    def get_subtree(self, parameters: Iterable[Parameter]) -> TransformationTree:
        params = []
        for p in parameters:
            if p.name in self.root_parameters:
                if self.root_parameters[p.name].annotation != p.annotation:
                    raise ValueError(f"Annotation mismatch for the parameter `{p.name}`")
                params.append(self.root_parameters[p.name])
            else:
                params.append(p)
        new = TransformationTree(params)
        new.reconnect(self)
        return new

    # -- code generation ---------------------------------------------------
    def get_kernel_declaration(self, kernel_name: str, *, skip_constants: bool = False) -> tuple[str, list[str]]:
        leaves = self.get_leaf_parameters()
        if skip_constants:
            leaves = [p for p in leaves if not p.annotation.constant]
        decl = f'extern "C" __global__ void {kernel_name}({", ".join(param_cname(p, qualified=True) for p in leaves)})'
        return decl, [p.name for p in leaves]

    def get_kernel_argobjects(self) -> list[KernelParameter]:
        return [self._argobject(n, self.root_parameters[n].annotation) for n in self.root_names]

    def _argobject(self, name: str, annotation: Annotation) -> KernelParameter:
        if annotation.is_scalar():
            return KernelParameter(name, annotation)
        load_idx, load_same, load_comb = self._connection_modules(name, annotation, output=False)
        store_idx, store_same, store_comb = self._connection_modules(name, annotation, output=True)
        return KernelParameter(
            name,
            annotation,
            load_idx=load_idx,
            store_idx=store_idx,
            load_same=load_same,
            store_same=store_same,
            load_combined_idx=load_comb,
            store_combined_idx=store_comb,
        )

    def _connection_modules(self, name: str, annotation: Annotation, *, output: bool):
        node = self.nodes[name]
        param = Parameter(name, annotation)
        ntr = node.output_ntr if output else node.input_ntr
        subtree = self.get_leaf_parameters([name])
        m_idx = module_leaf_macro(param, output=output) if ntr is None else self._trf_module(annotation, ntr)
        m_same = module_same_indices(param, m_idx, output=output)
        m_comb = _CombinedFactory(param, subtree, m_idx, output=output)
        return m_idx, m_same, m_comb

    def _trf_module(self, annotation: Annotation, ntr: NodeTransformation) -> Module:
        param = Parameter(ntr.connector_node_name, annotation)
        args: list[Any] = [Indices(annotation.shape)]
        for tr_param in ntr.trf.signature.parameters.values():
            node_name = ntr.node_from_tr[tr_param.name]
            if node_name == ntr.connector_node_name:
                if ntr.output:
                    args.append(KernelParameter(param.name, annotation, load_same=VALUE_NAME))
                else:
                    args.append(KernelParameter(param.name, annotation, store_same=VALUE_NAME + " ="))
            else:
                args.append(self._argobject(node_name, tr_param.annotation))
        subtree = self.get_leaf_parameters([ntr.connector_node_name])
        return module_transformation(ntr.output, param, subtree, ntr.trf.snippet, args)


class Indices:
    """Names of the index variables available to a transformation snippet."""

    def __init__(self, shape: Sequence[int]):
        self._names = index_cnames(shape)

    def __getitem__(self, dim: int) -> str:
        return self._names[dim]

    def __len__(self) -> int:
        return len(self._names)

    def __iter__(self) -> Iterator[str]:
        return iter(self._names)

    def all(self) -> str:
        return ", ".join(self._names)


class KernelParameter(Type):
    """
    A kernel argument as seen from a template: ``str()`` is its C name; arrays
    carry the load/store modules (see the module docstring).
    """

    def __init__(self, name: str, type_: Type, **modules: Any):
        super().__init__(type_)
        self.name = name
        self._leaf_name = leaf_name(name)
        for k, v in modules.items():
            if v is not None:
                setattr(self, k, v)

    def __repr__(self) -> str:
        return f"KernelParameter({self.name})"

    def __str__(self) -> str:
        return self._leaf_name


# ---------------------------------------------------------------------------
# generated modules
# ---------------------------------------------------------------------------
def param_cname(param: Parameter, *, qualified: bool = False) -> str:
    name = leaf_name(param.name)
    if not qualified:
        return name
    ctype = param.annotation.ctype
    if param.annotation.is_array():
        # no __restrict__: callers may pass one buffer for several parameters
        const = "const " if param.annotation.constant or (param.annotation.role == "i") else ""
        return f"{const}{ctype}* {name}"
    return f"{ctype} {name}"


def _flat_index_expr(param: Parameter) -> str:
    t = param.annotation
    item = t.dtype.itemsize
    if t.offset % item:
        raise ValueError("Offset must be a multiple of the itemsize")
    if any(s % item for s in t.strides):
        raise ValueError(f"Some of the strides {t.strides} are not multiples of the itemsize {item}")
    names = index_cnames(t.shape)
    terms = [f"(long long)({n}) * ({s // item}LL)" for n, s in zip(names, t.strides)]
    return " + ".join(terms + [f"({t.offset // item}LL)"])


def module_leaf_macro(param: Parameter, *, output: bool) -> Module:
    names = index_cnames(param.annotation.shape)
    lname = leaf_name(param.name)
    expr = _flat_index_expr(param)

    def render(prefix: str) -> str:
        if output:
            return f"#define {prefix}({', '.join(names + [VALUE_NAME])}) {lname}[{expr}] = ({VALUE_NAME})\n"
        return f"#define {prefix}({', '.join(names)}) ({lname}[{expr}])\n"

    return Module(render)


def module_same_indices(param: Parameter, m_idx: Module, *, output: bool) -> Module:
    names = index_cnames(param.annotation.shape)

    def render(prefix: str) -> str:
        if output:
            return f"#define {prefix}({VALUE_NAME}) {m_idx}({', '.join(names + [VALUE_NAME])})\n"
        return f"#define {prefix} {m_idx}({', '.join(names)})\n"

    return Module(render)


def module_transformation(
    output: bool, param: Parameter, subtree: list[Parameter], snippet: Snippet, args: list[Any]
) -> Module:
    names = index_cnames(param.annotation.shape)
    ctype = param.annotation.ctype

    def render(prefix: str) -> str:
        sig = [param_cname(p, qualified=True) for p in subtree] + [f"long long {n}" for n in names]
        if output:
            sig.append(f"{ctype} {VALUE_NAME}")
        body = snippet(*args)
        head = f"// {'output' if output else 'input'} transformation of \"{param.name}\"\n"
        if output:
            fn = f"__device__ __forceinline__ void {prefix}func({', '.join(sig)})\n{{\n{body}\n}}\n"
        else:
            fn = (
                f"__device__ __forceinline__ {ctype} {prefix}func({', '.join(sig)})\n{{\n"
                f"    {ctype} {VALUE_NAME};\n{body}\n    return {VALUE_NAME};\n}}\n"
            )
        call_args = [leaf_name(p.name) for p in subtree] + names + ([VALUE_NAME] if output else [])
        macro_args = names + ([VALUE_NAME] if output else [])
        return head + fn + f"#define {prefix}({', '.join(macro_args)}) {prefix}func({', '.join(call_args)})\n"

    return Module(render)


class _CombinedFactory:
    """``load_combined_idx(slices)`` / ``store_combined_idx(slices)``: index by
    groups of dimensions (e.g. shape (2, 3, 4, 5, 6) with slices (2, 2, 1) is
    indexed as shape (6, 20, 6))."""

    def __init__(self, param: Parameter, subtree: list[Parameter], m_idx: Module, *, output: bool):
        self.param = param
        self.subtree = subtree
        self.m_idx = m_idx
        self.output = output
        self._cache: dict[tuple[int, ...], Module] = {}

    def __call__(self, slices: Sequence[int]) -> Module:
        slices = tuple(int(s) for s in slices)
        if slices not in self._cache:
            self._cache[slices] = module_combined(self.param, self.subtree, self.m_idx, slices, output=self.output)
        return self._cache[slices]


def module_combined(
    param: Parameter, subtree: list[Parameter], m_idx: Module, slices: Sequence[int], *, output: bool
) -> Module:
    shape = param.annotation.shape
    if sum(slices) != len(shape):
        raise ValueError(f"slices {tuple(slices)} do not cover shape {shape}")
    names = index_cnames(shape)
    ctype = param.annotation.ctype
    cnames = [f"_c_idx{k}" for k in range(len(slices))]

    def render(prefix: str) -> str:
        sig = [param_cname(p, qualified=True) for p in subtree] + [f"long long {c}" for c in cnames]
        if output:
            sig.append(f"{ctype} {VALUE_NAME}")
        lines = []
        start = 0
        for k, n in enumerate(slices):
            dims = list(range(start, start + n))
            for d in dims:
                stride = 1
                for e in range(d + 1, start + n):
                    stride *= shape[e]
                lines.append(f"    long long {names[d]} = {cnames[k]} / {stride}LL;")
                if d != dims[-1]:
                    lines.append(f"    {cnames[k]} -= {names[d]} * {stride}LL;")
            start += n
        body = "\n".join(lines)
        call = f"{m_idx}({', '.join(names + ([VALUE_NAME] if output else []))})"
        ret = "void" if output else ctype
        stmt = f"    {call};" if output else f"    return {call};"
        fn = f"__device__ __forceinline__ {ret} {prefix}func({', '.join(sig)})\n{{\n{body}\n{stmt}\n}}\n"
        call_args = [leaf_name(p.name) for p in subtree] + cnames + ([VALUE_NAME] if output else [])
        macro_args = cnames + ([VALUE_NAME] if output else [])
        return fn + f"#define {prefix}({', '.join(macro_args)}) {prefix}func({', '.join(call_args)})\n"

    return Module(render)
