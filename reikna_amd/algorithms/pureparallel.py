"""PureParallel: an elementwise computation over an index space.

The public interface is reikna's (reikna/algorithms/pureparallel.py:19-151:
the constructor's ``parameters, code, guiding_array, render_kwds``, the
``from_trf`` factory and its two ValueError texts); the FFT uses it for
trivial problems (every transformed axis of length 1) so that attached
transformations still run (reikna/fft/fft.py:647-662).

The kernel is this project's own: one thread per point of the index space,
launched as a flat 1-D grid of 256-thread blocks.  A thread takes its flat
position and decodes it into the multi-index with the extents compiled in as
literals (the compiler turns each division by a constant into a multiply and
shift), innermost dimension first, so neighbouring lanes get neighbouring
innermost indices and the snippet's accesses to C-ordered arrays coalesce.
"""

from __future__ import annotations

from collections.abc import Callable, Iterable, Mapping
from typing import Any

from .. import helpers
from ..core import Computation, ComputationPlan, Indices, KernelArguments, Parameter, Transformation
from ..core.transformation import TransformationParameter
from ..runtime import DeviceParameters
from ..template import DefTemplate, Snippet

_BLOCK = 256
_KERNEL_NAME = "rk_elementwise"


def _decode_lines(names: list[str], shape: tuple[int, ...], flat: str) -> list[str]:
    """C statements declaring the index variables `names` from the flat
    row-major position `flat` over `shape` (extents as literals)."""
    if not shape:
        return []
    out = [f"    unsigned long long rk_rest = {flat};"]
    for k in range(len(shape) - 1, 0, -1):
        n = int(shape[k])
        out.append(f"    const VSIZE_T {names[k]} = (VSIZE_T)(rk_rest % {n}ull);")
        out.append(f"    rk_rest /= {n}ull;")
    out.append(f"    const VSIZE_T {names[0]} = (VSIZE_T)rk_rest;")
    return out


class PureParallel(Computation):
    """
    An elementwise computation: ``code`` runs once for every point of the
    index space.

    :param parameters: the computation's parameters (a list of Parameter).
    :param code: a Snippet taking ``(idxs, *parameters)``, or the source of
        such a snippet's body (``idxs`` is the list of index variable names).
    :param guiding_array: the index space: a shape, or the name of an array
        parameter whose shape it takes (default: the first parameter).
    :param render_kwds: extra names for rendering ``code`` when it is source.
    """

    def __init__(
        self,
        parameters: Iterable[Parameter],
        code: Snippet | str,
        guiding_array: str | tuple[int, ...] | None = None,
        render_kwds: Mapping[str, Any] | None = None,
    ):
        Computation.__init__(self, parameters)
        names = list(self.signature.parameters)
        self._param_names = names
        if isinstance(guiding_array, str) or guiding_array is None:
            ann = self.signature.parameters[guiding_array or names[0]].annotation
            self._space = tuple(ann.shape)
        else:
            self._space = tuple(int(n) for n in guiding_array)
        self._body = code if isinstance(code, Snippet) else Snippet(
            DefTemplate.from_string("rk_elementwise_body", ["idxs", *names], code), dict(render_kwds or {})
        )

    @classmethod
    def from_trf(cls, trf: Transformation, guiding_array: str | TransformationParameter | None = None) -> PureParallel:
        """The elementwise computation that runs ``trf`` over the shape of one
        of its array parameters (default: its first parameter)."""
        if isinstance(guiding_array, TransformationParameter):
            if not guiding_array.belongs_to(trf):
                raise ValueError("The transformation parameter must belong to the provided transformation")
            guiding_array = str(guiding_array)
        params = trf.signature.parameters
        name = guiding_array if guiding_array is not None else next(iter(params))
        if not params[name].annotation.is_array():
            raise ValueError("The parameter serving as a guiding array cannot be a scalar")
        return cls(list(params.values()), trf.snippet, guiding_array=name)

    def _build_plan(
        self, plan_factory: Callable[[], ComputationPlan], _device_params: DeviceParameters, args: KernelArguments
    ) -> ComputationPlan:
        plan = plan_factory()
        idxs = Indices(self._space)
        total = helpers.product(self._space) if self._space else 1
        body = [
            "${kernel_declaration}",
            "{",
            "    const unsigned long long rk_flat = ((unsigned long long)blockIdx.y * gridDim.x + blockIdx.x) * %dull"
            " + threadIdx.x;" % _BLOCK,
            f"    if (rk_flat >= {total}ull) return;",
            *_decode_lines(list(idxs), self._space, "rk_flat"),
            "    ${body(idxs, %s)}" % ", ".join(self._param_names),
            "}",
        ]
        template = DefTemplate.from_string(
            _KERNEL_NAME, ["kernel_declaration", *self._param_names], "\n".join(body) + "\n"
        )
        # a flat grid: `total` threads in blocks of _BLOCK
        blocks = (total + _BLOCK - 1) // _BLOCK
        plan.kernel_call(
            template,
            args.all(),
            kernel_name=_KERNEL_NAME,
            global_size=(blocks * _BLOCK,),
            local_size=(_BLOCK,),
            render_kwds=dict(idxs=idxs, body=self._body),
        )
        return plan
