"""PureParallel: an elementwise computation generated from a snippet
(reikna/algorithms/pureparallel.py:19-151).  The FFT uses it for trivial
problems (every transformed axis of length 1) so that attached
transformations still run (reikna/fft/fft.py:647-662)."""

from __future__ import annotations

from collections.abc import Callable, Iterable, Mapping
from typing import Any

from .. import helpers
from ..core import Computation, ComputationPlan, Indices, KernelArguments, Parameter, Transformation
from ..core.transformation import TransformationParameter
from ..runtime import DeviceParameters
from ..template import DefTemplate, Snippet


class PureParallel(Computation):
    """
    An elementwise computation.

    :param parameters: list of Parameter objects.
    :param code: a Snippet taking ``(idxs, *parameters)``, or its template source.
    :param guiding_array: the shape of the index space, or the name of a parameter.
    :param render_kwds: extra names for the template.
    """

    def __init__(
        self,
        parameters: Iterable[Parameter],
        code: Snippet | str,
        guiding_array: str | tuple[int, ...] | None = None,
        render_kwds: Mapping[str, Any] | None = None,
    ):
        Computation.__init__(self, parameters)
        self._root_parameters = list(self.signature.parameters.keys())
        if isinstance(code, Snippet):
            self._snippet = code
        else:
            self._snippet = Snippet(
                DefTemplate.from_string("pure_parallel_inner", ["idxs", *self._root_parameters], code),
                render_kwds or {},
            )
        if guiding_array is None:
            guiding_array = self._root_parameters[0]
        if isinstance(guiding_array, str):
            self._guiding_shape = self.signature.parameters[guiding_array].annotation.shape
        else:
            self._guiding_shape = tuple(guiding_array)

    @classmethod
    def from_trf(cls, trf: Transformation, guiding_array: str | TransformationParameter | None = None) -> PureParallel:
        """A PureParallel running ``trf`` over the shape of one of its arrays."""
        if guiding_array is None:
            guiding_array = next(iter(trf.signature.parameters.keys()))
        if isinstance(guiding_array, TransformationParameter):
            if not guiding_array.belongs_to(trf):
                raise ValueError("The transformation parameter must belong to the provided transformation")
            guiding_array = str(guiding_array)
        param = trf.signature.parameters[guiding_array]
        if not param.annotation.is_array():
            raise ValueError("The parameter serving as a guiding array cannot be a scalar")
        return cls(list(trf.signature.parameters.values()), trf.snippet, guiding_array=param.name)

    def _build_plan(
        self, plan_factory: Callable[[], ComputationPlan], _device_params: DeviceParameters, args: KernelArguments
    ) -> ComputationPlan:
        plan = plan_factory()
        idxs = Indices(self._guiding_shape)
        arglist = ", ".join(self._root_parameters)
        lines = [
            "${kernel_declaration}",
            "{",
            "    if (${static.skip}()) return;",
        ]
        lines += [f"    VSIZE_T {idx} = ${{static.global_id}}({i});" for i, idx in enumerate(idxs)]
        lines += ["    ${snippet(idxs, " + arglist + ")}", "}"]
        template = DefTemplate.from_string("pure_parallel", ["kernel_declaration", *self._root_parameters],
                                           "\n".join(lines) + "\n")
        plan.kernel_call(
            template,
            args.all(),
            kernel_name="kernel_pure_parallel",
            global_size=(1,) if len(self._guiding_shape) == 0 else self._guiding_shape,
            render_kwds=dict(idxs=idxs, product=helpers.product, snippet=self._snippet),
        )
        return plan
