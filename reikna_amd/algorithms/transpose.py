"""Transpose: permute the axes of an array (reikna/algorithms/transpose.py:100-244).

Same interface as reikna's ``Transpose(arr_t, output_arr_t=None, axes=None,
block_width_override=None)`` with signature ``(output, input)``.  A general
permutation is split into a shortest sequence of batched 2-D transposes
``(batch, height, width) -> (batch, width, height)`` over groups of adjacent
axes (the same decomposition reikna searches for, reikna/algorithms/transpose.py:
28-97; here a breadth-first search), each one kernel.

The kernel is written for gfx950: a work-group of 256 threads moves a
64 x 64 tile; each wave reads 64 consecutive elements of a row (one
coalesced access per row), the tile goes through LDS with a pitch of 65
elements (rows and columns fall in different banks), and each wave writes
64 consecutive elements of an output row.  Attached transformations run on
the global loads and stores as in reikna (``load_combined_idx`` /
``store_combined_idx`` over the (batch, height, width) groups).
"""

from __future__ import annotations

from collections import deque
from collections.abc import Callable, Iterable

import numpy

from .. import helpers
from ..core import Annotation, Computation, ComputationPlan, KernelArguments, Parameter
from ..runtime import ArrayMetadata, DeviceParameters
from ..template import DefTemplate

__all__ = ["Transpose", "transpose_operations"]

TILE = 64          # tile edge (elements)
ROWS = 4           # rows of the tile one pass of 256 threads covers


def _swap(axes: tuple[int, ...], b: int, c: int) -> tuple[int, ...]:
    """Exchange the adjacent axis groups [b, c) and [c, end)."""
    return axes[:b] + axes[c:] + axes[b:c]


def transpose_operations(n: int, target: tuple[int, ...]) -> list[tuple[int, int]]:
    """Shortest sequence of group swaps (b, c) turning (0..n-1) into `target`
    (axes already in place at the front are left out of the search)."""
    source = tuple(range(n))
    if source == target:
        return []
    keep = 0
    while keep < n and target[keep] == keep:
        keep += 1
    start = source[keep:]
    goal = tuple(target[keep:])
    m = len(start)
    moves = [(b, c) for b in range(m - 1) for c in range(b + 1, m)]
    prev: dict[tuple[int, ...], tuple[tuple[int, ...], tuple[int, int]] | None] = {start: None}
    queue = deque([start])
    while queue:
        node = queue.popleft()
        if node == goal:
            break
        for b, c in moves:
            nxt = _swap(node, b, c)
            if nxt not in prev:
                prev[nxt] = (node, (b, c))
                queue.append(nxt)
    ops = []
    node = goal
    while prev[node] is not None:
        parent, op = prev[node]
        ops.append((op[0] + keep, op[1] + keep))
        node = parent
    return ops[::-1]


def _transposes(shape: tuple[int, ...], axes: tuple[int, ...]) -> list[tuple[tuple, tuple, tuple]]:
    ops = transpose_operations(len(shape), axes)
    out = []
    for b, c in ops:
        out.append((shape[:b], shape[b:c], shape[c:]))
        shape = _swap(shape, b, c)
    return out


_KERNEL = r"""
${kernel_declaration}
{
    // one 64 x 64 tile of one matrix of the batch per work-group
    LOCAL_MEM ${output.ctype} tile[${tile} * (${tile} + 1)];
    const int tid = threadIdx.x;
    const int lx = tid % ${tile}, ly = tid / ${tile};
    const VSIZE_T g = (VSIZE_T)blockIdx.x + (VSIZE_T)blockIdx.y * gridDim.x;
    if (g >= ${blocks_total}) return;
    const VSIZE_T per_matrix = ${blocks_h} * ${blocks_w};
    const VSIZE_T b = g / per_matrix;
    const VSIZE_T r = g - b * per_matrix;
    const VSIZE_T h0 = (r / ${blocks_w}) * ${tile};
    const VSIZE_T w0 = (r % ${blocks_w}) * ${tile};
    // read: rows h0 + ly + k*${rows}, columns w0 + lx
    #pragma unroll
    for (int k = 0; k < ${tile} / ${rows}; ++k) {
        const VSIZE_T h = h0 + ly + k * ${rows};
        const VSIZE_T w = w0 + lx;
        if (h < ${height} && w < ${width})
            tile[(ly + k * ${rows}) * (${tile} + 1) + lx] = ${input.load_combined_idx(in_slices)}(b, h, w);
    }
    LOCAL_BARRIER;
    // write: output rows w0 + ly + k*${rows} (input columns), columns h0 + lx
    #pragma unroll
    for (int k = 0; k < ${tile} / ${rows}; ++k) {
        const VSIZE_T w = w0 + ly + k * ${rows};
        const VSIZE_T h = h0 + lx;
        if (h < ${height} && w < ${width})
            ${output.store_combined_idx(out_slices)}(b, w, h, tile[lx * (${tile} + 1) + ly + k * ${rows}]);
    }
}
"""


class Transpose(Computation):
    """
    Changes the order of axes of an array, as ``numpy.transpose``.

    :param arr_t: an array-like defining the initial array.
    :param output_arr_t: an array-like defining the output array; if ``None``
        its shape is the permuted shape and its dtype that of ``arr_t``.
    :param axes: the new axes order (reversed if ``None``).
    :param block_width_override: accepted for interface compatibility; the
        tile is fixed at 64 x 64 (one 64-lane wave per tile row).

    Signature: ``(output:o, input:i)``.
    """

    def __init__(
        self,
        arr_t,
        output_arr_t=None,
        axes: Iterable[int] | None = None,
        block_width_override: int | None = None,
    ):
        inp = arr_t.as_array_metadata() if hasattr(arr_t, "as_array_metadata") else ArrayMetadata.from_arraylike(arr_t)
        n = len(inp.shape)
        if axes is None:
            axes = tuple(reversed(range(n)))
        else:
            axes = tuple(axes)
            if set(axes) != set(range(n)) or len(axes) != n:
                raise ValueError("All transpose axes must be distinct")
        self._axes = axes
        self._transposes = _transposes(tuple(inp.shape), axes)
        out_shape = tuple(inp.shape[i] for i in axes)
        if output_arr_t is not None:
            out = (output_arr_t.as_array_metadata() if hasattr(output_arr_t, "as_array_metadata")
                   else ArrayMetadata.from_arraylike(output_arr_t))
        else:
            out = ArrayMetadata(out_shape, inp.dtype)
        if tuple(out.shape) != out_shape:
            raise ValueError(f"Expected output array shape: {out_shape}, got {tuple(out.shape)}")
        if numpy.dtype(out.dtype) != numpy.dtype(inp.dtype):
            raise ValueError("Input and output array must have the same dtype")
        Computation.__init__(
            self,
            [
                Parameter("output", Annotation(out, "o")),
                Parameter("input", Annotation(inp, "i")),
            ],
        )

    def _add_transpose(self, plan: ComputationPlan, mem_out, mem_in, batch_shape, height_shape, width_shape) -> None:
        height = helpers.product(height_shape)
        width = helpers.product(width_shape)
        batch = helpers.product(batch_shape)
        blocks_h = helpers.min_blocks(height, TILE)
        blocks_w = helpers.min_blocks(width, TILE)
        total = batch * blocks_h * blocks_w
        template = DefTemplate.from_string("transpose", ["kernel_declaration", "output", "input"], _KERNEL)
        plan.kernel_call(
            template,
            [mem_out, mem_in],
            kernel_name="kernel_transpose",
            global_size=(total * TILE * ROWS,),
            local_size=(TILE * ROWS,),
            render_kwds=dict(
                tile=TILE,
                rows=ROWS,
                height=height,
                width=width,
                blocks_h=blocks_h,
                blocks_w=blocks_w,
                blocks_total=total,
                in_slices=(len(batch_shape), len(height_shape), len(width_shape)),
                out_slices=(len(batch_shape), len(width_shape), len(height_shape)),
            ),
        )

    def _build_plan(
        self, plan_factory: Callable[[], ComputationPlan], _device_params: DeviceParameters, args: KernelArguments
    ) -> ComputationPlan:
        plan = plan_factory()
        output, inp = args.output, args.input
        if not self._transposes:
            from ..transformations import copy
            from .pureparallel import PureParallel

            trf = copy(inp, out_arr_t=output)
            plan.computation_call(PureParallel.from_trf(trf, trf.input), output, inp)
            return plan
        mem_out = output
        for i, (batch_shape, height_shape, width_shape) in enumerate(self._transposes):
            mem_in = inp if i == 0 else mem_out
            if i == len(self._transposes) - 1:
                mem_out = output
            else:
                mem_out = plan.temp_array(batch_shape + width_shape + height_shape, output.dtype)
            self._add_transpose(plan, mem_out, mem_in, batch_shape, height_shape, width_shape)
        return plan
