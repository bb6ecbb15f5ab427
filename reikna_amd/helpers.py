"""Small integer helpers used by the planners (reikna/helpers.py:87-208)."""

from __future__ import annotations

import collections
import functools
import itertools
from collections.abc import Callable, Iterable, Sequence
from typing import Any


def product(seq: Iterable[int]) -> int:
    """Product of the elements of ``seq`` (1 for an empty sequence)."""
    return functools.reduce(lambda a, b: a * b, seq, 1)


def min_blocks(length: int, block: int) -> int:
    """Number of blocks of ``block`` elements needed to cover ``length``."""
    return (length - 1) // block + 1


def log2(num: int) -> int:
    """Integer base-2 logarithm, rounded down."""
    if num < 1:
        return 0
    return int(num).bit_length() - 1


def bounding_power_of_2(num: int) -> int:
    """Smallest ``2**m >= num``."""
    if num <= 1:
        return 1
    return 1 << (int(num) - 1).bit_length()


def factors(num: int, limit: int | None = None) -> list[tuple[int, int]]:
    """Pairs ``(f, num // f)`` for every divisor ``f <= limit`` of ``num``, ascending."""
    if limit is None or limit > num:
        limit = num
    return [(f, num // f) for f in range(1, limit + 1) if num % f == 0]


def wrap_in_tuple(seq_or_elem: Any) -> tuple:
    if seq_or_elem is None:
        return ()
    if isinstance(seq_or_elem, int):
        return (seq_or_elem,)
    return tuple(seq_or_elem)


def normalize_axes(ndim: int, axes: Any) -> tuple[int, ...]:
    """Non-negative axes (negative ones counted from the end); all axes if ``None``."""
    if axes is None:
        return tuple(range(ndim))
    axes = wrap_in_tuple(axes)
    axes = tuple(a if a >= 0 else ndim + a for a in axes)
    if any(a < 0 or a >= ndim for a in axes):
        raise IndexError("Array index out of range")
    return axes


def are_axes_innermost(ndim: int, axes: Sequence[int]) -> bool:
    return tuple(axes) == tuple(range(ndim - len(axes), ndim))


def make_axes_innermost(ndim: int, axes: Sequence[int]) -> tuple[tuple[int, ...], tuple[int, ...]]:
    """Permutation moving ``axes`` to the end (in order) and its inverse."""
    outer = [i for i in range(ndim) if i not in axes]
    to = tuple(outer + list(axes))
    back = [0] * ndim
    for i, a in enumerate(to):
        back[a] = i
    return to, tuple(back)


class Graph:
    """Undirected graph of temporary-buffer dependencies (reikna/helpers.py:29-84)."""

    def __init__(self, pairs: Iterable[tuple[Any, Any]] | None = None):
        self._nodes: dict[Any, set] = collections.defaultdict(set)
        if pairs:
            for a, b in pairs:
                self.add_edge(a, b)

    def add_edge(self, a: Any, b: Any) -> None:
        if a == b:
            raise ValueError("nodes must be distinct")
        self._nodes[a].add(b)
        self._nodes[b].add(a)

    def add_cluster(self, nodes: Iterable[Any]) -> None:
        for a, b in itertools.combinations(list(nodes), 2):
            self.add_edge(a, b)

    def __getitem__(self, node: Any) -> set:
        return self._nodes[node]

    def pairs(self) -> list[tuple[Any, Any]]:
        out = set()
        for a, ns in self._nodes.items():
            for b in ns:
                out.add((a, b) if str(a) < str(b) else (b, a))
        return sorted(out, key=str)

    def translate(self, fn: Callable[[Any], Any]) -> Graph:
        return Graph((fn(a), fn(b)) for a, b in self.pairs())
