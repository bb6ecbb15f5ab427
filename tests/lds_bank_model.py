"""LDS bank model of the exchanges of one FFT pass tile (fft_device.hpp).

The row pitch of a tile's LDS (``lds_pitch_for`` in fft_device.hpp, mirrored
by ``tile_pitch`` in build.py) was chosen with this model: it replays, for
every wave of a tile of C instances of length L, the LDS address of every
lane of every exchange instruction of ``fft_pass_tile`` (staged or flat
loads, the exchanges between Stockham stages, staged or flat stores) under
either compute mapping, and prices each instruction by its worst bank:

* 64-bit accesses: two half-waves, each over 32 slots of two banks;
* 32-bit accesses (tiles split into re/im halves): 64 lanes over 64 banks.

``python tests/lds_bank_model.py`` prints, for every catalogue tile
whose pitch the rule changes, the model's cycles per ideal cycle at the base
(odd) pitch and at the rule's pitch, averaged over the load/store modes and
both mappings.
"""

from __future__ import annotations

import sys
from pathlib import Path

if __package__ in (None, ""):   # run as a script: the repository root holds the package
    sys.path.insert(0, str(Path(__file__).resolve().parent.parent))

from reikna_amd.csrc.build import catalog_entries, lds_pitch, radix_list, tile_pitch  # noqa: E402

# (column mapping, load walk, store walk) combinations the planner produces
MODES = [
    (True, "direct", "direct"), (True, "direct", "flat"), (True, "flat", "direct"),
    (True, "staged", "direct"), (True, "direct", "staged"),
    (False, "direct", "direct"), (False, "flat", "flat"), (False, "flat", "direct"), (False, "direct", "flat"),
]


def pad(p: int) -> int:
    return p + (p >> 4)


class Stages:
    """Stockham stages of a length-L instance (Plan<L> in fft_device.hpp)."""

    def __init__(self, length: int):
        self.L = length
        self.radices = radix_list(length)
        self.T = length // self.radices[0]
        self.ns, self.nb, self.q = [], [], []
        ns = 1
        for r in self.radices:
            self.ns.append(ns)
            self.nb.append(length // r)
            self.q.append((length // r + self.T - 1) // self.T)
            ns *= r

    def valid(self, s: int, t: int, q: int) -> bool:
        return self.nb[s] % self.T == 0 or t + q * self.T < self.nb[s]

    def in_pos(self, s: int, t: int, i: int) -> int:
        q, r = divmod(i, self.radices[s])
        return t + q * self.T + r * self.nb[s]

    def out_pos(self, s: int, t: int, i: int) -> int:
        q, r = divmod(i, self.radices[s])
        j = t + q * self.T
        ns = self.ns[s]
        return (j // ns) * ns * self.radices[s] + j % ns + r * ns


def instruction_cycles(addrs: list[int], split: bool) -> int:
    if not addrs:
        return 0
    if split:
        count: dict[int, int] = {}
        for a in addrs:
            count[a % 64] = count.get(a % 64, 0) + 1
        return max(count.values())
    cycles = 0
    for half in (addrs[:32], addrs[32:]):
        if half:
            count = {}
            for a in half:
                count[a % 32] = count.get(a % 32, 0) + 1
            cycles += max(count.values())
    return cycles


def tile_cycles(length: int, c: int, pitch: int, colmap: bool, load: str, store: str) -> tuple[int, int]:
    """(modelled cycles, ideal cycles) of all LDS exchange instructions of one tile."""
    st = Stages(length)
    T, nt, r0 = st.T, c * st.T, st.radices[0]
    last = len(st.radices) - 1
    split = c * lds_pitch(length) * 8 > 80 * 1024
    total = ideal = 0

    def lanes(w):
        for tid in range(64 * w, min(64 * w + 64, nt)):
            if colmap:
                yield tid, tid % c, tid // c, tid // T, tid % T
            else:
                yield tid, tid // T, tid % T, tid % c, tid // c

    def flat(tid, ci, t, cs, ps, it):
        e = tid + it * nt
        return (e // length) * pitch + pad(e % length)

    def own(tid, ci, t, cs, ps, i):
        return ci * pitch + pad(st.in_pos(0, t, i))

    def staged(tid, ci, t, cs, ps, it):
        return cs * pitch + pad(ps + it * T)

    def stage_out(s):
        return lambda tid, ci, t, cs, ps, i: (
            ci * pitch + pad(st.out_pos(s, t, i)) if st.valid(s, t, i // st.radices[s]) else None)

    def stage_in(s):
        return lambda tid, ci, t, cs, ps, i: (
            ci * pitch + pad(st.in_pos(s, t, i)) if st.valid(s, t, i // st.radices[s]) else None)

    for w in range((nt + 63) // 64):
        ln = list(lanes(w))

        def run(fn, n):
            nonlocal total, ideal
            for i in range(n):
                addrs = [a for a in (fn(*x, i) for x in ln) if a is not None]
                total += instruction_cycles(addrs, split)
                ideal += 1 if split else 2

        if load == "flat":
            run(flat, r0)
            run(own, r0)
        elif load == "staged":
            run(staged, r0)
            run(own, r0)
        for s in range(last):
            run(stage_out(s), st.q[s] * st.radices[s])
            run(stage_in(s + 1), st.q[s + 1] * st.radices[s + 1])
        n_last = st.q[last] * st.radices[last]
        if store == "flat":
            run(stage_out(last), n_last)
            run(flat, r0)
        elif store == "staged":
            run(stage_out(last), n_last)
            run(staged, r0)
    return total, ideal


def score(length: int, c: int, pitch: int) -> float:
    """Modelled cycles per ideal cycle, averaged over MODES."""
    s = 0.0
    for colmap, load, store in MODES:
        t, i = tile_cycles(length, c, pitch, colmap, load, store)
        s += t / max(i, 1)
    return s / len(MODES)


def main() -> None:
    seen = set()
    for length, c, _ in catalog_entries():
        if (length, c) in seen or length > 4096:
            continue
        seen.add((length, c))
        base, rule = lds_pitch(length), tile_pitch(length, c)
        if base != rule:
            print(f"L={length:5d} C={c:3d}: pitch {base} -> {score(length, c, base):.2f}, "
                  f"pitch {rule} -> {score(length, c, rule):.2f}", flush=True)


if __name__ == "__main__":
    main()
