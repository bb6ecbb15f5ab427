"""Writes tests/golden/fft_golden.npz: seeded complex64 inputs and their FP64
numpy.fft outputs (the reference's oracle, tests/test_fft/test_fft.py:163,169).

The reference holds no golden vectors for the FFT (its tests draw unseeded
random data, tests/helpers.py:39), so these fixtures are generated from the
oracle library; run `python tests/golden/make_golden.py` to regenerate them.
"""

from pathlib import Path

import numpy

CASES = [
    # name, shape, axes, seed
    ("n8", (3, 8), (1,), 101),
    ("n60", (4, 60), (1,), 102),
    ("n997", (2, 997), (1,), 103),
    ("n1000", (2, 1000), (1,), 104),
    ("n1024", (2, 1024), (1,), 105),
    ("n12288", (1, 12288), (1,), 106),
    ("d2_17x200", (17, 200), (0, 1), 107),
    ("d3_8x12x10", (8, 12, 10), (0, 1, 2), 108),
]


def make() -> dict:
    out = {}
    for name, shape, axes, seed in CASES:
        rng = numpy.random.default_rng(seed)
        x = (rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape)).astype(numpy.complex64)
        out[f"{name}_in"] = x
        out[f"{name}_axes"] = numpy.array(axes)
        out[f"{name}_fwd"] = numpy.fft.fftn(x.astype(numpy.complex128), axes=axes)
        out[f"{name}_inv"] = numpy.fft.ifftn(x.astype(numpy.complex128), axes=axes)
    return out


if __name__ == "__main__":
    path = Path(__file__).resolve().parent / "fft_golden.npz"
    numpy.savez_compressed(path, **make())
    print(f"wrote {path} ({path.stat().st_size} bytes)")
