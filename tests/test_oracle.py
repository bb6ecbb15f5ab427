"""Pins the CPU oracle (oracle/fft_oracle.py) before it is trusted as the checker.

Each oracle function is checked against something it does not compute itself:
closed forms of the DFT, the O(N^2) definition, numpy.fft, the reference's
documented radix plans (reikna/fft/fft.py:30-49, 67-79, 92-110), and the
committed golden vectors (tests/golden/fft_golden.npz).  Several checks plant a
plausible mistake (wrong sign, missing conjugate, transposed index) in a copy
of the computation and require the check to catch it.
"""

import math
from pathlib import Path

import numpy
import pytest

import oracle
from oracle import fft_oracle as fo

GOLDEN = Path(__file__).resolve().parent / "golden" / "fft_golden.npz"


# -- closed forms ---------------------------------------------------------------
@pytest.mark.parametrize("n", [1, 2, 7, 16, 60, 97, 1000])
def test_delta_and_constant(n):
    d = numpy.zeros(n, complex)
    d[0] = 1
    assert numpy.allclose(oracle.fftn(d), numpy.ones(n))
    c = numpy.ones(n, complex)
    expect = numpy.zeros(n, complex)
    expect[0] = n
    assert numpy.allclose(oracle.fftn(c), expect)


@pytest.mark.parametrize("n,k0", [(8, 3), (60, 7), (97, 96), (1000, 1)])
def test_single_frequency(n, k0):
    x = numpy.exp(2j * numpy.pi * k0 * numpy.arange(n) / n)
    y = oracle.fftn(x)
    expect = numpy.zeros(n, complex)
    expect[k0] = n
    assert numpy.allclose(y, expect, atol=1e-9 * n)
    # a sign mistake in the kernel puts the peak at n - k0 instead
    assert not numpy.allclose(numpy.fft.ifft(x) * n, expect, atol=1e-6) or k0 == n - k0


@pytest.mark.parametrize("n", [5, 64, 97, 360])
def test_linearity_parseval_shift_roundtrip(n):
    rng = numpy.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    y = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    a, b = 0.3 - 1.2j, 2.5 + 0.1j
    assert numpy.allclose(oracle.fftn(a * x + b * y), a * oracle.fftn(x) + b * oracle.fftn(y))
    X = oracle.fftn(x)
    assert math.isclose(numpy.vdot(x, x).real * n, numpy.vdot(X, X).real, rel_tol=1e-12)
    s = 3
    k = numpy.arange(n)
    assert numpy.allclose(oracle.fftn(numpy.roll(x, s)), X * numpy.exp(-2j * numpy.pi * k * s / n))
    assert numpy.allclose(oracle.ifftn(X), x)
    r = rng.standard_normal(n)
    R = oracle.fftn(r)
    assert numpy.allclose(R[1:], numpy.conj(R[1:][::-1]))


@pytest.mark.parametrize("n", [2, 3, 5, 8, 13, 31, 64, 97])
def test_direct_dft_matches_fft(n):
    rng = numpy.random.default_rng(n + 1)
    x = rng.standard_normal((3, n)) + 1j * rng.standard_normal((3, n))
    assert numpy.allclose(fo.dft_direct(x), oracle.fftn(x, (1,)), atol=1e-10 * n)
    assert numpy.allclose(fo.dft_direct(x, inverse=True), oracle.ifftn(x, (1,)), atol=1e-10)


def test_multi_axis_separable():
    rng = numpy.random.default_rng(3)
    x = rng.standard_normal((4, 6, 5)) + 1j * rng.standard_normal((4, 6, 5))
    y = fo.dft_direct(x)
    y = numpy.moveaxis(fo.dft_direct(numpy.moveaxis(y, 1, 2)), 2, 1)
    y = numpy.moveaxis(fo.dft_direct(numpy.moveaxis(y, 0, 2)), 2, 0)
    assert numpy.allclose(oracle.fftn(x), y)
    # a transposed-axis mistake changes the answer
    assert not numpy.allclose(oracle.fftn(x, (0, 2)), y)


# -- Bluestein as the reference computes it ----------------------------------
@pytest.mark.parametrize("n", [1, 2, 7, 60, 97, 997, 1000])
def test_bluestein_matches_dft(n):
    rng = numpy.random.default_rng(n)
    x = rng.standard_normal((2, n)) + 1j * rng.standard_normal((2, n))
    assert numpy.allclose(fo.bluestein_fft(x), oracle.fftn(x, (1,)), atol=1e-9 * max(n, 1))
    assert numpy.allclose(fo.bluestein_fft(x, inverse=True), oracle.ifftn(x, (1,)), atol=1e-9)


def test_chirp_integer_phase():
    """exp(-i pi (m^2 mod 2n)/n) equals exp(-i pi m^2 / n) (checked where the
    direct FP64 phase is still exact) and stays accurate where it is not."""
    n = 997
    m = numpy.arange(n)
    direct = numpy.exp(-1j * numpy.pi * (m.astype(float) ** 2) / n)
    assert numpy.allclose(oracle.chirp(n), direct, atol=1e-9)
    n = 3_000_017   # m^2 beyond 2^43: the direct phase loses digits, the integer one does not
    m = numpy.array([n - 1, n - 2, 2_999_999])
    exact = [complex(math.cos(math.pi * ((int(k) ** 2) % (2 * n)) / n), -math.sin(math.pi * ((int(k) ** 2) % (2 * n)) / n)) for k in m]
    assert numpy.allclose(fo.chirp(n)[m], exact, atol=1e-12)


@pytest.mark.parametrize("n", [7, 60, 255, 1000])
def test_kweights_definition(n):
    """Row 0 of kweights (fft.py:216-236) is the M-point spectrum of the
    wrapped conjugate chirp: used as Bluestein's kernel it reproduces the DFT."""
    m = 2 * fo.bounding_power_of_2(n)
    kw = fo.kweights(n, m)
    c = fo.chirp(n)
    rng = numpy.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    a = numpy.zeros(m, complex)
    a[:n] = x * c
    y = numpy.fft.ifft(numpy.fft.fft(a) * kw[0])[:n] * c
    assert numpy.allclose(y, numpy.fft.fft(x), atol=1e-9 * n)
    # row 1 serves the inverse direction (fft.mako picks kweights[(1 - dir) / 2]):
    # with the conjugate chirp it yields the normalised inverse DFT
    ci = numpy.conj(c)
    a_inv = numpy.zeros(m, complex)
    a_inv[:n] = x * ci
    y_inv = numpy.fft.ifft(numpy.fft.fft(a_inv) * kw[1])[:n] * ci
    assert numpy.allclose(y_inv, numpy.fft.ifft(x), atol=1e-12 * n)
    # dropping the conjugate breaks it
    y_bad = numpy.fft.ifft(numpy.fft.fft(a) * numpy.conj(kw[0]))[:n] * c
    assert not numpy.allclose(y_bad, numpy.fft.fft(x), atol=1e-6)


# -- the reference's radix planners ------------------------------------------------
def test_radix_array_documented_values():
    # explicit table and docstring examples of reikna/fft/fft.py:30-49, 67-79
    assert fo.radix_array(1024) == [16, 16, 4]
    assert fo.radix_array(2048) == [8, 8, 8, 4]
    assert fo.radix_array(32) == [8, 4]
    assert fo.radix_array(256) == [4, 4, 4, 4]
    for p in range(1, 20):
        assert math.prod(fo.radix_array(2**p)) == 2**p
        assert math.prod(fo.radix_array(2**p, use_max_radix=True)) == 2**p
    with pytest.raises(ValueError):
        fo.radix_array(12)


def test_global_radix_info_plan_counts():
    # pass counts reikna's global radix planner gives the benchmark sizes (fft.py:89-146)
    r, r1, r2 = fo.global_radix_info(2**20)
    assert r == [128, 128, 64] and len(r) == 3
    assert fo.global_radix_info(256)[0] == [128, 2]
    for p in range(1, 24):
        r, r1, r2 = fo.global_radix_info(2**p)
        assert math.prod(r) == 2**p
        assert all(b <= a <= 16 and a * b == c for a, b, c in zip(r1, r2, r))


def test_flops_convention():
    assert oracle.fft_flops((64, 2**20), (1,)) == 5 * 64 * 2**20 * 20
    assert oracle.fft_flops((256, 256, 256), (0, 1, 2)) == 5 * 256**3 * 24
    assert math.isclose(oracle.fft_flops((4096, 997), (1,)), 5 * 4096 * 997 * math.log2(997))


def test_tolerance_and_metric():
    assert math.isclose(oracle.tolerance((64, 2**20), (1,)), 4e-6)
    x = numpy.ones(10)
    assert oracle.rel_l2(x * (1 + 1e-7), x) == pytest.approx(1e-7)


# -- golden vectors --------------------------------------------------------------
def test_golden_vectors():
    g = numpy.load(GOLDEN)
    names = sorted({k.rsplit("_", 1)[0] for k in g.files})
    assert len(names) >= 8
    for name in names:
        x, axes = g[f"{name}_in"], tuple(int(a) for a in g[f"{name}_axes"])
        assert x.dtype == numpy.complex64
        assert numpy.allclose(oracle.fftn(x, axes), g[f"{name}_fwd"], rtol=1e-12, atol=1e-12)
        assert numpy.allclose(oracle.ifftn(x, axes), g[f"{name}_inv"], rtol=1e-12, atol=1e-12)
        if len(axes) == 1 and x.shape[-1] <= 1024:
            assert numpy.allclose(fo.dft_direct(x), g[f"{name}_fwd"], atol=1e-9 * x.shape[-1])
        if len(axes) == 1:
            assert numpy.allclose(fo.bluestein_fft(x), g[f"{name}_fwd"], atol=1e-9 * x.shape[-1])


def test_threaded_path_matches_numpy():
    """The workers= path (scipy.fft's pocketfft on several threads) gives the
    numpy.fft result to FP64 rounding, forward and inverse, over 1 and 2 axes;
    and it is a DFT: a planted single tone lands on its bin."""
    x = oracle.test_array((6, 96, 128), seed=3)
    for axes in [(2,), (1, 2)]:
        a = oracle.fftn(x, axes)
        b = oracle.fftn(x, axes, workers=4)
        assert numpy.abs(a - b).max() <= 1e-12 * numpy.abs(a).max()
        ai = oracle.ifftn(x, axes)
        bi = oracle.ifftn(x, axes, workers=4)
        assert numpy.abs(ai - bi).max() <= 1e-12 * numpy.abs(ai).max()
    n, k0 = 240, 37
    tone = numpy.exp(2j * numpy.pi * k0 * numpy.arange(n) / n)
    y = oracle.fftn(tone, (0,), workers=2)
    assert abs(y[k0] - n) < 1e-9 and numpy.abs(numpy.delete(y, k0)).max() < 1e-9
