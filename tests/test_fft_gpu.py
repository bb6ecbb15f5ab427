"""GPU parity of the FFT computation against the FP64 oracle.

Cases follow the reference's tests/test_fft/test_fft.py (local :28-60, global
:62-98, sequence :100-113, trivial :173-193, typecheck :139-141) plus the
configurations of the benchmark (1024x1024, 64x2^20, 256^3, Bluestein 4096x{997,
1000, 12288}).  The bar is rel-L2 <= 2e-7 * log2 N against numpy.fft in FP64;
the reference's own per-element check (numpy.isclose(atol=2e-5, rtol=1e-3) on
U[0,1) inputs, test_fft.py:148) is applied to its own cases as well.
"""

import numpy
import pytest

import oracle
from reikna_amd import Array
from reikna_amd.fft import FFT
from reikna_amd.transformations import div_const, mul_param

pytestmark = pytest.mark.gpu


def ref_test_array(shape, seed):
    """The reference's get_test_array for complex64 (tests/helpers.py:27-51),
    re and im U[0, 1), but seeded."""
    rng = numpy.random.default_rng(seed)
    return (rng.uniform(0, 1, shape) + 1j * rng.uniform(0, 1, shape)).astype(numpy.complex64)


def run_fft(queue, data, axes, inverse=False, inplace=True, flags=0):
    dev = Array.from_host(queue, data)
    fft = FFT(dev, axes=axes, flags=flags)
    fftc = fft.compile(queue.device)
    if inplace:
        fftc(queue, dev, dev, int(inverse))
        res = dev.get(queue)
    else:
        out = Array.empty_like(queue.device, dev)
        fftc(queue, out, dev, int(inverse))
        res = out.get(queue)
    return res, fft


def check(queue, shape, axes, seed=0, reference_check=True, inplace=True, flags=0):
    data = ref_test_array(shape, seed)
    fwd, _ = run_fft(queue, data, axes, inverse=False, inplace=inplace, flags=flags)
    ref = oracle.fftn(data, axes)
    err = oracle.rel_l2(fwd, ref)
    tol = oracle.tolerance(shape, axes)
    assert err <= tol, f"forward rel-L2 {err:.3g} > {tol:.3g} for {shape} over {axes}"
    if reference_check:
        assert numpy.isclose(fwd, ref.astype(numpy.complex64), atol=2e-5, rtol=1e-3).all()
    inv, _ = run_fft(queue, data, axes, inverse=True, inplace=inplace, flags=flags)
    iref = oracle.ifftn(data, axes)
    ierr = oracle.rel_l2(inv, iref)
    assert ierr <= tol, f"inverse rel-L2 {ierr:.3g} > {tol:.3g} for {shape} over {axes}"
    return err, ierr


# -- reference: test_local (fft.mako insertGlobalLoadsAndTranspose paths) ------
def local_cases():
    out = []
    for p in (3, 7, 8, 9, 10, 12):
        size = 2**p
        batch = 2**20 // size
        for s in (size, size // 2 - 1):
            for b in (1, batch - 1, batch):
                out.append(((b, s), (1,)))
    return out


@pytest.mark.parametrize("shape_axes", local_cases(), ids=lambda v: f"{v[0][0]}x{v[0][1]}")
def test_local(queue, shape_axes):
    shape, axes = shape_axes
    check(queue, shape, axes, seed=hash(shape) % 1000)


# -- reference: test_global (strided axes, multi-pass) --------------------------
def global_cases():
    out = []
    mem_limit = 2**22
    for p in (3, 7, 9, 10):
        size = 2**p
        batch = mem_limit // size // 64
        for s in (size, size // 2 - 1):
            for ib in (2, 3, 63, 64):
                for ob in (1, batch - 1, batch):
                    out.append(((ob, s, ib), (1,)))
    big = 2**15
    batch = mem_limit // big // 8
    out += [((1, big, 1), (1,)), ((1, big, 2), (1,)), ((1, big, 3), (1,)), ((batch, big, 3), (1,)),
            ((batch - 1, big, 3), (1,))]
    return out


@pytest.mark.parametrize("shape_axes", global_cases(), ids=lambda v: "x".join(map(str, v[0])))
def test_global(queue, shape_axes):
    shape, axes = shape_axes
    check(queue, shape, axes, seed=sum(shape))


# -- reference: test_sequence (multi-axis, non-powers of two) -----------------
@pytest.mark.parametrize(
    "shape_axes",
    [((3, 255, 7), (2, 0)), ((17, 200, 131), (1, 0)), ((7, 1000, 11), (1, 2)), ((15, 900, 57), (2, 0, 1))],
    ids=lambda v: f"{v[0]}over{v[1]}",
)
def test_sequence(queue, shape_axes):
    shape, axes = shape_axes
    # the reference loosens rtol to 1e-2 here (test_fft.py:204-207); the
    # rel-L2 bound is not loosened
    data = ref_test_array(shape, 7)
    fwd, _ = run_fft(queue, data, axes)
    ref = oracle.fftn(data, axes)
    assert oracle.rel_l2(fwd, ref) <= oracle.tolerance(shape, axes)
    assert numpy.isclose(fwd, ref.astype(numpy.complex64), atol=2e-5, rtol=1e-2).all()


def test_typecheck():
    from reikna_amd import ArrayMetadata

    with pytest.raises(ValueError, match="FFT computation requires array of a complex dtype"):
        FFT(ArrayMetadata(100, numpy.float32))


def test_trivial(queue):
    """Problem size 1 over the axes: transformations must still run."""
    shape, axes, param = (128, 1, 1, 128), (1, 2), 4
    data = ref_test_array(shape, 3)
    dev = Array.from_host(queue, data)
    res = Array.empty_like(queue.device, dev)
    fft = FFT(dev, axes=axes)
    scale = mul_param(dev, numpy.int32)
    fft.parameter.input.connect(scale, scale.output, input_prime=scale.input, param=scale.param)
    fftc = fft.compile(queue.device)
    fftc(queue, res, dev, param)
    assert numpy.allclose(res.get(queue), data * param, rtol=1e-6)


# -- benchmark configurations and the three planner branches ----------------
@pytest.mark.parametrize(
    "shape,axes,dist",
    [
        ((1024, 1024), (1,), "uniform"),      # C1
        ((64, 2**20), (1,), "normal"),        # C2b (two passes)
        ((256, 256, 256), (0, 1, 2), "normal"),  # C2a (three strided/contiguous axes)
        ((4096, 1000), (1,), "uniform"),      # C3: 2^3 5^3, no Bluestein
        ((4096, 997), (1,), "uniform"),       # C3: prime, single-kernel Bluestein
        ((4096, 12288), (1,), "uniform"),     # C3: 3 * 2^12, single pass
    ],
    ids=lambda v: str(v),
)
def test_bench_configs(queue, shape, axes, dist):
    data = oracle.test_array(shape, seed=11, dist=dist)
    fwd, fft = run_fft(queue, data, axes, inplace=False)
    ref = oracle.fftn(data, axes)
    err = oracle.rel_l2(fwd, ref)
    assert err <= oracle.tolerance(shape, axes), (err, [p.L for p in fft.native_plan.passes])
    inv, _ = run_fft(queue, data, axes, inverse=True, inplace=True)
    assert oracle.rel_l2(inv, oracle.ifftn(data, axes)) <= oracle.tolerance(shape, axes)


def test_headline_closed_forms(queue):
    """The 64 x 2^20 transform (two chunked 1024-point passes) against what
    the mathematics fixes, without numpy.fft: an impulse at position p gives
    exp(-2 pi i p m / N) exactly (rows in both 32-row chunks), Parseval's
    identity, linearity, and the forward/inverse round trip."""
    shape, n = (64, 2**20), 2**20
    tol = oracle.tolerance((n,), (0,))
    rows = (0, 17, 31, 32, 50, 63)
    pos = {r: (r * 7919 * 13 + 5) % n for r in rows}
    imp = numpy.zeros(shape, numpy.complex64)
    for r, p in pos.items():
        imp[r, p] = 1.0
    fwd, fft = run_fft(queue, imp, (1,), inplace=False)
    assert [p.L for p in fft.native_plan.passes] == [1024, 1024] and fft.native_plan.chunk_rows == 32
    m = numpy.arange(n, dtype=numpy.float64)
    for r, p in pos.items():
        want = numpy.exp(-2j * numpy.pi * ((p * m) % n) / n)   # exact integer phase p*m mod N
        assert oracle.rel_l2(fwd[r], want) <= tol, (r, p)
    assert not fwd[1].any() and not fwd[62].any()   # rows without an impulse stay zero
    x = oracle.test_array(shape, seed=97)
    y = oracle.test_array(shape, seed=98)
    fx, _ = run_fft(queue, x, (1,), inplace=False)
    fy, _ = run_fft(queue, y, (1,), inplace=False)
    a, b = numpy.complex64(0.75 - 0.5j), numpy.complex64(-1.25 + 2.0j)
    fxy, _ = run_fft(queue, (a * x + b * y).astype(numpy.complex64), (1,), inplace=False)
    lin = a.astype(numpy.complex128) * fx.astype(numpy.complex128) + b.astype(numpy.complex128) * fy
    assert oracle.rel_l2(fxy, lin) <= 2 * tol
    ex = numpy.sum(numpy.abs(x.astype(numpy.complex128)) ** 2, axis=1) * n
    efx = numpy.sum(numpy.abs(fx.astype(numpy.complex128)) ** 2, axis=1)
    assert numpy.max(numpy.abs(efx / ex - 1.0)) <= 2 * tol
    back, _ = run_fft(queue, fx, (1,), inverse=True, inplace=True)
    assert oracle.rel_l2(back, x) <= 2 * tol


@pytest.mark.parametrize("n", [997, 1009, 4093, 4099, 8191, 12289, 65521, 131071])
def test_bluestein_primes(queue, n):
    """Primes: single-kernel Bluestein up to M = 8192, two chains of passes above."""
    batch = max(1, 2**20 // n)
    data = oracle.test_array((batch, n), seed=n)
    fwd, fft = run_fft(queue, data, (1,), inplace=False)
    ref = oracle.fftn(data, (1,))
    assert oracle.rel_l2(fwd, ref) <= oracle.tolerance((batch, n), (1,))
    inv, _ = run_fft(queue, data, (1,), inverse=True)
    assert oracle.rel_l2(inv, oracle.ifftn(data, (1,))) <= oracle.tolerance((batch, n), (1,))


@pytest.mark.parametrize("n", [60, 1000, 3 * 2**12, 5 * 2**11, 7 * 2**9, 2**14, 2**16, 3 * 2**14, 2**20])
def test_smooth_lengths(queue, n):
    batch = max(1, 2**21 // n)
    data = oracle.test_array((batch, n), seed=n + 1)
    fwd, _ = run_fft(queue, data, (1,), inplace=False)
    assert oracle.rel_l2(fwd, oracle.fftn(data, (1,))) <= oracle.tolerance((batch, n), (1,))


@pytest.mark.parametrize("n", [256, 1000, 4096])
def test_forced_multipass_and_bluestein(queue, n):
    """Planner branches that the default plan would not take at these sizes."""
    import reikna_amd._lib as L

    data = oracle.test_array((64, n), seed=5)
    for flags in (L.RK_FFT_FORCE_MULTIPASS, L.RK_FFT_FORCE_BLUESTEIN,
                  L.RK_FFT_FORCE_MULTIPASS | L.RK_FFT_FORCE_BLUESTEIN):
        fwd, fft = run_fft(queue, data, (1,), inplace=False, flags=flags)
        err = oracle.rel_l2(fwd, oracle.fftn(data, (1,)))
        assert err <= oracle.tolerance((64, n), (1,)), (flags, err)


def test_fused_transformations(queue):
    """Input scaled by a run-time parameter and output divided by a constant,
    both fused into the FFT's first/last pass."""
    shape, axes = (64, 4096), (1,)
    data = oracle.test_array(shape, seed=21)
    dev = Array.from_host(queue, data)
    out = Array.empty_like(queue.device, dev)
    fft = FFT(dev, axes=axes)
    scale = mul_param(dev, numpy.float32)
    fft.parameter.input.connect(scale, scale.output, input_prime=scale.input, param=scale.param)
    div = div_const(dev, 8.0)
    fft.parameter.output.connect(div, div.input, output_prime=div.output)
    assert list(fft.signature.parameters) == ["output_prime", "input_prime", "param", "inverse"]
    fftc = fft.compile(queue.device)
    fftc(queue, out, dev, 3.0)
    ref = oracle.fftn(data * 3.0, axes) / 8.0
    assert oracle.rel_l2(out.get(queue), ref) <= oracle.tolerance(shape, axes)


@pytest.mark.parametrize(
    "shape,axes,rows",
    [
        ((4096, 997), (1,), None),              # single-kernel Bluestein (chirp pre-op + fused input)
        ((8, 65521), (1,), None),               # Bluestein as two multi-pass chains
        ((64, 2**20), (1,), [0, 17, 40, 63]),   # the chunked two-pass chain (16 x 1024 tiles)
        ((256, 256, 256), (0, 1, 2), None),     # three axes, the input fused into the first axis's pass
    ],
    ids=["4096x997", "8x65521", "64x2^20", "256^3"],
)
@pytest.mark.parametrize("inverse", [False, True])
def test_fused_input_on_every_plan_kind(queue, shape, axes, rows, inverse):
    """An input transformation (mul_param) and an output one (div_const)
    fused on the plans where the first pass has its own pre-op or is one of
    several: reikna fuses its load/store functions with pad, chirp and
    kweights (fft.mako:867-890, tests/test_fft/test_fft.py:173-193)."""
    data = oracle.test_array(shape, seed=31)
    dev = Array.from_host(queue, data)
    out = Array.empty_like(queue.device, dev)
    fft = FFT(dev, axes=axes)
    scale = mul_param(dev, numpy.float32)
    fft.parameter.input.connect(scale, scale.output, input_prime=scale.input, param=scale.param)
    div = div_const(dev, 4.0)
    fft.parameter.output.connect(div, div.input, output_prime=div.output)
    fftc = fft.compile(queue.device)
    fftc(queue, out, dev, -1.5, int(inverse))
    res = out.get(queue)
    x = data * numpy.float32(-1.5)
    trf = oracle.ifftn if inverse else oracle.fftn
    tol = oracle.tolerance(shape, axes)
    if rows is None:
        err = oracle.rel_l2(res, trf(x, axes) / 4.0)
    else:
        err = max(oracle.rel_l2(res[r], trf(x[r], [a - 1 for a in axes]) / 4.0) for r in rows)
    assert err <= tol, f"{shape} over {axes}: rel-L2 {err:.3g} > {tol:.3g}"


@pytest.mark.parametrize("shape,rows", [((16, 2048, 2048), (0, 15)), ((4, 8192, 2048), (1, 3))],
                         ids=["16x2048^2", "4x8192x2048"])
def test_chunked_batched_2d(queue, shape, rows):
    """Batches of 2-D transforms larger than the Infinity Cache run in chunks
    of the batch dimension (the second with a two-pass axis through the
    workspace); forward and inverse, in and out of place."""
    import reikna_amd._lib  # noqa: F401

    data = oracle.test_array(shape, seed=61)
    for inverse, inplace in ((False, False), (True, True)):
        res, fft = run_fft(queue, data, (1, 2), inverse=inverse, inplace=inplace)
        assert fft.native_plan.chunk_rows > 0
        trf = oracle.ifftn if inverse else oracle.fftn
        for r in rows:
            err = oracle.rel_l2(res[r], trf(data[r], (0, 1), workers=16))
            assert err <= oracle.tolerance(shape, (1, 2)), (shape, inverse, r, err)


def test_strided_input_and_output(queue):
    """Non-contiguous arrays: reikna honours arbitrary strides (transformation_modules.py:32-69)."""
    base = oracle.test_array((64, 2048), seed=9)
    data = base[:, ::2]                      # (64, 1024) with a 16-byte element stride
    from reikna_amd import ArrayMetadata

    in_meta = ArrayMetadata((64, 1024), numpy.complex64, strides=(2048 * 8, 16))
    dev = Array(queue.device, in_meta)
    dev.set(queue, data)
    out = Array.empty(queue.device, (64, 1024), numpy.complex64)
    fft = FFT(in_meta, axes=(1,))
    # the output type differs from the input (contiguous): connect a copy
    from reikna_amd.transformations import copy

    cp = copy(in_meta, out_arr_t=ArrayMetadata((64, 1024), numpy.complex64))
    fft.parameter.output.connect(cp, cp.input, output_contig=cp.output)
    fftc = fft.compile(queue.device)
    fftc(queue, out, dev)
    assert oracle.rel_l2(out.get(queue), oracle.fftn(data, (1,))) <= oracle.tolerance((64, 1024), (1,))


@pytest.mark.parametrize(
    "shape,axes",
    [((512, 3 * 2**16), (1,)), ((256, 5 * 2**16), (1,)), ((2**16, 512), (0,)), ((2**14, 4096), (0, 1))],
    ids=lambda v: str(v),
)
def test_large_streaming_passes(queue, shape, axes):
    """Transforms larger than the Infinity Cache: streaming multi-pass chains,
    whose column passes of lengths with a factor 3 or 5 run kernels with the
    load/store modes compiled in."""
    data = oracle.test_array(shape, seed=sum(shape))
    fwd, fft = run_fft(queue, data, axes, inplace=False)
    assert any(p.nt_load or p.nt_store for p in fft.native_plan.passes)
    assert oracle.rel_l2(fwd, oracle.fftn(data, axes)) <= oracle.tolerance(shape, axes)
    inv, _ = run_fft(queue, data, axes, inverse=True, inplace=True)
    assert oracle.rel_l2(inv, oracle.ifftn(data, axes)) <= oracle.tolerance(shape, axes)


@pytest.mark.parametrize(
    "shape,axes,lengths",
    [
        ((8, 2**20), (1,), [1024, 1024]),        # unchunked, both 16 x 1024 tiles under launch bounds
        ((3, 2**19), (1,), [1024, 512]),         # one bounded pass, one 512-point pass
        ((2**17, 4), (0,), [512, 256]),          # a strided axis: column passes with an inner dimension
        ((2, 3 * 2**18), (1,), [1024, 768]),     # a radix-3 pass beside a bounded one
        ((4, 10**6), (1,), [1000, 1000]),        # 1000-point passes (8-instance tiles)
    ],
    ids=["8x2^20", "3x2^19", "2^17x4-axis0", "2x3*2^18", "4x10^6"],
)
def test_two_pass_chains(queue, shape, axes, lengths):
    """Complex64 chains take passes of up to 1024 points; the 16 x 1024
    column tiles with compiled-in modes run with launch bounds for two
    work-groups per CU.  Forward out of place and inverse in place against
    the oracle."""
    data = oracle.test_array(shape, seed=sum(shape) + 7)
    fwd, fft = run_fft(queue, data, axes, inplace=False)
    ps = fft.native_plan.passes
    assert sorted(p.L for p in ps) == sorted(lengths)
    assert [p.occ for p in ps] == [p.L == 1024 and p.C == 16 for p in ps]
    assert oracle.rel_l2(fwd, oracle.fftn(data, axes)) <= oracle.tolerance(shape, axes)
    inv, _ = run_fft(queue, data, axes, inverse=True, inplace=True)
    assert oracle.rel_l2(inv, oracle.ifftn(data, axes)) <= oracle.tolerance(shape, axes)


@pytest.mark.parametrize(
    "shape,axes",
    [((192, 192, 192), (0, 1, 2)), ((96, 96, 96), (0, 1, 2)), ((64, 3 * 2**13), (1,)), ((64, 4000), (1,)),
     ((4096, 1000), (1,)), ((8192, 96), (1,)), ((4096, 160), (1,))],
    ids=lambda v: str(v),
)
def test_fixed_mode_passes(queue, shape, axes):
    """Column passes and single-pass rows of lengths with a factor 3 or 5 run
    kernels with their load/store modes compiled in."""
    data = oracle.test_array(shape, seed=sum(shape) + 1)
    fwd, fft = run_fft(queue, data, axes, inplace=False)
    assert any(p.fixed_modes for p in fft.native_plan.passes)
    assert oracle.rel_l2(fwd, oracle.fftn(data, axes)) <= oracle.tolerance(shape, axes)
    inv, _ = run_fft(queue, data, axes, inverse=True, inplace=True)
    assert oracle.rel_l2(inv, oracle.ifftn(data, axes)) <= oracle.tolerance(shape, axes)


@pytest.mark.parametrize("shape", [(4096, 40), (2048, 100), (16384, 5), (3 * 2**12, 7), (1024, 3, 5)])
def test_long_strided_axis(queue, shape):
    """Axis 0 long and strided: column tiles narrower than 16 instances, sent to
    one XCD in groups of consecutive tiles; above 4096 points two passes."""
    data = oracle.test_array(shape, seed=sum(shape))
    for inverse in (False, True):
        res, _ = run_fft(queue, data, (0,), inverse=inverse, inplace=False)
        ref = oracle.ifftn(data, (0,)) if inverse else oracle.fftn(data, (0,))
        assert oracle.rel_l2(res, ref) <= oracle.tolerance(shape, (0,))


def test_graph_capture_of_plans(queue):
    """A multi-pass plan (workspace ping-pong) and a plan with fused
    transformations recorded once as a HIP graph and replayed several times,
    with eager work and syncs in between, each replay equal to the eager run."""
    data = oracle.test_array((8, 2**16), seed=31)
    arr = Array.from_host(queue, data)
    out = Array.empty_like(queue.device, arr)
    fft = FFT(arr, axes=(1,)).compile(queue.device)
    fft(queue, out, arr, 0)
    eager = out.get(queue)
    assert oracle.rel_l2(eager, oracle.fftn(data, (1,))) <= oracle.tolerance((8, 2**16), (1,))

    arr2 = Array.from_host(queue, data)
    fused = FFT(arr2, axes=(1,))
    scale = div_const(arr2, 4.0)
    fused.parameter.output.connect(scale, scale.input, scaled=scale.output)
    fusedc = fused.compile(queue.device)
    out2 = Array.empty_like(queue.device, arr2)
    fusedc(queue, out2, arr2, 0)
    eager2 = out2.get(queue)

    with queue.capture() as cap:
        fft(queue, out, arr, 0)
        fusedc(queue, out2, arr2, 0)
    for k in range(3):
        out.set(queue, numpy.zeros_like(data))
        out2.set(queue, numpy.zeros_like(data))
        cap.graph.launch(queue)
        assert numpy.array_equal(out.get(queue), eager), k
        assert numpy.array_equal(out2.get(queue), eager2), k
        fft(queue, out, arr, 1)          # eager work between replays
        queue.synchronize()


@pytest.mark.parametrize("case", ["n8", "n60", "n997", "n1000", "n1024", "n12288", "d2_17x200", "d3_8x12x10"])
def test_golden_fixtures(queue, case):
    """The HIP path against the committed FP64 fixtures (tests/golden/, written
    by make_golden.py from numpy.fft), forward and inverse."""
    from pathlib import Path

    g = numpy.load(Path(__file__).parent / "golden" / "fft_golden.npz")
    data, axes = g[case + "_in"], tuple(int(a) for a in g[case + "_axes"])
    tol = oracle.tolerance(data.shape, axes)
    for inverse, key in ((False, "_fwd"), (True, "_inv")):
        res, _ = run_fft(queue, data, axes, inverse=inverse, inplace=False)
        assert oracle.rel_l2(res, g[case + key]) <= tol, (case, inverse)


def test_checked_execute(queue):
    """rk_fft_execute_checked runs the plan when the buffers hold its layouts'
    extents and refuses shorter ones before launching anything."""
    import numpy as np

    data = oracle.test_array((16, 4096), seed=81)
    inp = Array.from_host(queue, data)
    out = Array.from_host(queue, np.zeros_like(data))
    fft = FFT(inp, axes=(1,))
    fft.compile(queue.device)
    native = fft.native_plan
    with pytest.raises(ValueError, match="output buffer of"):
        native.execute_checked(out.ptr, out.nbytes - 8, inp.ptr, inp.nbytes, 0, None, 0, queue.stream)
    queue.synchronize()
    assert not np.any(out.get(queue))          # nothing was launched
    native.execute_checked(out.ptr, out.nbytes, inp.ptr, inp.nbytes, 0, None, 0, queue.stream)
    err = oracle.rel_l2(out.get(queue), oracle.fftn(data, (1,)))
    assert err <= oracle.tolerance((16, 4096), (1,))


def test_one_plan_on_two_streams(queue):
    """A plan is immutable: the same native plan runs on two streams at once,
    each call with its own workspace, and both results are correct."""
    from reikna_amd import ArrayMetadata
    from reikna_amd.fft.native import NativeFftPlan
    from reikna_amd.runtime import Queue

    shape = (16, 2**16)
    meta = ArrayMetadata(shape, numpy.complex64)
    plan = NativeFftPlan(meta, meta, (1,))
    assert plan.workspace_bytes > 0
    q2 = Queue(queue.device)
    datas = [oracle.test_array(shape, seed=s) for s in (101, 102)]
    ins = [Array.from_host(queue, d) for d in datas]
    outs = [Array.empty(queue.device, shape, numpy.complex64) for _ in range(2)]
    wss = [Array.empty(queue.device, (plan.workspace_bytes,), numpy.uint8) for _ in range(2)]
    queue.synchronize()
    for _ in range(8):
        for k, q in enumerate((queue, q2)):
            plan.execute(outs[k].ptr, ins[k].ptr, 0, wss[k].ptr, q.stream)
    queue.synchronize()
    q2.synchronize()
    for k in range(2):
        err = oracle.rel_l2(outs[k].get(queue), oracle.fftn(datas[k], (1,)))
        assert err <= oracle.tolerance(shape, (1,)), (k, err)


@pytest.mark.parametrize("n", [2**26, 5 * 2**23, 2**24 + 43],
                         ids=["2^26", "5x2^23", "prime 2^24+43"])
def test_single_long_transform(queue, n):
    """One 1-D transform of 2^24 - 2^26 points (no batch dimension, so no
    chunks): four streaming passes, a radix-5 chain, and a prime through
    Bluestein chains over M = 5 * 2^23."""
    data = oracle.test_array((n,), seed=n % 1000)
    fwd, fft = run_fft(queue, data, (0,), inplace=False)
    assert len(fft.native_plan.passes) >= 3 and fft.native_plan.chunk_rows == 0
    assert oracle.rel_l2(fwd, oracle.fftn(data, (0,), workers=16)) <= oracle.tolerance((n,), (0,))
