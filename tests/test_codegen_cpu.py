"""Planner structure and generated-kernel compilation without a GPU.

Compiling against a VirtualDevice runs the native planner in dry-run mode
(RK_FFT_DRY_RUN) and compiles every generated kernel for gfx950 with hiprtc,
so the pass chains the reference's planner would build (reikna/fft/fft.py:487-584)
and every fused-transformation kernel are checked on the CPU.
"""

import numpy
import pytest

from reikna_amd import ArrayMetadata
from reikna_amd.algorithms import PureParallel
from reikna_amd.fft import FFT
from reikna_amd.runtime import VirtualDevice
from reikna_amd.transformations import combine_complex, copy, div_const, mul_param, norm_const, split_complex


@pytest.fixture(scope="module")
def vdev():
    return VirtualDevice()


def passes(shape, axes, flags=0, dev=None):
    fft = FFT(ArrayMetadata(shape, numpy.complex64), axes=axes, flags=flags)
    fft.compile(dev or VirtualDevice())
    return fft.native_plan.passes


def test_plan_single_pass_rows(vdev):
    ps = passes((1024, 1024), (1,), dev=vdev)
    assert [(p.L, p.load_mode, p.store_mode) for p in ps] == [(1024, "direct", "direct")]


def test_plan_multi_pass_large_axis(vdev):
    # fewest passes with every pass length <= 1024 (16-column tiles)
    for shape, n_pass in (((64, 2**20), 2), ((16, 2**22), 3)):
        ps = passes(shape, (1,), dev=vdev)
        assert len(ps) == n_pass and numpy.prod([p.L for p in ps]) == shape[1] and max(p.L for p in ps) <= 1024
        assert ps[0].load_mode == "direct" and ps[0].store_mode == "flat"     # column load, contiguous store
        for p in ps[1:]:
            assert p.load_mode == "direct" and p.store_mode == "direct"     # column both sides
            assert p.variant & 3 == 1                                        # inter-pass twiddle variant
        ws = [("in", "ws0"), ("ws0", "out")] if n_pass == 2 else [("in", "ws0"), ("ws0", "ws1"), ("ws1", "out")]
        assert [(p.src, p.dst) for p in ps] == ws
    ps = passes((64, 2**16), (1,), dev=vdev)
    assert [p.L for p in ps] == [256, 256]


def test_plan_three_axes(vdev):
    ps = passes((256, 256, 256), (0, 1, 2), dev=vdev)
    assert [(p.axis, p.L) for p in ps] == [(2, 256), (1, 256), (0, 256)]
    assert all(p.dst == "out" for p in ps)   # single-pass axes run in place on `out`


@pytest.mark.parametrize("n,expect", [(997, 2048), (1009, 2048), (4093, 8192), (5003, 10240), (12289, 32768),
                                      (65521, 131072)])
def test_plan_bluestein_length(vdev, n, expect):
    ps = passes((4, n), (1,), dev=vdev)
    assert all(p.axis_len == expect for p in ps)
    if expect <= 8192:
        assert len(ps) == 1 and ps[0].kind == 1   # single-kernel Bluestein
    else:
        # forward chain over M, then a second forward chain of the same passes
        assert len(ps) % 2 == 0 and len(ps) >= 4
        half = len(ps) // 2
        assert [p.L for p in ps[:half]] == [p.L for p in ps[half:]]
        assert ps[0].pre_ops == 1 and ps[half - 1].post_ops == 1 and ps[-1].post_ops == 6


def test_plan_smooth_non_pow2_is_not_bluestein(vdev):
    for n in (1000, 7 * 2**9):
        ps = passes((4, n), (1,), dev=vdev)
        assert len(ps) == 1 and ps[0].kind == 0 and ps[0].L == n
    # above 4096 points: two Stockham passes, no Bluestein padding
    for n in (3 * 2**12, 5 * 2**11):
        ps = passes((4, n), (1,), dev=vdev)
        assert len(ps) == 2 and all(p.kind == 0 and p.axis_len == n for p in ps)
        assert ps[0].L * ps[1].L == n


def test_plan_flags(vdev):
    import reikna_amd._lib as L

    assert len(passes((8, 1024), (1,), L.RK_FFT_FORCE_MULTIPASS, dev=vdev)) == 2
    ps = passes((8, 1024), (1,), L.RK_FFT_FORCE_BLUESTEIN, dev=vdev)
    assert ps[0].kind == 1 and ps[0].axis_len == 2048


def test_dry_plan_cannot_execute(vdev):
    fft = FFT(ArrayMetadata((4, 64), numpy.complex64), axes=(1,))
    fft.compile(vdev)
    with pytest.raises(ValueError, match="dry-run"):
        fft.native_plan.execute(1, 1, 0, None, None)


# -- generated kernels ---------------------------------------------------------
def test_fused_input_and_output_compile(vdev):
    arr = ArrayMetadata((64, 4096), numpy.complex64)
    fft = FFT(arr, axes=(1,))
    scale = mul_param(arr, numpy.float32)
    fft.parameter.input.connect(scale, scale.output, input_prime=scale.input, param=scale.param)
    div = div_const(arr, 8.0)
    fft.parameter.output.connect(div, div.input, output_prime=div.output)
    assert list(fft.signature.parameters) == ["output_prime", "input_prime", "param", "inverse"]
    fft.compile(vdev)


def test_fused_multi_pass_and_multi_axis_compile(vdev):
    for shape, axes in (((4, 2**16), (1,)), ((32, 64, 48), (0, 1, 2)), ((8, 997), (1,)), ((2, 65521), (1,))):
        arr = ArrayMetadata(shape, numpy.complex64)
        fft = FFT(arr, axes=axes)
        real = ArrayMetadata(shape, numpy.float32)
        comb = combine_complex(arr)
        fft.parameter.input.connect(comb, comb.output, re=comb.real, im=comb.imag)
        nrm = norm_const(arr, 2)
        fft.parameter.output.connect(nrm, nrm.input, power=nrm.output)
        assert list(fft.signature.parameters) == ["power", "re", "im", "inverse"]
        assert fft.signature.parameters["re"].annotation.array_metadata == real
        fft.compile(vdev)


def test_split_complex_output(vdev):
    arr = ArrayMetadata((16, 256), numpy.complex64)
    fft = FFT(arr, axes=(1,))
    sp = split_complex(arr)
    fft.parameter.output.connect(sp, sp.input, re=sp.real, im=sp.imag)
    assert list(fft.signature.parameters) == ["re", "im", "input", "inverse"]
    fft.compile(vdev)


def test_pure_parallel_from_transformation_compiles(vdev):
    arr = ArrayMetadata((100, 30), numpy.complex64)
    trf = copy(arr)
    pp = PureParallel.from_trf(trf, trf.input)
    c = pp.compile(vdev)
    assert len(c.kernel_sources) == 1
    src = c.kernel_sources[0]
    # one thread per element over a flat grid, the index decoded with literal extents
    assert "rk_elementwise" in src and "rk_flat >= 3000ull" in src and "% 30ull" in src


def test_pure_parallel_snippet_and_errors(vdev):
    import pytest

    from reikna_amd.core import Parameter, Annotation

    arr = ArrayMetadata((7, 5, 3), numpy.float32)
    pp = PureParallel(
        [Parameter("out", Annotation(arr, "o")), Parameter("inp", Annotation(arr, "i"))],
        "${out.store_idx}(${idxs[0]}, ${idxs[1]}, ${idxs[2]}, ${inp.load_idx}(${idxs[0]}, ${idxs[1]}, ${idxs[2]}) * 2);",
    )
    src = pp.compile(vdev).kernel_sources[0]
    assert "rk_flat >= 105ull" in src and "% 3ull" in src and "% 5ull" in src
    trf = copy(arr)
    with pytest.raises(ValueError, match="cannot be a scalar"):
        from reikna_amd.transformations import mul_param

        m = mul_param(arr, numpy.float32)
        PureParallel.from_trf(m, m.param)
    other = copy(arr)
    with pytest.raises(ValueError, match="must belong"):
        PureParallel.from_trf(trf, other.input)


def chunked(shape, axes, dev):
    fft = FFT(ArrayMetadata(shape, numpy.complex64), axes=axes)
    fft.compile(dev)
    return fft.native_plan.chunk_rows, fft.native_plan.passes


def test_plan_streaming_accesses(vdev):
    # chains whose data exceed the 256 MB Infinity Cache run in chunks of
    # 128 MiB of intermediate: the caller's input is loaded and the output
    # stored with the nontemporal hint, the intermediate stays cached
    for shape, rows in (((64, 2**20), 32), ((4096, 12288), 2730), ((16, 2**22), 4)):
        n, ps = chunked(shape, (1,), vdev)
        assert n == rows, (shape, n)
        assert ps[0].nt_load and not ps[0].nt_store and ps[0].src == "in"
        assert ps[-1].nt_store and not ps[-1].nt_load and ps[-1].dst == "out"
        assert not any(p.nt_load or p.nt_store for p in ps[1:-1])
    # the in-place passes of a 128 MiB 3-D transform and small transforms
    # keep plain accesses, unchunked
    assert chunked((256, 256, 256), (0, 1, 2), vdev)[0] == 0
    assert not any(p.stream for p in passes((256, 256, 256), (0, 1, 2), dev=vdev))
    assert not any(p.stream for p in passes((1024, 1024), (1,), dev=vdev))
    # Bluestein passes have no streaming twins
    assert not any(p.stream for p in passes((4096, 65521), (1,), dev=vdev) if p.variant & 2)


def test_plan_fixed_mode_kernels(vdev):
    # the column passes of large multi-pass (chunked) transforms run kernels
    # with their load/store modes compiled in, power-of-two lengths included;
    # single-pass power-of-two rows keep the generic kernel
    for shape in ((4096, 12288), (512, 3 * 2**16), (256, 5 * 2**16), (64, 2**20)):
        ps = passes(shape, (1,), dev=vdev)
        assert len(ps) >= 2 and ps[0].nt_load and ps[-1].nt_store   # chunked chains
        assert all(p.fixed_modes for p in ps), shape
        assert (ps[0].load_mode, ps[0].store_mode) == ("direct", "flat")
        assert all((p.load_mode, p.store_mode) == ("direct", "direct") for p in ps[1:])
        assert not any(p.jit for p in ps), shape   # all in the ahead-of-time catalogue
    assert not any(p.fixed_modes for p in passes((1024, 1024), (1,), dev=vdev))
    # single-pass rows of 3/5-smooth lengths with >= 32 threads per instance
    for n in (1000, 2000, 1536, 4000):
        ps = passes((64, n), (1,), dev=vdev)
        assert [(p.load_mode, p.store_mode, p.fixed_modes) for p in ps] == [("direct", "direct", True)]
    # short rows stream their tiles flat on both sides, with FLAT/FLAT compiled in
    ps = passes((64, 96), (1,), dev=vdev)
    assert [(p.load_mode, p.store_mode, p.fixed_modes) for p in ps] == [("flat", "flat", True)]
    assert not passes((64, 256), (1,), dev=vdev)[0].fixed_modes   # power of two: generic kernel
    # the strided (column) passes of 256^3 have DIRECT/DIRECT compiled in, its row pass not
    assert [p.fixed_modes for p in passes((256, 256, 256), (0, 1, 2), dev=vdev)] == [False, True, True]


def test_plan_two_pass_chains(vdev):
    # complex64 chains take passes of up to 1024 points: 2^17 - 2^20 points in
    # two passes; the 16 x 1024 column tiles (1024 threads) with their modes
    # compiled in run with launch bounds for two work-groups per CU, all from
    # the ahead-of-time catalogue
    for shape, ls in (((64, 2**20), [1024, 1024]), ((256, 2**18), [512, 512]), ((512, 3 * 2**16), [512, 384]),
                      ((8, 2**20), [1024, 1024]), ((16, 2**22), [128, 256, 128])):
        ps = passes(shape, (1,), dev=vdev)
        assert sorted(p.L for p in ps) == sorted(ls), shape   # (384 and 512 are equally balanced)
        assert [p.occ for p in ps] == [p.L == 1024 for p in ps], shape
        assert all(p.fixed_modes and not p.jit for p in ps), shape
    # no bound where the kernel would spill (generic modes) or for other widths
    assert not any(p.occ for p in tuned_plan((64, 2**20), (1,), vdev, fixed_modes=0).passes)
    assert not any(p.occ for p in tuned_plan((64, 2**20), (1,), vdev, col_tile=8).passes)
    assert not any(p.occ for p in passes((1024, 1024), (1,), dev=vdev))
    # complex128 chains keep passes of at most 256 points
    fft = FFT(ArrayMetadata((64, 2**20), numpy.complex128), axes=(1,))
    fft.compile(vdev)
    assert [p.L for p in fft.native_plan.passes] == [128, 128, 64]
    assert not any(p.occ for p in fft.native_plan.passes)


def test_plan_xcd_groups_for_narrow_column_tiles(vdev):
    # a 4096-point strided axis: column tiles of 4 instances (the pass is planned;
    # its XCD grouping is a launch detail checked on the GPU)
    ps = passes((4096, 40), (0,), dev=vdev)
    assert [(p.L, p.C, p.load_mode) for p in ps] == [(4096, 4, "direct")]


def test_other_dtypes_refused():
    # the kernels compute complex64 and complex128; a wider complex type is
    # refused, not misread, and a real array is not an FFT argument
    if numpy.dtype(numpy.clongdouble).itemsize > 16:
        with pytest.raises(NotImplementedError, match="complex128"):
            FFT(ArrayMetadata((4, 64), numpy.clongdouble))
    with pytest.raises(ValueError, match="complex dtype"):
        FFT(ArrayMetadata((4, 64), numpy.float64))


def test_complex128_plans(vdev):
    # complex128 plans: the same pass chains (passes of at most 256 points,
    # the complex64 planner held to that limit here) with twice the bytes per
    # element, so tiles of half the instances in the same LDS, twice the
    # workspace, chunks of half the rows, and fused transformations in double2
    from reikna_amd.fft import PlanTuning

    fm256 = PlanTuning(factor_max=256)
    for shape, axes in (((64, 2**20), (1,)), ((4, 65521), (1,)), ((32, 64, 48), (0, 1, 2)), ((1024, 1024), (1,))):
        p64 = FFT(ArrayMetadata(shape, numpy.complex64), axes=axes, tuning=fm256)
        p128 = FFT(ArrayMetadata(shape, numpy.complex128), axes=axes)
        p64.compile(vdev)
        p128.compile(vdev)
        a, b = p64.native_plan, p128.native_plan
        assert [p.L for p in a.passes] == [p.L for p in b.passes]
        assert all(q.C == max(1, p.C // 2) for p, q in zip(a.passes, b.passes))
        assert all(q.lds_bytes <= 160 * 1024 for q in b.passes)
        assert b.chunk_rows * 2 == a.chunk_rows
    big = FFT(ArrayMetadata((4, 65521), numpy.complex128), axes=(1,))
    big.compile(vdev)
    small = FFT(ArrayMetadata((4, 65521), numpy.complex64), axes=(1,), tuning=fm256)
    small.compile(vdev)
    assert big.native_plan.workspace_bytes == 2 * small.native_plan.workspace_bytes
    # single-kernel Bluestein up to M = 2048 (longer convolutions as chains)
    blue = FFT(ArrayMetadata((8, 997), numpy.complex128), axes=(1,))
    blue.compile(vdev)
    assert [p.kind for p in blue.native_plan.passes] == [1]
    chains = FFT(ArrayMetadata((8, 2003), numpy.complex128), axes=(1,))
    chains.compile(vdev)
    assert len(chains.native_plan.passes) >= 2
    arr = ArrayMetadata((64, 4096), numpy.complex128)
    fft = FFT(arr, axes=(1,))
    scale = mul_param(arr, numpy.float64)
    fft.parameter.input.connect(scale, scale.output, input_prime=scale.input, param=scale.param)
    div = div_const(arr, 8.0)
    fft.parameter.output.connect(div, div.input, output_prime=div.output)
    fft.compile(vdev)


def test_array_arguments_checked_against_the_compiled_layout():
    from reikna_amd.core.computation import _check_array_argument
    from reikna_amd.runtime import Array, _ForeignBuffer

    meta = ArrayMetadata((512,), numpy.complex64, strides=(16,), buffer_size=8192)
    dev = VirtualDevice()

    def arr(m, size):
        return Array(dev, m, _ForeignBuffer(dev, 0, size, None))

    _check_array_argument("a", arr(meta, 8192), meta)
    with pytest.raises(ValueError, match="strides"):
        _check_array_argument("a", arr(ArrayMetadata((512,), numpy.complex64), 4096), meta)
    with pytest.raises(ValueError, match="bytes"):
        _check_array_argument("a", arr(meta, 4096), meta)
    with pytest.raises(ValueError, match="shape"):
        _check_array_argument("a", arr(ArrayMetadata((256,), numpy.complex64), 8192), meta)


def test_plan_chunks_batched_multi_axis(vdev):
    # a batch of 2-D transforms larger than the Infinity Cache runs in chunks
    # of the batch dimension: the row pass writes the output, the column pass
    # re-reads it from the cache, so only the plan's first loads and last
    # stores are nontemporal
    n, ps = chunked((64, 2048, 2048), (1, 2), vdev)
    assert n == 8 and len(ps) == 2
    assert ps[0].nt_load and not ps[0].nt_store and ps[1].nt_store and not ps[1].nt_load
    # a multi-pass axis (workspace) inside a batched 2-D transform
    n, ps = chunked((4, 8192, 2048), (1, 2), vdev)
    assert n == 2 and len(ps) == 3
    # the transformed leading axis cannot be chunked
    assert chunked((2**14, 4096), (0, 1), vdev)[0] == 0


def tuned_plan(shape, axes, dev, **knobs):
    from reikna_amd.fft import PlanTuning

    fft = FFT(ArrayMetadata(shape, numpy.complex64), axes=axes, tuning=PlanTuning(**knobs))
    fft.compile(dev)
    return fft.native_plan


def test_plan_tuning_knobs(vdev):
    # rk_fft_plan_create_tuned: each knob changes the plan it names
    base = [p.L for p in tuned_plan((64, 2**20), (1,), vdev).passes]
    assert base == [1024, 1024]
    assert [p.L for p in tuned_plan((64, 2**20), (1,), vdev, factor_max=256).passes] == [128, 128, 64]
    assert [p.L for p in tuned_plan((64, 2**20), (1,), vdev, factor_max=256, first_factor=64).passes] == [64, 128, 128]
    assert [p.L for p in tuned_plan((64, 2**20), (1,), vdev, factor_max=128).passes] == [128, 128, 64]
    ps = tuned_plan((64, 2**20), (1,), vdev, factor_max=256, col_tile=64).passes
    assert [p.C for p in ps] == [64, 64, 64]
    assert tuned_plan((64, 2**20), (1,), vdev).chunk_rows == 32
    assert tuned_plan((64, 2**20), (1,), vdev, factor_max=256).chunk_rows == 16
    assert tuned_plan((64, 2**20), (1,), vdev, chunk_bytes=0).chunk_rows == 0
    assert tuned_plan((64, 2**20), (1,), vdev, chunk_bytes=64 << 20).chunk_rows == 8
    assert not any(p.fixed_modes for p in tuned_plan((64, 2**20), (1,), vdev, fixed_modes=0).passes)
    # knobs the plan cannot honour are ignored: a first factor that does not
    # divide the axis, a tile wider than 1024 threads
    assert [p.L for p in tuned_plan((64, 2**20), (1,), vdev, first_factor=3).passes] == base
    assert [p.C for p in tuned_plan((4, 4096), (1,), vdev, row_tile=64).passes] == \
        [p.C for p in tuned_plan((4, 4096), (1,), vdev).passes]


def test_plan_tuning_out_of_range_refused(vdev):
    from reikna_amd.fft import PlanTuning

    for knobs in ({"fixed_modes": 2}, {"chunk_bytes": -2}, {"col_tile": -1}, {"factor_max": 1}):
        with pytest.raises(ValueError, match="knob out of range|factor_max"):
            tuned_plan((64, 4096), (1,), vdev, **knobs)
    assert PlanTuning().to_c().reserved == 0


def test_checked_execute_refuses_short_buffers(vdev):
    # rk_fft_execute_checked compares the buffer sizes with the extents the
    # plan's layouts reach before launching anything (here a dry-run plan,
    # which then refuses to execute)
    from reikna_amd.runtime import ArrayMetadata as AM

    plan = FFT(AM((8, 4096), numpy.complex64), axes=(1,))
    plan.compile(vdev)
    native = plan.native_plan
    full = 8 * 4096 * 8
    fake = 1 << 20   # never dereferenced: every check comes first
    with pytest.raises(ValueError, match="input buffer of"):
        native.execute_checked(fake, full, fake, full - 8, 0, None, 0, None)
    with pytest.raises(ValueError, match="output buffer of"):
        native.execute_checked(fake, full - 1, fake, full, 0, None, 0, None)
    with pytest.raises(ValueError, match="dry-run plan"):
        native.execute_checked(fake, full, fake, full, 0, None, 0, None)
    # a strided view reaches one past its last element from the buffer start
    view = AM((8, 4096), numpy.complex64, strides=(2 * 4096 * 8, 8), first_element_offset=4096 * 8)
    pv = FFT(view, axes=(1,))
    pv.compile(vdev)
    reach = (4096 + 7 * 2 * 4096 + 4096) * 8
    with pytest.raises(ValueError, match="input buffer of"):
        pv.native_plan.execute_checked(fake, reach, fake, reach - 8, 0, None, 0, None)
    with pytest.raises(ValueError, match="dry-run plan"):
        pv.native_plan.execute_checked(fake, reach, fake, reach, 0, None, 0, None)
    # a multi-pass plan needs its workspace
    big = FFT(AM((64, 2**20), numpy.complex64), axes=(1,))
    big.compile(vdev)
    n = 64 * 2**20 * 8
    ws = big.native_plan.workspace_bytes
    assert ws > 0
    with pytest.raises(ValueError, match="workspace of"):
        big.native_plan.execute_checked(fake, n, fake, n, 0, fake, ws - 1, None)


def test_large_prime_plans_quickly(vdev):
    # the Bluestein kernel spectrum of a prime near 2^24 (M = 5 * 2^23) is an
    # FP64 host FFT of 42 M points: iterative radix passes take ~5 s for the
    # whole plan (a recursive one took ~50 s)
    import time

    t0 = time.perf_counter()
    fft = FFT(ArrayMetadata((16777259,), numpy.complex64))
    fft.compile(vdev)
    assert time.perf_counter() - t0 < 25.0
    assert fft.native_plan.passes[0].axis_len == 5 * 2**23
