"""GPU parity of batched uneven transforms (BatchFFT) against the FP64 oracle.

Every item of a batch is checked against numpy.fft in FP64 at the complex64
bound (rel-L2 <= 2e-7 * log2 N).  Beyond parity: a transform computed alone
and inside a mixed batch is bit-identical (its pass chain depends only on its
length), HIP-graph replays match eager runs, repeated launches reuse the
queue state, and the per-item path (grouped mode) agrees within tolerance.
"""

import numpy
import pytest

import oracle
from reikna_amd import Array
from reikna_amd.fft import BatchFFT, BatchItem, BatchOptions
from reikna_amd.fft.batch import packed_items, sample_lengths

pytestmark = pytest.mark.gpu

SPECIAL = [64, 1000, 997, 4096, 8192, 65521, 49152, 40960, 2, 3, 7, 13, 1, 5003, 12289, 65536]


def packed_data(lengths, counts, seed):
    total = sum(n * c for n, c in zip(lengths, counts))
    return oracle.test_array((total,), seed=seed)


def check_items(items, host_in, host_out, inverse=False):
    worst = 0.0
    for it in items:
        for j in range(it.count):
            x = host_in[it.in_offset + j * it.in_dist: it.in_offset + j * it.in_dist + it.length]
            y = host_out[it.out_offset + j * it.out_dist: it.out_offset + j * it.out_dist + it.length]
            if inverse:
                ref = oracle.ifftn(x, (0,))
                if it.scale:
                    ref = ref * it.length
            else:
                ref = oracle.fftn(x, (0,))
                if it.scale:
                    ref = ref / it.length
            err = oracle.rel_l2(y, ref)
            tol = oracle.tolerance((it.length,))
            assert err <= tol, f"length {it.length} (item {it}, copy {j}): rel-L2 {err:.3g} > {tol:.3g}"
            worst = max(worst, err / tol)
    return worst


def run_batch(queue, items, host_in, n_out, inverse=False, options=None, inplace=False):
    batch = BatchFFT(queue.device, items, options)
    inp = Array.from_host(queue, host_in)
    out = inp if inplace else Array.from_host(queue, numpy.zeros(n_out, numpy.complex64))
    batch(queue, out, inp, int(inverse))
    res = out.get(queue)
    batch.check()
    return res, batch


@pytest.mark.parametrize("inverse", [False, True])
def test_mixed_batch_matches_oracle(queue, inverse):
    rng = numpy.random.default_rng(21)
    lengths = sample_lengths(40, seed=11) + SPECIAL
    counts = [int(c) for c in rng.integers(1, 4, len(lengths))]
    items = packed_items(lengths, counts)
    data = packed_data(lengths, counts, seed=6)
    res, batch = run_batch(queue, items, data, data.size, inverse=inverse)
    assert batch.info["launches"] == 2   # the queue launch + the copy of the length-1 item
    check_items(items, data, res, inverse=inverse)


def test_scaled_items(queue):
    lengths = [64, 997, 3 * 2**12, 65536, 65521]
    items = packed_items(lengths, None, scale=True)
    data = packed_data(lengths, [1] * len(lengths), seed=7)
    fwd, _ = run_batch(queue, items, data, data.size)
    check_items(items, data, fwd)
    inv, _ = run_batch(queue, items, data, data.size, inverse=True)
    check_items(items, data, inv, inverse=True)


def test_strided_items(queue):
    # distances larger than the length, separate in/out layouts, gaps
    items = [
        BatchItem(1024, 3, in_offset=5, out_offset=0, in_dist=1100, out_dist=1024),
        BatchItem(997, 2, in_offset=4000, out_offset=3072, in_dist=1000, out_dist=2000),
        BatchItem(8192, 2, in_offset=7000, out_offset=7072, in_dist=9000, out_dist=8192),
    ]
    n_in = 7000 + 9000 + 8192
    n_out = 7072 + 8192 * 2
    data = oracle.test_array((n_in,), seed=8)
    res, _ = run_batch(queue, items, data, n_out)
    check_items(items, data, res)


def test_inplace(queue):
    lengths = [256, 1000, 4099, 16384, 3 * 2**13]
    counts = [2, 1, 2, 1, 1]
    items = packed_items(lengths, counts)
    data = packed_data(lengths, counts, seed=9)
    res, _ = run_batch(queue, items, data, data.size, inplace=True)
    check_items(items, data, res)


def test_alone_and_in_mix_identical(queue):
    lengths = sample_lengths(24, seed=12) + [997, 65521, 8192]
    items = packed_items(lengths)
    data = packed_data(lengths, [1] * len(lengths), seed=10)
    mixed, _ = run_batch(queue, items, data, data.size)
    for it in items[::3] + items[-3:]:
        x = data[it.in_offset: it.in_offset + it.length]
        alone, _ = run_batch(queue, [BatchItem(it.length)], x, it.length)
        assert numpy.array_equal(alone, mixed[it.out_offset: it.out_offset + it.length]), f"length {it.length}"


def test_grouped_mode_agrees(queue):
    lengths = sample_lengths(16, seed=13) + [997, 12289]
    items = packed_items(lengths)
    data = packed_data(lengths, [1] * len(lengths), seed=11)
    q, bq = run_batch(queue, items, data, data.size)
    g, bg = run_batch(queue, items, data, data.size, options=BatchOptions(mode="grouped"))
    assert bq.info["launches"] == 1 and bg.info["launches"] >= len(lengths)
    check_items(items, data, g)
    for it in items:
        a = q[it.out_offset: it.out_offset + it.length]
        b = g[it.out_offset: it.out_offset + it.length]
        assert oracle.rel_l2(a, b.astype(numpy.complex128)) <= 2 * oracle.tolerance((it.length,))


def test_repeated_launches_and_graph_replay(queue):
    lengths = sample_lengths(32, seed=14) + [65521, 65536]
    items = packed_items(lengths)
    data = packed_data(lengths, [1] * len(lengths), seed=12)
    batch = BatchFFT(queue.device, items)
    inp = Array.from_host(queue, data)
    out = Array.from_host(queue, numpy.zeros_like(data))
    batch(queue, out, inp)
    eager = out.get(queue)
    check_items(items, data, eager)
    # back-to-back launches without a host sync in between
    for _ in range(5):
        batch(queue, out, inp)
    assert numpy.array_equal(out.get(queue), eager)
    # one capture, several replays with eager work and syncs in between
    with queue.capture() as cap:
        batch(queue, out, inp)
    other = Array.from_host(queue, numpy.zeros_like(data))
    for _ in range(3):
        out.set(queue, numpy.zeros_like(data))
        cap.graph.launch(queue)
        assert numpy.array_equal(out.get(queue), eager)
        batch(queue, other, inp)
        queue.synchronize()
        assert numpy.array_equal(other.get(queue), eager)
    batch.check()


def test_autotune_small(queue):
    from reikna_amd.fft.autotune import autotune_batch

    lengths = sample_lengths(16, seed=15)
    items = packed_items(lengths)
    cands = [BatchOptions(), BatchOptions(single_max=1024, factor_max=64), BatchOptions(order=1)]
    best, results = autotune_batch(queue.device, items, queue, candidates=cands, reps=3, use_cache=False)
    assert len(results) == 3 and best in cands
    assert all(ms > 0 for _, ms in results)
    data = packed_data(lengths, [1] * len(lengths), seed=13)
    res, _ = run_batch(queue, items, data, data.size, options=best)
    check_items(items, data, res)


# ---------------------------------------------------------------------------
# real-to-complex items (numpy.fft.rfft) in the same launch as complex ones
# ---------------------------------------------------------------------------
R2C_LENGTHS = [4, 6, 10, 64, 1000, 1024, 1994, 6000, 8192, 24576, 131072]


def r2c_batch_data(items, seed):
    """One float32 buffer: complex items' values as (re, im) pairs, real
    items' values as they are (item offsets count complex64 elements)."""
    total = max(it.in_offset + (it.count - 1) * it.in_dist + it.in_length for it in items)
    rng = numpy.random.default_rng(seed)
    return rng.uniform(-1, 1, 2 * total).astype(numpy.float32)


def check_r2c_items(items, flat_in, host_out):
    cin = flat_in.view(numpy.complex64)
    worst = 0.0
    for it in items:
        for j in range(it.count):
            o = it.in_offset + j * it.in_dist
            y = host_out[it.out_offset + j * it.out_dist: it.out_offset + j * it.out_dist + it.out_length]
            if it.r2c:
                x = flat_in[2 * o: 2 * o + it.length].astype(numpy.float64)
                ref = numpy.fft.rfft(x)
            else:
                ref = oracle.fftn(cin[o: o + it.length], (0,))
            if it.scale:
                ref = ref / it.length
            err = oracle.rel_l2(y, ref)
            tol = oracle.tolerance((it.length,))
            assert err <= tol, f"length {it.length} r2c={it.r2c} (copy {j}): rel-L2 {err:.3g} > {tol:.3g}"
            worst = max(worst, err / tol)
    return worst


@pytest.mark.parametrize("scale", [False, True])
def test_r2c_items_in_mixed_batch(queue, scale):
    rng = numpy.random.default_rng(31)
    c2c = sample_lengths(12, seed=16) + [997, 65521]
    lengths = R2C_LENGTHS + c2c
    kinds = [True] * len(R2C_LENGTHS) + [False] * len(c2c)
    counts = [int(c) for c in rng.integers(1, 4, len(lengths))]
    items = packed_items(lengths, counts, scale=scale, r2c=kinds)
    flat = r2c_batch_data(items, seed=32)
    batch = BatchFFT(queue.device, items)
    inp = Array.from_host(queue, flat)   # float32 input
    n_out = max(it.out_offset + (it.count - 1) * it.out_dist + it.out_length for it in items)
    out = Array.from_host(queue, numpy.zeros(n_out, numpy.complex64))
    batch(queue, out, inp)
    res = out.get(queue)
    batch.check()
    check_r2c_items(items, flat, res)
    assert batch.info["launches"] == 1


def test_r2c_alone_equals_in_mix(queue):
    n = 24576
    alone_items = packed_items([n], [2], r2c=True)
    flat = r2c_batch_data(alone_items, seed=33)
    b1 = BatchFFT(queue.device, alone_items)
    inp = Array.from_host(queue, flat)
    out = Array.from_host(queue, numpy.zeros(2 * (n // 2 + 1), numpy.complex64))
    b1(queue, out, inp)
    alone = out.get(queue)
    # the same item at other offsets inside a mixed batch
    mix = packed_items([1000, n, 65521, 8192], [3, 2, 1, 2], r2c=[False, True, False, True])
    it = mix[1]
    flat2 = r2c_batch_data(mix, seed=34)
    flat2[2 * it.in_offset: 2 * it.in_offset + 2 * n] = flat[: 2 * n]
    b2 = BatchFFT(queue.device, mix)
    inp2 = Array.from_host(queue, flat2)
    n_out = max(i.out_offset + (i.count - 1) * i.out_dist + i.out_length for i in mix)
    out2 = Array.from_host(queue, numpy.zeros(n_out, numpy.complex64))
    b2(queue, out2, inp2)
    res = out2.get(queue)
    assert numpy.array_equal(res[it.out_offset: it.out_offset + 2 * (n // 2 + 1)], alone)
    check_r2c_items(mix, flat2, res)


def test_r2c_forward_only(queue):
    items = packed_items([64, 128], r2c=[True, False])
    batch = BatchFFT(queue.device, items)
    inp = Array.from_host(queue, r2c_batch_data(items, seed=35))
    out = Array.from_host(queue, numpy.zeros(200, numpy.complex64))
    with pytest.raises(Exception, match="forward only"):
        batch(queue, out, inp, 1)


def test_wait_give_up_is_reported_and_cleared(queue, monkeypatch):
    """A poll limit of 1 makes the second-pass tiles give up their waits (every
    wait is bounded: the kernel still ends); check() reports it once and
    clears the plan's flag and counters."""
    monkeypatch.setenv("RK_BATCH_SPIN_LIMIT", "1")
    items = packed_items([65536] * 8)
    batch = BatchFFT(queue.device, items)
    monkeypatch.delenv("RK_BATCH_SPIN_LIMIT")
    data = packed_data([65536] * 8, [1] * 8, seed=42)
    inp = Array.from_host(queue, data)
    out = Array.from_host(queue, numpy.zeros_like(data))
    batch(queue, out, inp)
    with pytest.raises(RuntimeError, match="poll limit"):
        batch.check()
    batch.check()   # cleared
    # a plan with the default limit is unaffected
    res, _ = run_batch(queue, items, data, data.size)
    check_items(items, data, res)


def test_stated_512_uneven_batch(queue):
    """The batched workload exactly as stated: 512 transforms, lengths
    log-uniform in [64, 65536] over 2^k, 3*2^k, 5*2^k and primes (seed 5),
    U[-1, 1) data (seed 6), fused 1/N, all in one launch; every transform
    against numpy.fft in FP64."""
    lengths = sample_lengths(512, 64, 65536, seed=5)
    assert len(lengths) == 512 and min(lengths) >= 64 and max(lengths) <= 65536
    items = packed_items(lengths, scale=True)
    data = oracle.test_array((sum(lengths),), seed=6)
    res, batch = run_batch(queue, items, data, data.size)
    assert batch.info["launches"] == 1
    worst = check_items(items, data, res)
    assert worst <= 1.0


def test_stated_512_uneven_batch_r2c(queue):
    """The same 512 lengths with the real-to-complex packing: every even
    length an r2c item (numpy.fft.rfft of its real input), the odd ones
    complex, fused 1/N, one launch; every transform checked."""
    lengths = sample_lengths(512, 64, 65536, seed=5)
    kinds = [n % 2 == 0 for n in lengths]
    items = packed_items(lengths, scale=True, r2c=kinds)
    flat = r2c_batch_data(items, seed=6)
    batch = BatchFFT(queue.device, items)
    inp = Array.from_host(queue, flat)
    n_out = max(it.out_offset + (it.count - 1) * it.out_dist + it.out_length for it in items)
    out = Array.from_host(queue, numpy.zeros(n_out, numpy.complex64))
    batch(queue, out, inp)
    res = out.get(queue)
    batch.check()
    assert batch.info["launches"] == 1
    assert sum(kinds) > 300 and sum(not k for k in kinds) > 50
    check_r2c_items(items, flat, res)
