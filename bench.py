"""Benchmark of the FFT hot path on one MI355X.

Default workload: complex64 C2C FFT of a 64 x 2^20 array over axis 1
(out-of-place, inputs resident in HBM, N(0,1) data from seed 3), the
configuration the headline GFLOP/s (5 N log2 N convention) is quoted on.

Prints ONE JSON line with the metric, GFLOP/s, ms per step, a roofline
object for the whole transform (algorithmic bytes: 16 B per complex element
per transform, one read + one write, however many passes the plan makes,
over the measured time per step; the dominant pass's own rate is reported
beside it as pass_GBs / pass_frac) and the CPU baseline (the FP64 oracle
library, scipy.fft / pocketfft, on a bounded sample of the same workload on
this host).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
"""

from __future__ import annotations

import argparse
import dataclasses
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "c64 C2C FFT GFLOP/s (5Nlog2N) + %HBM roofline; p99 latency mixed-size stream"
HBM_PEAK_GBS = 8000.0

WORKLOADS = {
    # name: (shape, axes, distribution, seed, description)
    "c2b": ((64, 2**20), (1,), "normal", 3, "c64 C2C FFT 64 x 2^20 over axis 1"),
    "c1": ((1024, 1024), (1,), "uniform", 1, "c64 C2C FFT 1024 x 1024 over axis 1, forward + inverse"),
    "c2a": ((256, 256, 256), (0, 1, 2), "normal", 2, "c64 C2C FFT 256^3 over axes (0, 1, 2)"),
    "c3-1000": ((4096, 1000), (1,), "uniform", 4, "c64 C2C FFT 4096 x 1000 over axis 1"),
    "c3-997": ((4096, 997), (1,), "uniform", 4, "c64 C2C FFT 4096 x 997 (prime, Bluestein) over axis 1"),
    "c3-12288": ((4096, 12288), (1,), "uniform", 4, "c64 C2C FFT 4096 x 12288 over axis 1"),
    # batched uneven lengths: handled by bench_batch
    "c4": (None, None, "uniform", 6, "512 c64 C2C transforms, lengths log-uniform in [64, 65536] "
           "(2^k, 3*2^k, 5*2^k, primes; seed 5), fused 1/N, one launch"),
    # request serving: handled by bench_serve
    "c5": (None, None, "normal", 8, "Poisson stream of c64 FFT requests: 1-D 2^8..2^20 and 2-D 128^2..2048^2, "
           "batch 1..64 (log-uniform), arrivals seed 7"),
    # the computations beside the FFT (transpose, real-input transforms, shift, spectrogram)
    "transpose": ((8192, 8192), None, "normal", 11, "c64 Transpose 8192 x 8192"),
    "rfft": ((64, 2**20), (1,), "normal", 12, "f32 RFFT 64 x 2^20 over axis 1"),
    "irfft": ((64, 2**19 + 1), (1,), "normal", 13, "c64 IRFFT 64 x (2^19 + 1) over axis 1"),
    "fftshift": ((8192, 8192), (0, 1), "normal", 14, "c64 FFTShift 8192 x 8192 in place"),
    "spectrogram": ((2**24,), None, "normal", 15, "f32 Spectrogram of 2^24 samples, 256-point frames, overlap 128"),
}
COMPUTATIONS = ("transpose", "rfft", "irfft", "fftshift", "spectrogram")
FWD_INV = ("c1",)   # a step is the forward and the inverse transform

# Memory traffic per step (every launch of one step; for c4 per launch of the
# batch kernel), from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs of this
# benchmark: FETCH_SIZE reads half of a copy kernel's known read size on this
# GPU and WRITE_SIZE all of it, so traffic = 2 * FETCH_SIZE + WRITE_SIZE (KiB
# per dispatch), summed over the step's dispatches.  These count the L2's
# requests to the memory fabric, reads served by the Infinity Cache included.
# Each entry names the profile summary it comes from.
MEASURED_TRAFFIC = {
    "c2b": (2.181e9, "profiles/fft_pass_c2b_64x2p20.md"),
    "c1": (3.407e7, "profiles/fft_pass_c1_1024x1024.md"),
    "c2a": (5.372e8, "profiles/fft_pass_c2a_256cubed.md"),
    "c3-1000": (6.581e7, "profiles/fft_pass_c3_4096x1000.md"),
    "c3-997": (6.629e7, "profiles/fft_pass_c3_4096x997_bluestein.md"),
    "c3-12288": (1.612e9, "profiles/fft_pass_c3_4096x12288.md"),
    "c4": (4.134e8, "profiles/fft_batch_c4_512_uneven.md"),
}


def cpu_baseline(shape, axes, data, budget_s: float = 12.0, fwd_inv: bool = False) -> dict:
    """The oracle library on this host's cores: scipy.fft (pocketfft) with one
    worker per core over a sample of the rows of the workload (forward and
    inverse when the workload's step is both)."""
    import numpy

    try:
        import scipy.fft as sfft

        workers = len(os.sched_getaffinity(0))
        fwd = lambda x: sfft.fftn(x, axes=axes, workers=workers)  # noqa: E731
        inv = lambda x: sfft.ifftn(x, axes=axes, workers=workers)  # noqa: E731
        lib = f"scipy.fft {__import__('scipy').__version__} (pocketfft)"
    except Exception:   # pragma: no cover
        workers = 1
        fwd = lambda x: numpy.fft.fftn(x, axes=axes)  # noqa: E731
        inv = lambda x: numpy.fft.ifftn(x, axes=axes)  # noqa: E731
        lib = f"numpy.fft {numpy.__version__} (pocketfft)"
    fn = (lambda x: inv(fwd(x))) if fwd_inv else fwd  # noqa: E731
    rows = shape[0] if 0 not in axes else shape[0]
    sample_rows = rows if 0 in axes else max(1, min(rows, (8 * 2**20) // max(1, math.prod(shape[1:]))))
    sample = data if 0 in axes else data[:sample_rows]
    times = []
    t_end = time.perf_counter() + budget_s
    while len(times) < 10 and (time.perf_counter() < t_end or len(times) < 2):
        t0 = time.perf_counter()
        fn(sample)
        times.append(time.perf_counter() - t0)
    t = min(times)
    flops = (2.0 if fwd_inv else 1.0) * 5.0 * sample.size * sum(math.log2(shape[a]) for a in axes)
    # numpy.fft (pocketfft, one thread) on a smaller slice of the same sample
    one = sample if 0 in axes else sample[: max(1, sample_rows // 4)]
    one_flops = (2.0 if fwd_inv else 1.0) * 5.0 * one.size * sum(math.log2(shape[a]) for a in axes)
    one_times = []
    t_end = time.perf_counter() + budget_s / 2
    while len(one_times) < 5 and (time.perf_counter() < t_end or len(one_times) < 1):
        t0 = time.perf_counter()
        y = numpy.fft.fftn(one, axes=axes)
        if fwd_inv:
            numpy.fft.ifftn(y, axes=axes)
        one_times.append(time.perf_counter() - t0)
    cpu_model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu_model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {
        "value": round(flops / t / 1e9, 2),
        "unit": "GFLOP/s",
        "cores": workers,
        "kind": "port",
        "sample": f"{lib}, {sample.shape} {sample.dtype} over axes {tuple(axes)} ({sample_rows} of {shape[0]} rows), "
        f"min of {len(times)} runs; host {cpu_model}, os.cpu_count()={os.cpu_count()}",
        "numpy_1_thread": {
            "value": round(one_flops / min(one_times) / 1e9, 2),
            "sample": f"numpy.fft {numpy.__version__} (pocketfft, one thread), {one.shape}, min of {len(one_times)} runs",
        },
    }


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # a few dozen untimed steps: the first hundred milliseconds of work on a
    # fresh process run ~6 % slower than the steady state that follows
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--workload", default="c2b", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--dtype", choices=("c64", "c128"), default="c64",
                    help="single transforms: complex64 (the configured workloads) or complex128")
    ap.add_argument("--autotune", action="store_true",
                    help="pick planner knobs by measurement (c4: batch knobs; single transforms: plan knobs)")
    ap.add_argument("--requests", type=int, default=1000, help="c5: requests per rate point")
    ap.add_argument("--rates", default="1000,2000,5000,10000,20000,50000", help="c5: offered rates (requests/s)")
    ap.add_argument("--small-rates", default="20000,50000,100000,200000,400000,800000",
                    help="c5: offered rates of the small-request stream")
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(
            f"bench.py runs on one GPU only: batched transforms shard over GPUs with no communication and "
            f"multi-axis ones with one all-to-all, neither of which is built here (asked for --gpus {args.gpus})",
            file=sys.stderr,
        )
        return 2

    import numpy

    import oracle
    from reikna_amd import Array, Event, Queue
    from reikna_amd.fft import FFT
    from reikna_amd.runtime import Device

    if args.workload == "c4":
        return bench_batch(args)
    if args.workload == "c5":
        return bench_serve(args)
    if args.workload in COMPUTATIONS:
        return bench_computation(args)
    shape, axes, dist, seed, desc = WORKLOADS[args.workload]
    dev = Device(0)
    queue = Queue(dev)
    print(f"[bench] {desc} on {dev.name} ({dev.arch}, {dev.params.compute_units} CUs)", file=sys.stderr, flush=True)
    data = oracle.test_array(shape, seed=seed, dist=dist)
    if args.dtype == "c128":
        data = data.astype(numpy.complex128)
        desc = desc.replace("c64 ", "c128 ", 1)
    arr_in = Array.from_host(queue, data)
    arr_out = Array.empty_like(dev, arr_in)
    tuning, tuned = None, []
    if args.autotune:
        from reikna_amd.fft.autotune import autotune_fft

        print("[bench] autotuning the plan ...", file=sys.stderr, flush=True)
        tuning, tuned = autotune_fft(queue, shape, axes, use_cache=False)
        print(f"[bench] chosen {tuning}", file=sys.stderr, flush=True)
    fft = FFT(arr_in, axes=axes, tuning=tuning)
    fftc = fft.compile(dev)
    native = fft.native_plan
    passes = native.passes
    print(f"[bench] passes: {[(p.L, p.C, p.load_mode, p.store_mode) for p in passes]}", file=sys.stderr, flush=True)

    fwd_inv = args.workload in FWD_INV
    arr_back = Array.empty_like(dev, arr_in) if fwd_inv else None

    def step():
        fftc(queue, arr_out, arr_in, 0)
        if fwd_inv:
            fftc(queue, arr_back, arr_out, 1)

    for _ in range(args.warmup):
        step()
    queue.synchronize()
    e0, e1 = Event(), Event()
    e0.record(queue)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    e1.record(queue)
    queue.synchronize()
    wall = time.perf_counter() - t0
    ms_total = e0.elapsed_ms(e1)
    ms_per_step = ms_total / args.steps

    # individually timed steps: min and median
    step_ms = []
    for _ in range(max(20, args.steps)):
        s0, s1 = Event(), Event()
        s0.record(queue)
        step()
        s1.record(queue)
        s1.synchronize()
        step_ms.append(s0.elapsed_ms(s1))
    step_ms.sort()

    # per-pass kernel durations, HIP events on the stream the kernels run on
    native.set_timing(True)
    per_pass = [[] for _ in passes]
    for _ in range(max(5, args.steps // 2)):
        fftc(queue, arr_out, arr_in, 0)
        for k, t in enumerate(native.pass_times_ms()):
            per_pass[k].append(t)
    native.set_timing(False)
    pass_ms = [sum(v) / len(v) for v in per_pass]
    dom = max(range(len(pass_ms)), key=lambda k: pass_ms[k])

    # device-to-device copy bandwidth of the same array on this box (reference point)
    for _ in range(3):
        arr_out.copy_from(queue, arr_in)
    queue.synchronize()
    c0, c1 = Event(), Event()
    c0.record(queue)
    ncopy = 10
    for _ in range(ncopy):
        arr_out.copy_from(queue, arr_in)
    c1.record(queue)
    queue.synchronize()
    copy_gbs = 2.0 * arr_in.nbytes * ncopy / (c0.elapsed_ms(c1) / 1e3) / 1e9
    step()

    # correctness spot check against the FP64 oracle
    res = arr_out.get(queue)
    if 0 in axes:
        err = oracle.rel_l2(res, oracle.fftn(data, axes))
    else:
        rows = [0, shape[0] - 1]
        err = max(oracle.rel_l2(res[r], oracle.fftn(data[r], [a - 1 for a in axes])) for r in rows)
    tol = oracle.tolerance(shape, axes)
    if args.dtype == "c128":   # the complex128 bound of the GPU tests
        tol = 5e-16 * max(1.0, math.log2(math.prod(shape[a] for a in axes)))
    if fwd_inv:   # the inverse of the forward result gives the input back
        err = max(err, oracle.rel_l2(arr_back.get(queue), data) / 2)

    elems = math.prod(shape)
    directions = 2 if fwd_inv else 1
    flops = directions * oracle.fft_flops(shape, axes)
    gflops = flops / (ms_per_step / 1e3) / 1e9
    # algorithmic bytes: one read and one write of every element per
    # transform, however many passes the plan makes
    alg_bytes = 2.0 * data.dtype.itemsize * elems
    transform_gbs = directions * alg_bytes / (ms_per_step / 1e3) / 1e9
    # the same bytes over the dominant pass's time: what one pass streams
    pass_gbs = alg_bytes / (pass_ms[dom] / 1e3) / 1e9
    traffic = MEASURED_TRAFFIC.get(args.workload) if args.dtype == "c64" else None
    out = {
        "metric": METRIC,
        "value": round(gflops, 1),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": args.dtype,
        "data": "synthetic",
        "config": {
            "workload": desc,
            "shape": list(shape),
            "axes": list(axes),
            "distribution": f"{dist}, seed {seed}",
            "directions": "forward + inverse" if fwd_inv else "forward",
            "passes": [{"L": p.L, "C": p.C, "load": p.load_mode, "store": p.store_mode, "nt_load": p.nt_load,
                        "nt_store": p.nt_store, "wave": p.wave, "pipe": p.pipe, "jit": p.jit} for p in passes],
            "chunk_rows": native.chunk_rows,
            **({"tuning": dataclasses.asdict(tuning),
                "autotune": [[dataclasses.asdict(t), round(ms, 4)] for t, ms in tuned]} if args.autotune else {}),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(transform_gbs, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(transform_gbs / HBM_PEAK_GBS, 4),
            "traffic": traffic[0] if traffic else None,
            "traffic_per": "step (all passes)",
            "traffic_source": traffic[1] if traffic else None,
            "kernel": f"whole transform: {len(passes)} pass(es), {2 * data.dtype.itemsize} B per element per transform over ms_per_step",
            "dominant_pass": f"pass {dom} (L={passes[dom].L}, C={passes[dom].C}), {pass_ms[dom] * 1e3:.1f} us avg",
            "pass_GBs": round(pass_gbs, 1),
            "pass_frac": round(pass_gbs / HBM_PEAK_GBS, 4),
            "pass_us": [round(t * 1e3, 2) for t in pass_ms],
            "d2d_copy_GBs": round(copy_gbs, 1),
            "frac_of_copy": round(transform_gbs / copy_gbs, 4),
        },
        "step_ms_min": round(step_ms[0], 5),
        "step_ms_median": round(step_ms[len(step_ms) // 2], 5),
        "rel_l2": err,
        "rel_l2_bound": tol,
        "wall_s": round(wall, 4),
    }
    if not args.no_cpu:
        print("[bench] CPU baseline ...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(shape, axes, data, fwd_inv=fwd_inv)
    print(json.dumps(out), flush=True)
    return 0 if err <= tol else 1


def bench_batch(args) -> int:
    """512 transforms of mixed lengths in one launch of the batch kernel,
    against the per-item path (one plan and its launches per transform)."""
    import numpy

    import oracle
    from reikna_amd import Array, Event, Queue
    from reikna_amd.fft import BatchFFT, BatchOptions
    from reikna_amd.fft.batch import packed_items, sample_lengths
    from reikna_amd.runtime import Device

    _, _, dist, seed, desc = WORKLOADS["c4"]
    dev = Device(0)
    queue = Queue(dev)
    lengths = sample_lengths(512, 64, 65536, seed=5)
    items = packed_items(lengths, scale=True)
    total = sum(lengths)
    data = oracle.test_array((total,), seed=seed, dist=dist)
    inp = Array.from_host(queue, data)
    out = Array.from_host(queue, numpy.zeros_like(data))
    print(f"[bench] {desc} on {dev.name}: {total} elements", file=sys.stderr, flush=True)
    opts = BatchOptions()
    if args.autotune:
        from reikna_amd.fft.autotune import autotune_batch

        opts, _ = autotune_batch(dev, items, queue)
        print(f"[bench] autotuned options: {opts}", file=sys.stderr, flush=True)
    batch = BatchFFT(dev, items, opts)
    info = batch.info

    def timed(fn, steps):
        for _ in range(args.warmup):
            fn()
        queue.synchronize()
        e0, e1 = Event(), Event()
        e0.record(queue)
        for _ in range(steps):
            fn()
        e1.record(queue)
        queue.synchronize()
        return e0.elapsed_ms(e1) / steps

    # the batch kernel is the only launch of a step, so the timed region's
    # average (HIP events on the launch stream) is its average duration
    ms = timed(lambda: batch(queue, out, inp), args.steps)
    kernel_ms = ms
    # one launch alone, synchronised (includes the launch latency)
    single = []
    for _ in range(max(5, args.steps // 2)):
        e0, e1 = Event(), Event()
        e0.record(queue)
        batch(queue, out, inp)
        e1.record(queue)
        e1.synchronize()
        single.append(e0.elapsed_ms(e1))
    single_ms = sorted(single)[len(single) // 2]
    batch.check()
    res = out.get(queue)
    worst = 0.0
    for it in items:
        x = data[it.in_offset: it.in_offset + it.length]
        ref = oracle.fftn(x, (0,)) / it.length
        worst = max(worst, oracle.rel_l2(res[it.out_offset: it.out_offset + it.length], ref)
                    / oracle.tolerance((it.length,)))
    # HIP graph replay of the same launch
    with queue.capture() as cap:
        batch(queue, out, inp)
    graph_ms = timed(lambda: cap.graph.launch(queue), args.steps)
    r2c = bench_batch_r2c(queue, lengths, timed, args.steps)
    # the per-item path
    grouped = BatchFFT(dev, items, BatchOptions(mode="grouped"))
    grouped_ms = timed(lambda: grouped(queue, out, inp), max(3, args.steps // 4))

    flops = sum(5.0 * n * math.log2(n) for n in lengths)
    alg_bytes = 16.0 * total
    achieved = alg_bytes / (kernel_ms / 1e3) / 1e9
    kname = f"rk_fft_batch_{('light', 'blue')[info['flavour']]}"
    traffic, traffic_source = MEASURED_TRAFFIC.get("c4", (None, None))
    out_json = {
        "metric": METRIC,
        "value": round(flops / (ms / 1e3) / 1e9, 1),
        "unit": "GFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64",
        "data": "synthetic",
        "config": {
            "workload": desc,
            "transforms": len(lengths),
            "elements": total,
            "distribution": f"{dist}, seed {seed}",
            "options": {k: getattr(opts, k) for k in ("single_max", "factor_max", "blue_max", "order")},
            "plan": info,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "traffic_per": "launch",
            "traffic_source": traffic_source,
            "kernel": f"{kname}, {kernel_ms * 1e3:.1f} us avg",
            "traffic_GBs": round(traffic / (kernel_ms / 1e3) / 1e9, 1) if traffic else None,
        },
        "single_launch_ms": round(single_ms, 5),
        "r2c_mix": r2c,
        "graph_ms_per_step": round(graph_ms, 5),
        "grouped_ms_per_step": round(grouped_ms, 5),
        "speedup_vs_grouped": round(grouped_ms / ms, 2),
        "worst_rel_l2_over_bound": round(worst, 4),
    }
    if not args.no_cpu:
        print("[bench] CPU baseline ...", file=sys.stderr, flush=True)
        import scipy.fft as sfft

        workers = len(os.sched_getaffinity(0))
        times = []
        for _ in range(5):
            t0 = time.perf_counter()
            for it in items:
                sfft.fft(data[it.in_offset: it.in_offset + it.length], norm="forward", workers=workers)
            times.append(time.perf_counter() - t0)
        out_json["cpu_baseline"] = {
            "value": round(flops / min(times) / 1e9, 2),
            "unit": "GFLOP/s",
            "cores": workers,
            "kind": "port",
            "sample": f"scipy.fft {__import__('scipy').__version__} (pocketfft), all 512 transforms one by one, "
            f"min of {len(times)} runs",
        }
    print(json.dumps(out_json), flush=True)
    return 0 if worst <= 1.0 and r2c["worst_rel_l2_over_bound"] <= 1.0 else 1


def bench_computation(args) -> int:
    """One of the computations beside the FFT at a size that streams through
    HBM: the whole computation (all its launches) timed with HIP events,
    checked against numpy in FP64, priced by its algorithmic bytes."""
    import numpy

    from reikna_amd import Array, Event, Queue
    from reikna_amd.algorithms import Transpose
    from reikna_amd.fft import FFTShift
    from reikna_amd.fft.rfft import IRFFT, RFFT
    from reikna_amd.fft.spectrogram import Spectrogram
    from reikna_amd.runtime import Device

    name = args.workload
    shape, axes, dist, seed, desc = WORKLOADS[name]
    dev = Device(0)
    queue = Queue(dev)
    rng = numpy.random.default_rng(seed)
    print(f"[bench] {desc} on {dev.name}", file=sys.stderr, flush=True)
    flops = None
    inplace = False
    if name == "transpose":
        host = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(numpy.complex64)
        comp = Transpose(host)
        ref_fn = lambda x: x.T  # noqa: E731
        out_shape, out_dtype = shape[::-1], numpy.complex64
        alg_bytes = 16.0 * host.size
        tol = 0.0
    elif name == "rfft":
        host = rng.standard_normal(shape).astype(numpy.float32)
        comp = RFFT(host)
        ref_fn = lambda x: numpy.fft.rfft(x.astype(numpy.float64), axis=-1)  # noqa: E731
        out_shape, out_dtype = (shape[0], shape[1] // 2 + 1), numpy.complex64
        alg_bytes = 4.0 * host.size + 8.0 * shape[0] * (shape[1] // 2 + 1)
        flops = 2.5 * host.size * math.log2(shape[1])
        tol = 2e-7 * math.log2(shape[1])
    elif name == "irfft":
        n = (shape[1] - 1) * 2
        real = rng.standard_normal((shape[0], n))
        host = numpy.fft.rfft(real, axis=-1).astype(numpy.complex64)
        comp = IRFFT(host)
        ref_fn = lambda x: numpy.fft.irfft(x.astype(numpy.complex128), axis=-1)  # noqa: E731
        out_shape, out_dtype = (shape[0], n), numpy.float32
        alg_bytes = 8.0 * host.size + 4.0 * shape[0] * n
        flops = 2.5 * shape[0] * n * math.log2(n)
        tol = 2e-7 * math.log2(n)
    elif name == "fftshift":
        host = (rng.standard_normal(shape) + 1j * rng.standard_normal(shape)).astype(numpy.complex64)
        comp = FFTShift(host, axes=axes)
        ref_fn = lambda x: numpy.fft.fftshift(x, axes=axes)  # noqa: E731
        out_shape, out_dtype = shape, numpy.complex64
        alg_bytes = 16.0 * host.size
        tol = 0.0
        inplace = True
    else:   # spectrogram
        host = rng.standard_normal(shape).astype(numpy.float32)
        comp = Spectrogram(host, fft_size=256, noverlap=128)
        frames = (shape[0] - 128) // 128
        out_shape, out_dtype = (129, frames), numpy.float32

        def ref_fn(x):
            x = x.astype(numpy.float64)
            idx = numpy.arange(256)[None, :] + 128 * numpy.arange(frames)[:, None]
            win = 0.5 * (1 - numpy.cos(2 * numpy.pi * numpy.arange(256) / 255))
            return numpy.abs(numpy.fft.rfft(x[idx] * win, axis=-1)).T

        alg_bytes = 4.0 * host.size + 4.0 * 129 * frames
        flops = 5.0 * frames * 256 * 8
        tol = 2e-7 * 8
    compiled = comp.compile(dev)
    inp = Array.from_host(queue, host)
    out = inp if inplace else Array.empty(dev, out_shape, out_dtype)

    def step():
        compiled(queue, out, inp)

    for _ in range(args.warmup):
        step()
    queue.synchronize()
    e0, e1 = Event(), Event()
    e0.record(queue)
    for _ in range(args.steps):
        step()
    e1.record(queue)
    queue.synchronize()
    ms = e0.elapsed_ms(e1) / args.steps
    # check: one more call from a fresh input
    inp.set(queue, host)
    step()
    res = out.get(queue)
    ref = ref_fn(host)
    if tol == 0.0:
        err = 0.0 if numpy.array_equal(res, ref.astype(res.dtype)) else float("inf")
    else:
        err = float(numpy.linalg.norm(res - ref) / numpy.linalg.norm(ref))
    achieved = alg_bytes / (ms / 1e3) / 1e9
    out_json = {
        "metric": METRIC,
        "value": round(flops / (ms / 1e3) / 1e9, 1) if flops else round(achieved, 1),
        "unit": "GFLOP/s" if flops else "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64" if out_dtype == numpy.complex64 or host.dtype == numpy.complex64 else "f32",
        "data": "synthetic",
        "config": {"workload": desc, "shape": list(shape), "distribution": f"{dist}, seed {seed}"},
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None,
            "kernel": "whole computation (all its launches)",
        },
        "rel_l2": err,
        "rel_l2_bound": tol,
    }
    print(json.dumps(out_json), flush=True)
    return 0 if err <= tol else 1


def bench_batch_r2c(queue, lengths, timed, steps) -> dict:
    """The same lengths as one launch with every even length as a
    real-to-complex item (numpy.fft.rfft of real input, scaled by 1/N) and the
    odd ones complex; FLOPs by the real-FFT convention 2.5 N log2 N."""
    import numpy

    from reikna_amd import Array
    from reikna_amd.fft import BatchFFT
    from reikna_amd.fft.batch import packed_items

    import oracle

    kinds = [n % 2 == 0 for n in lengths]
    items = packed_items(lengths, scale=True, r2c=kinds)
    n_in = max(it.in_offset + it.in_length for it in items)
    n_out = max(it.out_offset + it.out_length for it in items)
    rng = numpy.random.default_rng(6)
    flat = rng.uniform(-1, 1, 2 * n_in).astype(numpy.float32)
    batch = BatchFFT(queue.device, items)
    inp = Array.from_host(queue, flat)
    out = Array.from_host(queue, numpy.zeros(n_out, numpy.complex64))
    ms = timed(lambda: batch(queue, out, inp), steps)
    batch.check()
    res = out.get(queue)
    cin = flat.view(numpy.complex64)
    worst = 0.0
    for it in items:
        y = res[it.out_offset: it.out_offset + it.out_length]
        if it.r2c:
            ref = numpy.fft.rfft(flat[2 * it.in_offset: 2 * it.in_offset + it.length].astype(numpy.float64))
        else:
            ref = oracle.fftn(cin[it.in_offset: it.in_offset + it.length], (0,))
        worst = max(worst, oracle.rel_l2(y, ref / it.length) / oracle.tolerance((it.length,)))
    flops = sum((2.5 if k else 5.0) * n * math.log2(n) for n, k in zip(lengths, kinds))
    return {
        "r2c_items": int(sum(kinds)),
        "c2c_items": int(len(kinds) - sum(kinds)),
        "ms_per_step": round(ms, 5),
        "GFLOPs": round(flops / (ms / 1e3) / 1e9, 1),
        "plan": batch.info,
        "worst_rel_l2_over_bound": round(worst, 4),
    }


def bench_serve(args) -> int:
    """A Poisson request stream at several rates, through the continuous-
    batching server and through the plain per-request path (same admission,
    one plan execution per request)."""
    import numpy

    import oracle
    from reikna_amd import Queue
    from reikna_amd.fft.serve import FFTServer, make_specs, run_stream, sample_shapes
    from reikna_amd.runtime import Device

    _, _, dist, seed, desc = WORKLOADS["c5"]
    dev = Device(0)
    queue = Queue(dev)
    n_req = max(50, args.requests)
    shapes = sample_shapes(n_req, seed=7)
    specs = make_specs(queue, shapes, seed=seed)
    rates = [float(r) for r in args.rates.split(",")]
    print(f"[bench] {desc}: {n_req} requests per rate, rates {rates}", file=sys.stderr, flush=True)
    budget = 16 * 2**30
    plain_min = 2**22
    servers = {
        "batched": FFTServer(dev, budget, slots=4, mode="batched", plain_min_elements=plain_min),
        "plain": FFTServer(dev, budget, slots=4, mode="plain"),
    }
    # warm up: every shape's template / plan built, kernels loaded
    for srv in servers.values():
        run_stream(srv, specs, rate=1e6, seed=1)
    # correctness: a sample of the batched server's outputs against the oracle
    host_inputs = {}
    worst = 0.0
    sample_ids = set(range(0, n_req, max(1, n_req // 24)))

    def check(c, spec):
        nonlocal worst
        if c.id not in sample_ids:
            return
        key = id(spec.input)
        if key not in host_inputs:
            host_inputs[key] = spec.input.get(queue)
        x = host_inputs[key][: spec.elements].reshape(spec.count, *spec.shape)
        ref = oracle.fftn(x, tuple(range(1, 1 + len(spec.shape))))
        out = servers["batched"].output(c, spec).get(queue)
        worst = max(worst, oracle.rel_l2(out, ref) / oracle.tolerance(spec.shape))

    run_stream(servers["batched"], specs, rate=1e6, seed=2, on_complete=check)
    sweep = {}
    for mode, srv in servers.items():
        rows = []
        for r in rates:
            res = run_stream(srv, specs, rate=r, seed=7)
            rows.append({"rate": r, "throughput_rps": round(res.throughput_rps, 1), "gflops": round(res.gflops, 1),
                         "p50_ms": round(res.p50_ms, 3), "p99_ms": round(res.p99_ms, 3),
                         "launches": res.launches})
            print(f"[bench] {mode:7s} rate {r:8.0f}/s -> {res.throughput_rps:8.1f} req/s, {res.gflops:8.1f} GFLOP/s, "
                  f"p50 {res.p50_ms:.3f} ms, p99 {res.p99_ms:.3f} ms, {res.launches} launches",
                  file=sys.stderr, flush=True)
        sweep[mode] = rows
    best = {m: max(rows, key=lambda x: x["throughput_rps"]) for m, rows in sweep.items()}
    total_flops = sum(sp.flops for sp in specs)
    # the mix's HBM bound: one read and one write of every request's elements
    # at 8 TB/s is the least time any path can take for the stream
    mix_bytes = 16.0 * sum(sp.elements for sp in specs)
    bound_rps = n_req / (mix_bytes / (HBM_PEAK_GBS * 1e9))
    # the same servers on a stream of small requests (1-D 2^6..2^12, 2-D
    # 32^2..128^2, batch 1..16), where per-request launches dominate the
    # plain path
    small_specs = make_specs(queue, sample_shapes(n_req, seed=9, one_d=(6, 12), two_d=(5, 7), max_count=16), seed=10)
    for srv in servers.values():
        run_stream(srv, small_specs, rate=1e6, seed=1)
    small_rates = [float(r) for r in args.small_rates.split(",")]
    small = {}
    for mode, srv in servers.items():
        rows = []
        for r in small_rates:
            res = run_stream(srv, small_specs, rate=r, seed=11)
            rows.append({"rate": r, "throughput_rps": round(res.throughput_rps, 1), "gflops": round(res.gflops, 1),
                         "p50_ms": round(res.p50_ms, 3), "p99_ms": round(res.p99_ms, 3),
                         "launches": res.launches})
            print(f"[bench] small {mode:7s} rate {r:8.0f}/s -> {res.throughput_rps:8.1f} req/s, "
                  f"p50 {res.p50_ms:.3f} ms, p99 {res.p99_ms:.3f} ms, {res.launches} launches",
                  file=sys.stderr, flush=True)
        small[mode] = rows
    small_best = {m: max(rows, key=lambda x: x["throughput_rps"]) for m, rows in small.items()}
    out = {
        "metric": METRIC,
        "value": best["batched"]["throughput_rps"],
        "unit": "requests/s",
        "n_gpus": 1,
        "steps": len(rates),
        "warmup": 1,
        "ms_per_step": round(1e3 / best["batched"]["throughput_rps"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64",
        "data": "synthetic",
        "config": {
            "workload": desc,
            "requests_per_rate": n_req,
            "mean_elements_per_request": round(sum(sp.elements for sp in specs) / n_req),
            "mean_gflop_per_request": round(total_flops / n_req / 1e9, 4),
            "budget_bytes": budget,
            "slots": 4,
            "plain_min_elements": plain_min,
        },
        "sweep": sweep,
        "best": best,
        "speedup_vs_plain": round(best["batched"]["throughput_rps"] / best["plain"]["throughput_rps"], 2),
        "hbm_bound": {
            "mix_bytes": mix_bytes,
            "bound_rps": round(bound_rps, 1),
            "batched_frac_of_bound": round(best["batched"]["throughput_rps"] / bound_rps, 4),
            "plain_frac_of_bound": round(best["plain"]["throughput_rps"] / bound_rps, 4),
            "max_speedup_over_plain": round(bound_rps / best["plain"]["throughput_rps"], 2),
        },
        "small_mix": {
            "workload": "1-D 2^6..2^12 and 2-D 32^2..128^2, batch 1..16 (log-uniform), arrivals seed 11",
            "sweep": small,
            "best": small_best,
            "speedup_vs_plain": round(small_best["batched"]["throughput_rps"] / small_best["plain"]["throughput_rps"],
                                      2),
        },
        "worst_rel_l2_over_bound": round(worst, 4),
    }
    for srv in servers.values():
        srv.close()
    print(json.dumps(out), flush=True)
    return 0 if worst <= 1.0 else 1


if __name__ == "__main__":
    sys.exit(main())
