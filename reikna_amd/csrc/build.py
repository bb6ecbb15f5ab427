"""Build libreikna_amd.so: the gfx950 FFT kernels and the C-ABI host library.

Steps:
  1. generate the ahead-of-time kernel catalog (catalog_partK.hip): one
     fft_pass_body<L, C, V> instantiation per (length, tile width, variant)
     entry, spread over several translation units so hipcc runs in parallel;
  2. embed fft_args.h / fft_device.hpp as string literals for the run-time
     compiler (hiprtc) that covers lengths outside the catalog and fused
     transformations;
  3. compile everything with hipcc --offload-arch=gfx950 and link one shared
     library next to the Python package.

Usage: python -m reikna_amd.csrc.build [--jobs N] [--incremental] [--arch gfx950]
"""

from __future__ import annotations

import argparse
import concurrent.futures
import hashlib
import os
import shutil
import subprocess
import sys
import time
from pathlib import Path

HERE = Path(__file__).resolve().parent
PKG = HERE.parent
REPO = PKG.parent
BUILD = REPO / "_build"
LIB = PKG / "libreikna_amd.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# Packed FP32 instructions are disabled for all device code.
DEVICE_FLAGS = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]
COMMON_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wno-unused-result"]

N_PARTS = 12


# ---------------------------------------------------------------------------
# catalog: mirrors make_radix_list / lds_pitch / tile_c_row / tile_c_col
# ---------------------------------------------------------------------------
def radix_list(length: int) -> list[int]:
    x, p2, n3, n5, n7 = length, 0, 0, 0, 0
    while x % 2 == 0:
        x //= 2
        p2 += 1
    while x % 3 == 0:
        x //= 3
        n3 += 1
    while x % 5 == 0:
        x //= 5
        n5 += 1
    while x % 7 == 0:
        x //= 7
        n7 += 1
    r: list[int] = []
    while p2 >= 4:
        r.append(16)
        p2 -= 4
    if p2 > 0:
        r.append(1 << p2)
    r += [7] * n7 + [5] * n5 + [3] * n3
    if not r:
        r = [1]
    imax = max(range(len(r)), key=lambda k: (r[k], -k))
    r[0], r[imax] = r[imax], r[0]
    return r


def lds_pitch(length: int) -> int:
    s = (length - 1) + ((length - 1) >> 4) + 1
    return s + 1 if s % 2 == 0 else s


def tile_pitch(length: int, c: int) -> int:
    """Row pitch of a tile (fft_device.hpp lds_pitch_for)."""
    split = c * lds_pitch(length) * 8 > 80 * 1024
    pow2 = c > 0 and c & (c - 1) == 0
    if not pow2 or c < 2 or threads_per_instance(length) < 16 or c > (32 if split else 16):
        return lds_pitch(length)
    m = 64 if split else 32
    p = (length - 1) + ((length - 1) >> 4) + 1
    while p % m != m // c:
        p += 1
    return p


def tile_lds_bytes(length: int, c: int) -> int:
    """LDS of a tile: the exchanges go through halves (floats) when the tile
    needs more than 80 KB at the base pitch."""
    split = c * lds_pitch(length) * 8 > 80 * 1024
    return c * tile_pitch(length, c) * (4 if split else 8)


def threads_per_instance(length: int) -> int:
    return length // radix_list(length)[0]


def tile_c_row(length: int) -> int:
    t = threads_per_instance(length)
    c = max(1, 256 // t)
    while c > 1 and c * lds_pitch(length) * 8 > 64 * 1024:
        c //= 2
    return c


def tile_c_col(length: int) -> int:
    t = threads_per_instance(length)
    c = max(16, (tile_c_row(length) // 16) * 16)
    while c > 1 and (c * t > 1024 or c * lds_pitch(length) * 8 > 159 * 1024):
        c //= 2
    return c


def catalog_lengths() -> list[int]:
    ls = set()
    ls.update(2**k for k in range(1, 15))
    ls.update(3 * 2**k for k in range(0, 13))
    ls.update(5 * 2**k for k in range(0, 12))
    ls.update(125 * 2**k for k in range(0, 8))
    return sorted(x for x in ls if x <= 16384)


# kernel variants: pre-op | post-op << 2 | bluestein << 4 (fft_device.hpp, Variant)
VAR_PLAIN = 0
VAR_TWIDDLE = 1                      # passes after the first of a multi-pass chain
VAR_CHIRP = 2                        # first pass of a multi-pass Bluestein chain
VAR_TWIDDLE_KMUL = 1 | (1 << 2)      # last pass of the forward Bluestein chain
VAR_TWIDDLE_CONJ_CHIRP = 1 | (2 << 2)  # last pass of the second Bluestein chain
VAR_BLUE = 2 | (2 << 2) | (1 << 4)   # single-kernel Bluestein
VAR_STREAM = 1 << 7                  # catalogue kernel with nontemporal loads/stores (StreamIO)
STREAM_MAX = 4096                    # streaming kernels for plain/twiddle passes up to this length
BLUE_SINGLE_MAX = 8192               # longest single-kernel Bluestein convolution (rk_fft_plan.cpp)
VAR_MODE_DD = 1 << 8                 # load/store modes DIRECT/DIRECT compiled in (RK_VAR_MODE_DD)
VAR_MODE_DF = 1 << 9                 # load/store modes DIRECT/FLAT compiled in (RK_VAR_MODE_DF)
VAR_MODE_FF = 1 << 10                # load/store modes FLAT/FLAT compiled in (RK_VAR_MODE_FF)
VAR_NT_LOAD = 1 << 11                # nontemporal loads only (RK_VAR_NT_LOAD)
VAR_NT_STORE = 1 << 12               # nontemporal stores only (RK_VAR_NT_STORE)
VAR_OCC = 1 << 16                    # launch bounds for the LDS-allowed residency (RK_VAR_OCC)
FACTOR_MAX = 1024                    # longest pass of a complex64 multi-pass chain (rk_fft_plan.cpp factor_max)


def catalog_entries() -> list[tuple[int, int, int]]:
    ents = set()
    for length in catalog_lengths():
        ents.add((length, tile_c_row(length), VAR_PLAIN))
        ents.add((length, tile_c_col(length), VAR_PLAIN))
        if 8 <= length <= BLUE_SINGLE_MAX:
            ents.add((length, tile_c_row(length), VAR_BLUE))
        if 16 <= length <= 4096:
            for v in (VAR_TWIDDLE, VAR_CHIRP, VAR_TWIDDLE_KMUL, VAR_TWIDDLE_CONJ_CHIRP):
                ents.add((length, tile_c_col(length), v))
    # streaming twins of the passes of large transforms
    for length, c, v in list(ents):
        if length <= STREAM_MAX and v in (VAR_PLAIN, VAR_TWIDDLE):
            ents.add((length, c, v | VAR_STREAM))
    # chains run in chunks through the Infinity Cache (rk_fft_plan.cpp): the
    # first pass of a chain loads the caller's input with the nontemporal
    # hint and stores plainly, the last pass loads plainly and stores the
    # caller's output with the hint
    for length, c, v in list(ents):
        if length <= FACTOR_MAX and c == tile_c_col(length):
            if v == VAR_PLAIN:
                ents.add((length, c, v | VAR_NT_LOAD))
            elif v == VAR_TWIDDLE:
                ents.add((length, c, v | VAR_NT_STORE))
    # the column passes of large multi-pass transforms with their modes
    # compiled in: the first pass loads columns and stores the tile flat (or
    # in columns), the later passes load and store columns.  Only lengths
    # with a factor 3, 5 or 7: their tiles are register-bound (146 VGPRs with
    # every mode, 70-83 with two) and run 2-5 % faster with the modes fixed;
    # power-of-two tiles are LDS-bound at four waves per SIMD either way and
    # measured 1-2 % slower with them.
    # The same for the plain (non-streaming) column passes of such lengths:
    # strided axes of multi-axis transforms and chains below the cache size.
    # Power-of-two passes of chunked chains (the 64 x 2^20 chain) got the
    # same twins once their element addressing was per-thread walks: the
    # 128-point middle pass of 64 x 2^20 takes 41.7 us per 128 MiB chunk with
    # DIRECT/DIRECT compiled in against 44.8 us generic (its pattern copy
    # 38.8 us).
    nt = VAR_STREAM | VAR_NT_LOAD | VAR_NT_STORE
    for length, c, v in list(ents):
        pow2 = length & (length - 1) == 0
        if pow2 and 16 <= length <= FACTOR_MAX and c == tile_c_col(length) and v & ~nt in (VAR_PLAIN, VAR_TWIDDLE):
            if v & ~nt == VAR_PLAIN:
                ents.add((length, c, v | VAR_MODE_DF))
                ents.add((length, c, v | VAR_MODE_DD))
            else:
                ents.add((length, c, v | VAR_MODE_DD))
    for length, c, v in list(ents):
        pow2 = length & (length - 1) == 0
        if not pow2 and length <= FACTOR_MAX and c == tile_c_col(length) and v & ~nt in (VAR_PLAIN, VAR_TWIDDLE):
            if v & ~nt == VAR_PLAIN:
                ents.add((length, c, v | VAR_MODE_DF))
                ents.add((length, c, v | VAR_MODE_DD))
            else:
                ents.add((length, c, v | VAR_MODE_DD))
    # the 16 x 1024 column tiles of two-pass chains (work-groups of 1024
    # threads) with DIRECT/DIRECT or DIRECT/FLAT compiled in run with launch
    # bounds for two work-groups per CU (rk_fft_plan.cpp occ_default)
    for length, c, v in list(ents):
        base = v & ~(nt | VAR_MODE_DD | VAR_MODE_DF | VAR_MODE_FF)
        if (length <= 1024 and c * threads_per_instance(length) == 1024 and base in (VAR_PLAIN, VAR_TWIDDLE)
                and v & (VAR_MODE_DD | VAR_MODE_DF)):
            ents.add((length, c, v | VAR_OCC))
    # single-pass row tiles of such lengths (one thread per R0 elements of an
    # instance, at least 32 of them: shorter rows are streamed flat): the
    # 1000-point tile holds 80 VGPRs with every mode and 62 with DIRECT/DIRECT
    for length in catalog_lengths():
        if length & (length - 1) and length <= STREAM_MAX and threads_per_instance(length) >= 32:
            ents.add((length, tile_c_row(length), VAR_PLAIN | VAR_MODE_DD))
    # short rows of such lengths (under 32 threads per instance) stream their
    # tiles flat on both sides: the 96-, 160- and 192-point row tiles hold
    # 140-160 VGPRs with every mode and 120-124 with FLAT/FLAT
    for length in catalog_lengths():
        if length & (length - 1) and length >= 8 and threads_per_instance(length) < 32:
            ents.add((length, tile_c_row(length), VAR_PLAIN | VAR_MODE_FF))
    # single-kernel Bluestein tiles up to M = 1024 with DIRECT/DIRECT compiled
    # in and bounded to three waves per SIMD (at most 168 VGPRs, a few
    # spilled): 8192 x 251 (M = 512) 24.1 against 28.5 us, 4096 x 509 and
    # 16384 x 61 4 % faster; from M = 2048 the bound costs more than the
    # occupancy gives (4096 x 997 69 against 58 us), so those keep the generic
    # kernel
    for length, c, v in list(ents):
        if v == VAR_BLUE and length <= 1024 and c * threads_per_instance(length) <= 256:
            ents.add((length, c, v | VAR_MODE_DD))
    return sorted(ents)


# persistent batch kernel: fixed work-group size, one case per (L, C, V)
BATCH_THREADS = 256
BATCH_SINGLE_MAX = 4096
BATCH_FACTOR_MAX = 1024


def batch_lengths() -> list[int]:
    ls = {b * 2**k for b in (1, 3, 5) for k in range(0, 13)}
    return sorted(x for x in ls if 2 <= x <= BATCH_SINGLE_MAX)


def batch_tile_c(length: int) -> int:
    return max(1, BATCH_THREADS // threads_per_instance(length))


VAR_CHIRP_KMUL = 2 | (1 << 2)        # single-pass first Bluestein chain
VAR_CONJ_CHIRP = 2 << 2               # single-pass second Bluestein chain


BATCH_BLUE_MAX = 2048


def batch_entries(blue: bool = True) -> list[tuple[int, int, int]]:
    """Tile cases of the batch kernel.  Bluestein runs over power-of-two
    lengths: as one single-kernel pass up to BATCH_BLUE_MAX in the blue
    flavour, else as two chains (one pass each up to BATCH_SINGLE_MAX, passes
    of 16..BATCH_FACTOR_MAX above); multi-pass chains use pass lengths
    16..BATCH_FACTOR_MAX.  The single-kernel Bluestein cases need ~180-250
    VGPRs, and a kernel's register count (its largest case's) sets the
    occupancy of every case, so the light flavour leaves them out."""
    ents = set()
    for length in batch_lengths():
        c = batch_tile_c(length)
        pow2 = length & (length - 1) == 0
        ents.add((length, c, VAR_PLAIN))
        if pow2 and length >= 16:
            ents.add((length, c, VAR_CHIRP_KMUL))
            ents.add((length, c, VAR_CONJ_CHIRP))
            if blue and length <= BATCH_BLUE_MAX:
                ents.add((length, c, VAR_BLUE))
        if 16 <= length <= BATCH_FACTOR_MAX:
            ents.add((length, c, VAR_TWIDDLE))
            if pow2:
                for v in (VAR_CHIRP, VAR_TWIDDLE_KMUL, VAR_TWIDDLE_CONJ_CHIRP):
                    ents.add((length, c, v))
    return sorted(ents)


def generate_batch_cases(out_dir: Path) -> list[Path]:
    paths = []
    for flavour in ("light", "blue"):
        lines = [
            f"// generated by reikna_amd/csrc/build.py: tile cases of the batch kernel ({flavour})",
            "// X(case id, L, C, variant)",
            "#define RK_BATCH_CASES(X) \\",
        ]
        for k, (length, c, v) in enumerate(batch_entries(blue=flavour == "blue")):
            lines.append(f"    X({k}, {length}, {c}, {v}) \\")
        lines.append("")
        path = out_dir / f"batch_cases_{flavour}.inc"
        write_if_changed(path, "\n".join(lines) + "\n")
        paths.append(path)
    return paths


def variant_cost(v: int) -> float:
    v &= ~(VAR_STREAM | VAR_NT_LOAD | VAR_NT_STORE | VAR_MODE_DD | VAR_MODE_DF | VAR_MODE_FF)
    return 2.0 if v & (1 << 4) else 1.3 if v else 1.0


def generate_catalog(out_dir: Path) -> list[Path]:
    ents = catalog_entries()
    # balance parts by a rough cost (longer transforms compile longer)
    parts: list[list[tuple[int, int, int]]] = [[] for _ in range(N_PARTS)]
    cost = [0.0] * N_PARTS
    for e in sorted(ents, key=lambda e: -(len(radix_list(e[0])) * variant_cost(e[2]))):
        k = min(range(N_PARTS), key=lambda i: cost[i])
        parts[k].append(e)
        cost[k] += len(radix_list(e[0])) * variant_cost(e[2])
    files = []
    for k, part in enumerate(parts):
        lines = [
            "// generated by reikna_amd/csrc/build.py: ahead-of-time FFT pass kernels",
            '#include "fft_device.hpp"',
            '#include "rk_internal.hpp"',
            f"namespace rk_catalog_{k} {{",
            "template <int L, int C, int V>",
            "__global__ void __launch_bounds__((rkfft::pass_threads<L, C>()), (rkfft::pass_min_waves<L, C, V>()))",
            "rk_fft_pass(rk_pass_args a) {",
            "    rkfft::fft_pass_entry<L, C, V>(a);",
            "}",
            "static const rk::KernelEntry entries[] = {",
        ]
        for length, c, v in sorted(part):
            threads = c * threads_per_instance(length)
            lds = tile_lds_bytes(length, c)
            lines.append(
                f"    {{{length}, {c}, {v}, {threads}, {lds}, "
                f"(const void*)&rk_fft_pass<{length}, {c}, {v}>}},"
            )
        if not part:
            lines.append("    {0, 0, 0, 0, 0, nullptr},")
        lines += [
            "};",
            "}",
            f'extern "C" int rk_catalog_part_{k}(const rk::KernelEntry** e) {{',
            f"    *e = rk_catalog_{k}::entries;",
            f"    return {len(part)};",
            "}",
        ]
        path = out_dir / f"catalog_part{k}.hip"
        write_if_changed(path, "\n".join(lines) + "\n")
        files.append(path)
    idx = [
        "// generated by reikna_amd/csrc/build.py: the catalog index",
        '#include "rk_internal.hpp"',
    ]
    idx += [f'extern "C" int rk_catalog_part_{k}(const rk::KernelEntry** e);' for k in range(N_PARTS)]
    idx += [
        "namespace rk {",
        "const std::vector<KernelEntry>& catalog() {",
        "    static const std::vector<KernelEntry> all = [] {",
        "        std::vector<KernelEntry> v;",
        "        const KernelEntry* e = nullptr;",
        "        int n = 0;",
    ]
    for k in range(N_PARTS):
        idx.append(f"        n = rk_catalog_part_{k}(&e); for (int i = 0; i < n; ++i) if (e[i].func) v.push_back(e[i]);")
    idx += ["        return v;", "    }();", "    return all;", "}", "}"]
    path = out_dir / "catalog_index.cpp"
    write_if_changed(path, "\n".join(idx) + "\n")
    files.append(path)
    return files


def generate_embedded(out_dir: Path) -> Path:
    """Device headers as C++ raw strings for hiprtc."""
    parts = ["// generated by reikna_amd/csrc/build.py: device sources for hiprtc"]
    for name in ("fft_args.h", "fft_device.hpp"):
        text = (HERE / name).read_text()
        ident = "k_src_" + name.replace(".", "_")
        # raw strings are split into chunks below the compilers' literal limits
        chunks = [text[i : i + 12000] for i in range(0, len(text), 12000)]
        body = "\n".join(f'R"RKSRC({c})RKSRC"' for c in chunks)
        parts.append(f"static const char* {ident} =\n{body};")
    path = out_dir / "rk_embedded.inc"
    write_if_changed(path, "\n".join(parts) + "\n")
    return path


def write_if_changed(path: Path, text: str) -> None:
    if path.exists() and path.read_text() == text:
        return
    path.write_text(text)


# ---------------------------------------------------------------------------
# compilation
# ---------------------------------------------------------------------------
def compile_one(src: Path, obj: Path, arch: str, incremental: bool, deps: list[Path]) -> tuple[Path, float, str]:
    t0 = time.time()
    stamp = obj.with_suffix(".hash")
    h = hashlib.sha256()
    for p in [src, *deps]:
        h.update(p.read_bytes())
    h.update(arch.encode())
    digest = h.hexdigest()
    if incremental and obj.exists() and stamp.exists() and stamp.read_text() == digest:
        return obj, 0.0, "cached"
    cmd = [
        HIPCC,
        "-c",
        "-x",
        "hip",
        f"--offload-arch={arch}",
        *COMMON_FLAGS,
        *DEVICE_FLAGS,
        f"-I{HERE}",
        f"-I{BUILD}",
        f"-I{REPO / 'include'}",
        str(src),
        "-o",
        str(obj),
    ]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    stamp.write_text(digest)
    return obj, time.time() - t0, "built"


def build(jobs: int = 8, incremental: bool = False, arch: str = "gfx950", verbose: bool = True) -> Path:
    BUILD.mkdir(exist_ok=True)
    gen = generate_catalog(BUILD)
    generate_embedded(BUILD)
    generate_batch_cases(BUILD)
    batch_kernels = [HERE / "rk_batch_blue.hip", HERE / "rk_batch_light.hip"]
    hand = [HERE / "rk_runtime.cpp", HERE / "rk_fft_plan.cpp", HERE / "rk_jit.cpp", HERE / "rk_copy.hip",
            HERE / "rk_batch.cpp", HERE / "rk_serve.cpp"]
    deps = [HERE / "fft_args.h", HERE / "fft_device.hpp", HERE / "fft_batch.h", HERE / "rk_internal.hpp",
            HERE / "rk_batch_internal.hpp", HERE / "rk_batch_kernel.hpp",
            REPO / "include" / "reikna_amd.h", BUILD / "rk_embedded.inc",
            BUILD / "batch_cases_light.inc", BUILD / "batch_cases_blue.inc"]
    # the batch kernels are the longest compiles: start them first
    sources = batch_kernels + hand + gen
    t0 = time.time()
    objs = []
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = [ex.submit(compile_one, s, BUILD / (s.stem + ".o"), arch, incremental, deps) for s in sources]
        for f in concurrent.futures.as_completed(futs):
            obj, dt, status = f.result()
            objs.append(obj)
            if verbose and status == "built":
                print(f"  {obj.name}: {dt:.1f}s", flush=True)
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={arch}", "-o", str(tmp), *map(str, sorted(objs)),
           "-L/opt/rocm/lib", "-lhiprtc", "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} in {time.time() - t0:.1f}s ({len(catalog_entries())} catalog kernels)", flush=True)
    return LIB


def main(argv: list[str] | None = None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--jobs", type=int, default=min(8, os.cpu_count() or 1))
    ap.add_argument("--incremental", action="store_true")
    ap.add_argument("--arch", default="gfx950")
    ap.add_argument("--clean", action="store_true")
    args = ap.parse_args(argv)
    if args.clean and BUILD.exists():
        shutil.rmtree(BUILD)
    build(jobs=args.jobs, incremental=args.incremental, arch=args.arch)
    return 0


if __name__ == "__main__":
    sys.exit(main())
