"""The split LDS address forms of the pass tile equal the plain padded positions.

The stage exchanges and staged global sides of `fft_pass_tile` address LDS
through `stage_in_lds`, `stage_out_lds`, `flat_lds` and `col_lds`
(reikna_amd/csrc/fft_device.hpp), which split `lds_pad(position)` into a
per-thread part and a per-value constant where no padding slot falls between
them.  This test builds a host program from the header and compares every
form with `lds_pad` of the plain position (`stage_in_pos`, `stage_out_pos`,
the instance-linear e = tid + it * NT and the column position ps + it * T),
for every thread, value and stage of every 7-smooth length up to 16384 and
every tile width of the kernel catalogue and the batch kernel (1-256
instances, up to 1024 threads) at both element sizes and both exchange
layouts.  The layout
itself (where reikna pads its local buffers, reikna/fft/fft.py:149-213) is
unchanged; only its arithmetic form is checked here.
"""

import os
import shutil
import subprocess
import tempfile
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
CSRC = ROOT / "reikna_amd" / "csrc"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

PROGRAM = r"""
#include "fft_device.hpp"
#include <cstdio>
using namespace rkfft;

static int g_bad = 0;
static void bad(const char* what, int L, int a, int b, int c, int got, int want) {
    if (g_bad++ < 20) std::printf("MISMATCH %s L=%d (%d,%d,%d): %d != %d\n", what, L, a, b, c, got, want);
}

template <int L, int S> void check_stage() {
    using P = Plan<L>;
    constexpr int R = P::radix(S);
    for (int t = 0; t < P::T; ++t)
        for (int i = 0; i < P::q(S) * R; ++i) {
            if (!stage_valid<L, S>(t, i / R)) continue;
            const int wo = lds_pad(stage_out_pos<L, S>(t, i)), go = stage_out_lds<L, S>(t, i);
            if (go != wo) bad("stage_out", L, S, t, i, go, wo);
            const int wi = lds_pad(stage_in_pos<L, S>(t, i)), gi = stage_in_lds<L, S>(t, i);
            if (gi != wi) bad("stage_in", L, S, t, i, gi, wi);
        }
}

template <int L, int S = 0> void check_stages() {
    if constexpr (S < Plan<L>::NSTAGE) {
        check_stage<L, S>();
        check_stages<L, S + 1>();
    }
}

template <int L, int C, bool HALF, int ES> void check_tile() {
    using P = Plan<L>;
    constexpr int T = P::T, NT = C * T, LP = tile_pitch<L, C, HALF, ES>();
    for (int tid = 0; tid < NT; ++tid)
        for (int it = 0; it < P::R0; ++it) {
            const int e = tid + it * NT, ce = e / L;
            const int want = ce * LP + lds_pad(e - ce * L), got = flat_lds<L, NT, LP>(tid, it);
            if (got != want) bad("flat", L, C, tid, it, got, want);
        }
    for (int ps = 0; ps < T; ++ps)
        for (int it = 0; it < P::R0; ++it) {
            const int want = lds_pad(ps + it * T), got = col_lds<T>(ps, it);
            if (got != want) bad("col", L, C, ps, it, got, want);
        }
}

template <int L, int C> void check_width() {
    if constexpr (C * Plan<L>::T <= 1024) {
        check_tile<L, C, false, 8>();
        check_tile<L, C, true, 8>();
        check_tile<L, C, false, 16>();
    }
}

template <int L> void check_length() {
    check_stages<L>();
    check_width<L, 1>(); check_width<L, 2>(); check_width<L, 3>(); check_width<L, 4>();
    check_width<L, 5>(); check_width<L, 6>(); check_width<L, 8>(); check_width<L, 10>();
    check_width<L, 12>(); check_width<L, 16>(); check_width<L, 21>(); check_width<L, 25>();
    check_width<L, 32>(); check_width<L, 42>(); check_width<L, 48>(); check_width<L, 51>();
    check_width<L, 64>(); check_width<L, 80>(); check_width<L, 85>(); check_width<L, 128>();
    check_width<L, 256>();
}

int main() {
@CALLS@
    std::printf("checked @N@ lengths, %d mismatches\n", g_bad);
    return g_bad != 0;
}
"""


def smooth_lengths(limit: int) -> list[int]:
    out = []
    for n in range(2, limit + 1):
        x = n
        for p in (2, 3, 5, 7):
            while x % p == 0:
                x //= p
        if x == 1:
            out.append(n)
    return out


@pytest.mark.skipif(not Path(HIPCC).exists(), reason="hipcc not available")
def test_split_lds_forms_equal_plain_positions():
    lengths = smooth_lengths(16384)
    calls = "\n".join(f"    check_length<{n}>();" for n in lengths)
    src = PROGRAM.replace("@CALLS@", calls).replace("@N@", str(len(lengths)))
    tmp = Path(tempfile.mkdtemp(prefix="rk_lds_forms_"))
    try:
        (tmp / "check.hip").write_text(src)
        exe = tmp / "check"
        build = subprocess.run(
            [HIPCC, "--offload-arch=gfx950", "--offload-host-only", "-O0", "-std=c++17",
             f"-I{CSRC}", f"-I{ROOT / 'include'}", str(tmp / "check.hip"), "-o", str(exe)],
            capture_output=True, text=True, timeout=600)
        assert build.returncode == 0, build.stderr[-4000:]
        run = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
        assert run.returncode == 0, run.stdout[-4000:]
        assert f"checked {len(lengths)} lengths, 0 mismatches" in run.stdout
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
