"""CPU oracle for the FFT path — TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this package, and only as the checker (or the timed CPU baseline):
the product path (reikna_amd) never calls it and fails loudly without its
HIP library.

What it holds (each pinned in tests/test_oracle.py; see the function docstrings
for the reference lines they follow):
  * fftn / ifftn         — the reference's own oracle: numpy.fft in FP64
                           (tests/test_fft/test_fft.py:163,169)
  * dft_direct           — O(N^2) FP64 DFT from the definition
  * chirp, kweights      — reikna's Bluestein weights (fft.mako:196-202,
                           fft.py:216-236)
  * bluestein_fft        — reikna's chirp-z algorithm (fft.py:551-582)
  * radix_array, global_radix_info — reikna's radix planners (fft.py:24-146)
  * fft_flops            — the reference's FLOP convention (test_fft.py:234)
  * rel_l2, tolerance    — the error metric and the stated bound 2e-7 * log2 N
"""

from .fft_oracle import (
    bluestein_fft,
    chirp,
    dft_direct,
    fft_flops,
    fftn,
    global_radix_info,
    ifftn,
    kweights,
    radix_array,
    rel_l2,
    test_array,
    tolerance,
)

__all__ = [
    "bluestein_fft",
    "chirp",
    "dft_direct",
    "fft_flops",
    "fftn",
    "global_radix_info",
    "ifftn",
    "kweights",
    "radix_array",
    "rel_l2",
    "test_array",
    "tolerance",
]
