// rk_batch_blue.hip — the batch kernel with the Bluestein case list
// (with the single-kernel Bluestein cases; see rk_batch_kernel.hpp).  Its
// largest cases would take ~262 registers, one wave per SIMD; the bound keeps
// two waves per SIMD for all of its cases.
#include "batch_cases_blue.inc"
#define RK_BATCH_KERNEL rk_fft_batch_blue
#define RK_BATCH_HOST batch_blue
#define RK_BATCH_MIN_WAVES 2
#include "rk_batch_kernel.hpp"
