// rk_batch_blue.hip — the batch kernel with the Bluestein case list
// (with the single-kernel Bluestein cases; see rk_batch_kernel.hpp).
#include "batch_cases_blue.inc"
#define RK_BATCH_KERNEL rk_fft_batch_blue
#define RK_BATCH_HOST batch_blue
#include "rk_batch_kernel.hpp"
