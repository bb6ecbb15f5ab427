// rk_batch_light.hip — the batch kernel with the light case list
// (no single-kernel Bluestein cases; see rk_batch_kernel.hpp).
#include "batch_cases_light.inc"
#define RK_BATCH_KERNEL rk_fft_batch_light
#define RK_BATCH_HOST batch_light
#define RK_BATCH_MIN_WAVES 1
#include "rk_batch_kernel.hpp"
