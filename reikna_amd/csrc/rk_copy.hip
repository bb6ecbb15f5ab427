// rk_copy.hip — strided N-d copy used by trivial FFT plans (every transformed
// axis has length 1), of elements of `words` 8-byte words.  Mirrors FFT._build_trivial_plan (reikna/fft/fft.py:647-662),
// which copies through PureParallel so that attached transformations still run.
#include "rk_internal.hpp"

namespace rk {

__global__ void __launch_bounds__(256) rk_copy_kernel(CopyArgs a) {
    const rk_u64 stride = (rk_u64)gridDim.x * blockDim.x;
    for (rk_u64 idx = (rk_u64)blockIdx.x * blockDim.x + threadIdx.x; idx < a.total; idx += stride) {
        rk_u64 rem = idx;
        rk_i64 io = a.in_off, oo = a.out_off;
        for (int d = a.ndim - 1; d >= 0; --d) {
            const rk_u64 s = (rk_u64)a.shape[d];
            const rk_u64 k = rem % s;
            rem /= s;
            io += (rk_i64)k * a.in_str[d];
            oo += (rk_i64)k * a.out_str[d];
        }
        for (int w = 0; w < a.words; ++w) a.out[oo * a.words + w] = a.in[io * a.words + w];
    }
}

int launch_copy(const CopyArgs& a, hipStream_t stream) {
    if (a.total == 0) return RK_OK;
    const rk_u64 blocks = (a.total + 255) / 256;
    const unsigned grid = (unsigned)(blocks < 65536 ? blocks : 65536);
    hipLaunchKernelGGL(rk_copy_kernel, dim3(grid), dim3(256), 0, stream, a);
    RK_HIP(hipGetLastError());
    return RK_OK;
}

}  // namespace rk
