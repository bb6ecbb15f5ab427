// rk_internal.hpp — shared internals of the host library (not part of the C-ABI).
#ifndef RK_INTERNAL_HPP
#define RK_INTERNAL_HPP

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/reikna_amd.h"
#include "fft_args.h"

namespace rk {

// Sets the calling thread's last-error text and returns `code`.
int fail(int code, const char* fmt, ...);
void clear_error();

// Maps a failed HIP call to RK_ERR_HIP with the call site in the message.
#define RK_HIP(call)                                                                      \
    do {                                                                                  \
        hipError_t rk_e_ = (call);                                                        \
        if (rk_e_ != hipSuccess)                                                          \
            return ::rk::fail(RK_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(rk_e_)); \
    } while (0)

// One ahead-of-time compiled pass kernel (see the catalog_part*.hip files).
struct KernelEntry {
    int L;          // DFT length
    int C;          // instances per work-group tile
    int variant;    // pre-op | post-op << 2 | bluestein << 4 (fft_device.hpp Variant)
    int threads;    // work-group size
    int lds_bytes;  // static LDS of the work-group
    const void* func;
};

// Kernel variants (mirror of fft_device.hpp).
enum {
    VAR_PLAIN = 0,
    VAR_TWIDDLE = 1,
    VAR_CHIRP = 2,
    VAR_TWIDDLE_KMUL = 1 | (1 << 2),
    VAR_TWIDDLE_CONJ_CHIRP = 1 | (2 << 2),
    VAR_BLUE = 2 | (2 << 2) | (1 << 4),
};

// All ahead-of-time kernels compiled into this library.
const std::vector<KernelEntry>& catalog();

// Work-group tile widths the planner asks for (mirrors the catalog generator).
int tile_c_row(int L);
int tile_c_col(int L);
int lds_pitch_host(int L);
int threads_per_instance(int L);
bool is_smooth(int64_t n);

// A pass kernel from the run-time compiler: (L, C, variant) instantiated with
// the plain strided IO policy; cached per process and on disk.
int jit_pass_kernel(int L, int C, int variant, hipFunction_t* fn, int* threads, int* lds_bytes);

// Strided N-d copy (trivial plans: no axis longer than 1).
struct CopyArgs {
    const float2* in;
    float2* out;
    int ndim;
    rk_i64 shape[RK_MAX_DIMS];
    rk_i64 in_str[RK_MAX_DIMS];
    rk_i64 out_str[RK_MAX_DIMS];
    rk_i64 in_off, out_off;
    rk_u64 total;
};
int launch_copy(const CopyArgs& a, hipStream_t stream);

}  // namespace rk

#endif
