// rk_internal.hpp — shared internals of the host library (not part of the C-ABI).
#ifndef RK_INTERNAL_HPP
#define RK_INTERNAL_HPP

#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <functional>
#include <map>
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
    VAR_R2C = 1 << 6,   // batch kernel only: the real-to-complex combine pass
    VAR_STREAM = 1 << 7,   // catalogue entries: the kernel uses StreamIO (RK_VAR_STREAM)
    VAR_MODE_DD = 1 << 8,  // catalogue entries: DIRECT/DIRECT load/store modes compiled in (RK_VAR_MODE_DD)
    VAR_MODE_DF = 1 << 9,  // catalogue entries: DIRECT/FLAT load/store modes compiled in (RK_VAR_MODE_DF)
    VAR_MODE_FF = 1 << 10, // catalogue entries: FLAT/FLAT load/store modes compiled in (RK_VAR_MODE_FF)
    VAR_NT_LOAD = 1 << 11, // catalogue entries: nontemporal loads only (RK_VAR_NT_LOAD)
    VAR_NT_STORE = 1 << 12, // catalogue entries: nontemporal stores only (RK_VAR_NT_STORE)
    VAR_WAVE = 1 << 13,    // catalogue entries: row mapping, wave-local stage exchanges (RK_VAR_WAVE)
    VAR_PIPE = 1 << 14,    // catalogue entries: persistent software-pipelined body (RK_VAR_PIPE)
    VAR_HALF = 1 << 15,    // catalogue entries: exchanges through re/im halves (RK_VAR_HALF)
    VAR_OCC = 1 << 16,     // catalogue entries: registers bounded for the LDS-allowed residency (RK_VAR_OCC)
    VAR_F64 = 1 << 17,     // complex128 elements (RK_VAR_F64; run-time compiled only)
};

// All ahead-of-time kernels compiled into this library.
const std::vector<KernelEntry>& catalog();

// Work-group tile widths the planner asks for (mirrors the catalog generator).
int tile_c_row(int L);
int tile_c_col(int L);
int lds_pitch_host(int L);            // base pitch (tile width choices)
int tile_lds_bytes_host(int L, int C, int es = 8);   // LDS of a tile (fft_device.hpp tile_lds_bytes)
int threads_per_instance(int L);
bool is_smooth(int64_t n);

// A pass kernel from the run-time compiler: (L, C, variant) instantiated with
// the plain strided IO policy; cached per process and on disk.
int jit_pass_kernel(int L, int C, int variant, hipFunction_t* fn, int* threads, int* lds_bytes);

// ---------------------------------------------------------------------------
// planner pieces shared by single-shape plans and batch plans
// ---------------------------------------------------------------------------
struct Layout {   // an N-d strided view (element strides / offset)
    int ndim;
    int64_t shape[RK_MAX_DIMS];
    int64_t strides[RK_MAX_DIMS];
    int64_t offset;
};
Layout contiguous_layout(int ndim, const int64_t* shape);

struct HostLoop {   // batch dims that could not be folded into the kernel's O / I
    int64_t n, in_str, out_str;
};

enum BufId { BUF_IN = 0, BUF_OUT = 1, BUF_WS0 = 2, BUF_WS1 = 3 };

// One pass of an axis: a radix-L Stockham step over an axis of length Nax.
struct AxisPass {
    int L;
    int64_t Nax;      // axis length the Stockham step works on (N or M)
    int64_t NS, H;
    bool blue_single;
    uint32_t pre_ops, post_ops;
    int64_t n_valid_in, n_valid_out;
    int64_t in_len, out_len;   // axis length in the src / dst layouts
};

// Limits of the pass chain of one axis.
struct ChainLimits {
    int64_t single_max;    // longest single-pass DFT
    int64_t factor_max;    // longest pass of a multi-pass chain
    int64_t blue_single_max;   // longest single-kernel Bluestein length M
    bool pow2_bluestein;   // Bluestein lengths are powers of two
    bool (*length_ok)(int64_t L);   // pass lengths a kernel exists for (null: any smooth length)
    int64_t strided_single_max;     // single_max of axes that are not contiguous in the input (0: single_max)
    int64_t first_factor = 0;       // length of the first pass of a multi-pass split (0: the balanced split)
};
ChainLimits default_limits();

// The passes of one axis of length n.
int axis_passes(int64_t n, uint32_t flags, const ChainLimits& lim, std::vector<AxisPass>& out);

struct PassPlan {
    rk_pass_args args;
    int src, dst;
    int L, C, variant, threads, lds_bytes;
    const void* aot = nullptr;      // ahead-of-time kernel
    hipFunction_t jit = nullptr;    // run-time compiled kernel
    bool user_kernel = false;       // replaced with rk_fft_plan_set_pass_kernel
    bool is_first = false, is_last = false;
    bool nt_load = false;           // nontemporal loads (the kernel's IO policy)
    bool nt_store = false;          // nontemporal stores
    bool wave = false;              // row mapping with wave-local stage exchanges (RK_VAR_WAVE)
    bool pipe = false;              // persistent software-pipelined kernel: one launch per tile range
    bool half = false;              // exchanges through re/im halves (half the LDS)
    bool occ = false;               // registers bounded for the LDS-allowed residency
    int es = 8;                     // element bytes: 8 complex64, 16 complex128
    bool fixed_modes = false;       // the kernel has this pass's load/store modes compiled in
    uint64_t grid = 0;
    int axis = 0;
    int kind = 0;
    int64_t twl_off = -1, twhi_off = -1, twlo_off = -1, chirp_off = -1, bspec_off = -1;
    std::vector<HostLoop> loops;
    uint64_t inst_per_outer = 0;   // H * NS * I: instances per outer-batch index
    int64_t outer_count = 1;       // O
    int64_t src_stride0 = 0, dst_stride0 = 0;   // element stride of dim 0 in the src / dst layouts
};

// Device tables of a plan, appended once per key and 256-byte aligned.
struct TableBuilder {
    std::vector<unsigned char> bytes;
    std::map<std::string, int64_t> index;
    int64_t add(const std::string& key, const void* data, size_t n);
};

// Tile width C for a pass of length L under the row (colmap = false) or
// column compute mapping.
typedef int (*TileWidthFn)(int L, bool colmap);

// Fills pp.args (dims, strides, walk modes, ops), pp.variant, pp.C and pp.loops
// for pass `ap` from layout ls to ld along `axis`.
int describe_pass(const AxisPass& ap, const Layout& ls, const Layout& ld, int axis, TileWidthFn tile_width,
                  PassPlan& pp);

// Called for every pass route_passes builds, after describe_pass: picks the
// kernel and builds the tables (n: the axis length before Bluestein padding).
typedef std::function<int(PassPlan& pp, const AxisPass& ap, int64_t n)> PassHook;

// The passes of a transform over `work` (axes innermost first) from layout
// lin to lout: the first pass reads BUF_IN, each axis's last pass writes
// BUF_OUT (or only the very last one with RK_FFT_OUT_LAST_ONLY), passes in
// between ping-pong through BUF_WS0 / BUF_WS1, contiguous buffers of
// ws_elems elements each (two_ws: both are needed).
int route_passes(const rk_fft_desc& d, const std::vector<int>& work, const Layout& lin, const Layout& lout,
                 const ChainLimits& lim, TileWidthFn tile_width, const PassHook& hook, std::vector<PassPlan>& passes,
                 size_t& ws_elems, bool& two_ws);

// Stage twiddles, inter-pass twiddles, chirp and Bluestein spectrum of a pass
// (nreal: the transform length before Bluestein padding).
int build_tables(TableBuilder& tb, PassPlan& pp, const AxisPass& ap, int64_t nreal);
// Host-side multipliers of the kernels' instance decomposition (rk_pass_args
// mag_* / sh_*), set whenever a pass's dims change.
void set_div_magic(rk_pass_args& a);

// Tile cases of the batch kernel (rk_batch_kernel.hpp).
struct BatchCase {
    int id, L, C, variant, threads, lds_bytes;
};
// Two flavours of the kernel: light (no single-kernel Bluestein case) and
// blue (a superset with them); a batch uses blue only if it needs one.
enum BatchFlavour { BATCH_LIGHT = 0, BATCH_BLUE = 1 };
namespace batch_light {
const void* kernel();
const std::vector<BatchCase>& cases();
int lds_bytes();
}  // namespace batch_light
namespace batch_blue {
const void* kernel();
const std::vector<BatchCase>& cases();
int lds_bytes();
}  // namespace batch_blue
inline const void* batch_kernel(int flavour) { return flavour ? batch_blue::kernel() : batch_light::kernel(); }
inline const std::vector<BatchCase>& batch_cases(int flavour) {
    return flavour ? batch_blue::cases() : batch_light::cases();
}
inline int batch_lds_bytes(int flavour) { return flavour ? batch_blue::lds_bytes() : batch_light::lds_bytes(); }

// Strided N-d copy (trivial plans: no axis longer than 1).
struct CopyArgs {
    const float2* in;
    float2* out;
    int words = 1;   // 8-byte words per element (1 complex64, 2 complex128)
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
