// rk_batch_kernel.hip — the persistent kernel that runs a whole batch of
// transforms of uneven lengths in one launch.
//
// No reference counterpart: reikna compiles and launches one computation per
// shape (reikna/fft/fft.py:723-756).  The queue protocol is described in
// fft_batch.h; the grid has one work-group per queue position and no atomic
// is shared by the whole grid (same-address device atomics serialise at
// ~11 ns each on MI355X: 94 us for one per work-group at 8192 work-groups),
// and the tiles of one pass count themselves on up to 16 counters.
// All data accesses of the kernel are device-coherent (CoherentIO), so a
// pass hands its output to the next pass of the same transform through a
// tile counter alone: agent-scope fences would write back or invalidate a
// whole XCD's L2 per tile.  (A
// grid-stride loop over the queue was measured to lose per-case register
// allocation: loop-invariant code of all cases is hoisted in front of the
// loop, and 33 cases already spill.)  Every tile case is one fft_pass_tile<L, C, V> instantiation
// (the same code as the one-pass-per-launch kernels) with C chosen so that a
// tile needs at most RK_BATCH_THREADS threads; work-groups always have
// RK_BATCH_THREADS threads and the surplus threads of a smaller tile repeat
// the work of its last thread.  The case list comes from build.py
// (batch_cases.inc): every length 2^k, 3*2^k and 5*2^k up to 4096 as a single
// pass and as single-kernel Bluestein, and the pass lengths up to 256 of
// multi-pass chains.
//
// This file is the kernel body; rk_batch_light.hip and rk_batch_blue.hip
// instantiate it with two case lists (without / with the single-kernel
// Bluestein cases, which need ~250 VGPRs and so lower the occupancy of every
// case), defining RK_BATCH_CASES, RK_BATCH_KERNEL (the kernel's name) and
// RK_BATCH_HOST (the namespace of its host-side case table).
#include "fft_device.hpp"
#include "fft_batch.h"

namespace {

constexpr int kThreads = RK_BATCH_THREADS;

#define RK_CASE_LDS(id, L, C, V) rkfft::tile_lds_bytes<L, C>(),
#define RK_CASE_THREADS(id, L, C, V) rkfft::pass_threads<L, C>(),
constexpr int kCaseLds[] = {RK_BATCH_CASES(RK_CASE_LDS)};
constexpr int kCaseThreads[] = {RK_BATCH_CASES(RK_CASE_THREADS)};
#undef RK_CASE_LDS
#undef RK_CASE_THREADS

constexpr int max_of(const int* v, int n) {
    int m = 0;
    for (int i = 0; i < n; ++i) m = v[i] > m ? v[i] : m;
    return m;
}
constexpr int kNumCases = sizeof(kCaseLds) / sizeof(kCaseLds[0]);
constexpr int kLdsBytes = max_of(kCaseLds, kNumCases);
static_assert(max_of(kCaseThreads, kNumCases) <= kThreads, "a batch tile case needs more threads than a work-group has");
static_assert(kLdsBytes <= 64 * 1024, "batch tile cases must leave room for two work-groups per CU");

// A case whose tile and W_L table fit kBatchLds reads its stage twiddles from
// an LDS copy of the table behind its tile, as the one-pass kernels' tiles do
// (fft_device.hpp twl_in_lds).  40 KB keeps the LDS-limited residency of the
// largest tile alone (four work-groups per CU) and takes the chain passes of
// up to 512 points.
constexpr int kBatchLds = kLdsBytes > 40 * 1024 ? kLdsBytes : 40 * 1024;
template <int L, int C> constexpr int twl_offset() { return (rkfft::tile_lds_bytes<L, C>() + 15) / 16 * 16; }
template <int L, int C> constexpr bool batch_twl() {
    return rkfft::Plan<L>::NSTAGE >= 2 && twl_offset<L, C>() + L * (int)sizeof(float2) <= kBatchLds;
}

__device__ __forceinline__ void run_case(rk_u32 kc, const rk_pass_args& a, float2* lds, rk_u64 tile, int tid) {
    switch (kc) {
#define RK_CASE_RUN(id, L, C, V)                                                       \
    case id: {                                                                         \
        constexpr int nt = rkfft::pass_threads<L, C>();                                \
        if constexpr (batch_twl<L, C>()) {                                             \
            /* published by the tile's first exchange barrier */                        \
            float2* tw = reinterpret_cast<float2*>(reinterpret_cast<char*>(lds) + twl_offset<L, C>()); \
            for (int e = tid; e < L; e += kThreads) tw[e] = rkfft::table<float2>(a.twl)[e]; \
            rkfft::fft_pass_tile<L, C, V, rkfft::CoherentIO, 0, false, false, false, true>( \
                a, rkfft::CoherentIO{}, lds, tile, tid < nt ? tid : nt - 1, nullptr, tw); \
        } else {                                                                       \
            rkfft::fft_pass_tile<L, C, V>(a, rkfft::CoherentIO{}, lds, tile, tid < nt ? tid : nt - 1); \
        }                                                                              \
    } break;
        RK_BATCH_CASES(RK_CASE_RUN)
#undef RK_CASE_RUN
        default:
            break;
    }
}

// Real-to-complex combine tile (RK_BATCH_CASE_R2C, fft_batch.h): Z is the
// N/2-point spectrum of the real input read as complex pairs x[n] = a[2n] +
// i a[2n+1]; X[k] = Z[k] A[k] + conj(Z[(N/2 - k) mod N/2]) B[k] for k < N/2
// and X[N/2] = Re Z[0] - Im Z[0] (Rabiner's split of a real FFT, the one
// examples/demo_specialized_fft.py:156-331 fuses; A, B from FP64).  Every
// load is issued before the first store.
__device__ __forceinline__ void r2c_combine_tile(const rk_pass_args& a, rk_u64 tile, int tid) {
    const rkfft::CoherentIO io;
    const rk_u64 chunks = a.dim_I;
    const rk_u64 inst = tile / chunks;
    const rk_u64 k0 = (tile - inst * chunks) * RK_BATCH_R2C_SPAN;
    const rk_u64 half = a.N;
    const rk_i64 zb = a.in_off + (rk_i64)inst * a.in_so;
    const rk_i64 ob = a.out_off + (rk_i64)inst * a.out_so;
    float2 zk[4], zr[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const rk_u64 k = k0 + (rk_u64)(e * RK_BATCH_THREADS + tid);
        // k = N/2 reads Z[0]; lanes past the end read their own Z[k mod N/2]
        // (a valid address) and discard it
        const rk_u64 kk = k < half ? k : k % half;
        const rk_u64 kr = kk ? half - kk : 0;
        zk[e] = io.load(a, inst, kk, 0, zb + (rk_i64)kk);
        zr[e] = io.load(a, inst, kr, 0, zb + (rk_i64)kr);
    }
#pragma unroll
    for (int e = 0; e < 4; ++e) {
        const rk_u64 k = k0 + (rk_u64)(e * RK_BATCH_THREADS + tid);
        if (k > half) continue;
        float2 x;
        if (k == half)
            x = make_float2(zk[e].x - zk[e].y, 0.0f);
        else
            x = rkfft::c_add(rkfft::c_mul(zk[e], a.chirp[k]), rkfft::c_mul(rkfft::c_conj(zr[e]), a.bspec[k]));
        io.store(a, inst, k, 0, ob + (rk_i64)k, rkfft::c_scale(x, a.out_scale));
    }
}

__device__ __forceinline__ void clear_counters(rk_u32* c, rk_u32 groups) {
    for (rk_u32 g = 0; g <= groups; ++g) __hip_atomic_store(c + g * RK_BATCH_LINE, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace

extern "C" __global__ void __launch_bounds__(kThreads, RK_BATCH_MIN_WAVES) RK_BATCH_KERNEL(rk_batch_work w) {
    __shared__ float2 lds[(kBatchLds + 7) / 8];
    const int tid = threadIdx.x;
    // queue position = work-group index: work-groups are dispatched in index
    // order, so every earlier position is running or finished
    const rk_u32 item = blockIdx.x;
    if (item >= w.total) return;
    const rk_u32 p = __builtin_amdgcn_readfirstlane(w.tile_pass[item]);
    const rk_batch_pass* bp = w.passes + p;
    const rk_i32 wait = bp->wait;
    if (wait >= 0) {
        if (tid == 0) {
            // poll the awaited pass's finished-group count, backing off from
            // ~50 ns to poll_sleep_max x ~50 ns between polls (a queue holds
            // many waiting work-groups; fewer polls leave the memory channel
            // of the counter to the atomics that move it)
            const rk_u32* top = w.counters + (size_t)bp->wait_line * RK_BATCH_LINE;
            const rk_u32 need = bp->wait_groups;
            rk_u32 polls = 0, nap = 1;
            while (__hip_atomic_load(top, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need) {
                if (++polls > w.spin_limit) {
                    __hip_atomic_store(&w.state[2], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
                for (rk_u32 k = 0; k < nap; ++k) __builtin_amdgcn_s_sleep(2);
                if (nap < w.poll_sleep_max) nap <<= 1;
            }
        }
        __syncthreads();
    }
    rk_pass_args a = bp->a;
    const rk_u32 src = bp->src, dst = bp->dst;
    a.in = src == RK_BATCH_BUF_IN ? w.in : src == RK_BATCH_BUF_OUT ? w.out : src == RK_BATCH_BUF_WS ? w.ws : a.in;
    a.out = dst == RK_BATCH_BUF_OUT ? w.out : dst == RK_BATCH_BUF_WS ? w.ws : a.out;
    const bool inv = (w.inverse | bp->inverse) != 0;
    a.in_sign = (inv && bp->item_first) ? -1.0f : 1.0f;
    a.out_sign = (inv && bp->item_last) ? -1.0f : 1.0f;
    a.out_scale = inv ? bp->inv_scale : bp->fwd_scale;
    if (bp->kcase == RK_BATCH_CASE_R2C)
        r2c_combine_tile(a, (rk_u64)(item - bp->first), tid);
    else
        run_case(bp->kcase, a, lds, (rk_u64)(item - bp->first), tid);
    const rk_u32 groups = bp->groups;
    if (groups) {
        // every store of the tile has completed (they are device-coherent)
        // before the tile is counted
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        if (tid == 0) {
            rk_u32* own = w.counters + (size_t)bp->line * RK_BATCH_LINE;
            const rk_u32 t = item - (rk_u32)bp->first;
            const rk_u32 n = (rk_u32)bp->a.n_tiles;
            const rk_u32 g = t % groups;
            const rk_u32 size = n / groups + (g < n % groups ? 1u : 0u);
            if (__hip_atomic_fetch_add(own + (1 + g) * RK_BATCH_LINE, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1 == size &&
                __hip_atomic_fetch_add(own, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) + 1 == groups) {
                // every tile of this pass is done: nothing reads the awaited
                // pass's counters any more, nor this pass's when no later pass
                // waits for it
                if (wait >= 0) clear_counters(w.counters + (size_t)bp->wait_line * RK_BATCH_LINE, bp->wait_groups);
                if (!bp->awaited) clear_counters(own, groups);
            }
        }
    }
}

// Host view of the kernel and its case list.
#include "rk_internal.hpp"

namespace rk {
namespace RK_BATCH_HOST {

const void* kernel() { return (const void*)&RK_BATCH_KERNEL; }

const std::vector<BatchCase>& cases() {
    static const std::vector<BatchCase> all = [] {
        std::vector<BatchCase> v;
#define RK_CASE_ENTRY(id, L, C, V) v.push_back(BatchCase{id, L, C, V, rkfft::pass_threads<L, C>(), rkfft::tile_lds_bytes<L, C>()});
        RK_BATCH_CASES(RK_CASE_ENTRY)
#undef RK_CASE_ENTRY
        return v;
    }();
    return all;
}

int lds_bytes() { return kBatchLds; }

}  // namespace RK_BATCH_HOST
}  // namespace rk
