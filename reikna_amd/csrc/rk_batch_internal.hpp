// rk_batch_internal.hpp — pass chains for the batch kernel, shared by batch
// plans (rk_batch.cpp) and the request server (rk_serve.cpp).
#ifndef RK_BATCH_INTERNAL_HPP
#define RK_BATCH_INTERNAL_HPP

#include "rk_internal.hpp"
#include "fft_batch.h"

namespace rk {

// `count` transforms of shape (n[0]) or (n[0], n[1]) (row-major, last axis
// contiguous), transform j at in_offset + j * in_dist (elements).
struct ItemShape {
    int nd = 1;
    int64_t n[2] = {1, 1};
    int64_t count = 1;
    int64_t in_offset = 0, out_offset = 0;
    int64_t in_dist = 0, out_dist = 0;
    bool scale = false;   // forward result * 1/prod(n), inverse unscaled
    int64_t elements() const { return nd == 1 ? n[0] : n[0] * n[1]; }
};

struct ChainPass {
    PassPlan pp;       // args with offsets relative to the pass's buffers
    AxisPass ap;
    bool first = false, last = false;
    float fwd_scale = 1.0f, inv_scale = 1.0f;
};

// The passes of one item, in order; workspace buffers BUF_WS0 / BUF_WS1 are
// ws_elems elements each (BUF_WS1 only when two_ws).
struct Chain {
    std::vector<ChainPass> passes;
    size_t ws_elems = 0;
    bool two_ws = false;
    bool blue = false;   // uses a single-kernel Bluestein case (BATCH_BLUE flavour)
    size_t ws_total() const { return ws_elems * (two_ws ? 2 : 1); }
};

// Chain limits of the batch kernel's tile cases (options may be null).
ChainLimits batch_limits(const rk_batch_options* options);

// Builds the chain of `item` (falling back to Bluestein for smooth lengths
// whose passes have no tile case) and appends its tables to tb.
int build_chain(const ItemShape& item, const ChainLimits& lim, TableBuilder& tb, Chain& chain);

// Appends the combine pass of a real-to-complex item of N real points to
// the chain of its N/2-point complex FFT (whose BUF_OUT is then the
// half-length spectrum, N/2 points per transform back to back).
int append_r2c_combine(int64_t n, int64_t count, int64_t out_offset, int64_t out_dist, bool scale, TableBuilder& tb,
                       Chain& chain);

// One pass placed in a queue: its arguments with buffer offsets applied.
struct QueuedPass {
    const ChainPass* cp;
    rk_pass_args a;
    rk_u32 src, dst;   // RK_BATCH_BUF_*
    rk_u32 inverse;
    int item;          // passes of one item are consecutive and in chain order
    int level;
};

// Orders the passes into batch-kernel descriptors (with the tile cases of
// `flavour`) and the tile map; returns the number of tiles.  Order modes:
//   0  critical path: pass i of an item of n passes goes to rank
//      i + (longest chain - n), so the longest chains start first and every
//      chain ends in the last rank; inside a rank, largest tiles first;
//   1  level by level, items in order;
//   2  level by level, largest tiles first.
uint64_t assemble_queue(const std::vector<QueuedPass>& in, int order_mode, int flavour,
                        std::vector<rk_batch_pass>& passes, std::vector<rk_u32>& tile_pass, int& levels,
                        size_t& counter_lines);

// Longest back-off between two polls of a waiting tile, in units of 128
// clocks (RK_BATCH_POLL_MAX overrides; default 16, ~0.85 us).
rk_u32 batch_poll_sleep_max();

// Resolves table offsets of a pass (from build_chain's TableBuilder) against
// the device copy of the table block.
void resolve_tables(rk_pass_args& a, const PassPlan& pp, const char* tables);

}  // namespace rk

#endif
