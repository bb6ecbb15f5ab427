// rk_batch.cpp — batched transforms of uneven lengths behind
// rk_batch_plan_create / rk_batch_execute, and the pass chains of the batch
// kernel (rk_batch_internal.hpp) that the request server reuses.
//
// No reference counterpart: reikna compiles and launches one computation per
// shape (reikna/fft/fft.py:723-756).  Two execution modes:
//
//   queue (default): every item becomes a chain of passes (the same Stockham
//     and Bluestein chains as single-shape plans, restricted to the tile cases
//     of the batch kernel), every pass becomes tiles, and one launch of
//     rk_fft_batch runs all tiles of all items; see fft_batch.h for the queue
//     and its dependency counters.
//   grouped: one single-shape plan per item, executed one after the other
//     (one launch per pass per item: the per-request path).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <numeric>

#include "rk_batch_internal.hpp"

using rk::fail;

namespace {

constexpr int64_t kBatchSingleMax = 4096;
constexpr int64_t kBatchFactorMax = 1024;      // longest chain pass with tile cases
// chain pass limit when the options leave it at 0: the 512-transform mix of
// lengths 64..65536 took 99.7 us with 512, 100.3 with 1024, 109.9 with 256
// (its critical path, Bluestein chains over 2^17 points, is two passes per
// chain with 512 and three with 256)
constexpr int64_t kBatchFactorDefault = 512;
constexpr int64_t kBatchStridedMax = 256;   // column axes: >= 16 adjacent instances per tile
constexpr uint32_t kSpinLimit = 1u << 24;   // polls: ~1 s at the shortest back-off

const rk::BatchCase* find_case(int L, int variant, int flavour = rk::BATCH_BLUE) {
    for (const auto& c : rk::batch_cases(flavour))
        if (c.L == L && c.variant == variant) return &c;
    return nullptr;
}

bool batch_length_ok(int64_t L) { return L <= kBatchSingleMax && find_case((int)L, rk::VAR_PLAIN) != nullptr; }

int batch_tile_width(int L, bool) {
    const rk::BatchCase* c = find_case(L, rk::VAR_PLAIN);
    return c ? c->C : 1;
}

}  // namespace

namespace rk {

ChainLimits batch_limits(const rk_batch_options* o) {
    ChainLimits lim;
    const int32_t sm = o ? o->single_max : 0, fm = o ? o->factor_max : 0;
    lim.single_max = sm > 0 ? std::min<int64_t>(sm, kBatchSingleMax) : kBatchSingleMax;
    lim.factor_max = fm > 0 ? std::min<int64_t>(fm, kBatchFactorMax) : kBatchFactorDefault;
    // single-kernel Bluestein up to the longest power of two with a case
    static const int64_t blue_max = [] {
        int64_t m = 0;
        for (const auto& c : batch_cases(BATCH_BLUE))
            if (c.variant == VAR_BLUE && (c.L & (c.L - 1)) == 0) m = std::max<int64_t>(m, c.L);
        return m;
    }();
    lim.blue_single_max = blue_max;
    lim.pow2_bluestein = true;
    lim.length_ok = batch_length_ok;
    lim.strided_single_max = kBatchStridedMax;
    return lim;
}

int build_chain(const ItemShape& it, const ChainLimits& lim, TableBuilder& tb, Chain& chain) {
    rk_fft_desc d;
    memset(&d, 0, sizeof(d));
    d.dtype = RK_C64;
    d.ndim = 1 + it.nd;
    d.shape[0] = it.count;
    d.in_strides[0] = it.in_dist;
    d.out_strides[0] = it.out_dist;
    int64_t inner = 1;
    for (int k = it.nd; k >= 1; --k) {
        d.shape[k] = it.n[k - 1];
        d.in_strides[k] = inner;
        d.out_strides[k] = inner;
        inner *= it.n[k - 1];
    }
    d.in_offset = it.in_offset;
    d.out_offset = it.out_offset;
    Layout lin, lout;
    lin.ndim = lout.ndim = d.ndim;
    for (int k = 0; k < d.ndim; ++k) {
        lin.shape[k] = lout.shape[k] = d.shape[k];
        lin.strides[k] = d.in_strides[k];
        lout.strides[k] = d.out_strides[k];
    }
    lin.offset = d.in_offset;
    lout.offset = d.out_offset;
    std::vector<int> work;   // innermost first
    for (int k = d.ndim - 1; k >= 1; --k)
        if (d.shape[k] > 1) work.push_back(k);
    if (work.empty()) return fail(RK_ERR_INVALID, "batch: a transform of one element is a copy");
    // a smooth length whose passes have no tile case goes through Bluestein
    int rc = RK_OK;
    for (uint32_t force : {0u, (uint32_t)RK_FFT_FORCE_BLUESTEIN}) {
        d.flags = force;
        std::vector<PassPlan> passes;
        std::vector<AxisPass> aps;
        size_t ws = 0;
        bool two = false;
        PassHook hook = [&](PassPlan& pp, const AxisPass& ap, int64_t) -> int {
            const BatchCase* bc = find_case(ap.L, pp.variant);
            if (!bc || bc->C != pp.C || !pp.loops.empty())
                return fail(RK_ERR_NO_KERNEL, "batch: no tile case for L=%d variant %d", ap.L, pp.variant);
            aps.push_back(ap);
            return RK_OK;
        };
        rc = route_passes(d, work, lin, lout, lim, batch_tile_width, hook, passes, ws, two);
        if (rc != RK_OK) continue;
        chain.passes.clear();
        chain.ws_elems = ws;
        chain.two_ws = two;
        chain.blue = false;
        for (const PassPlan& pp : passes)
            if (pp.src == BUF_WS1 || pp.dst == BUF_WS1) chain.two_ws = true;
        const double norm = 1.0 / (double)it.elements();
        for (size_t i = 0; i < passes.size(); ++i) {
            ChainPass cp;
            cp.pp = passes[i];
            cp.ap = aps[i];
            cp.first = i == 0;
            if (cp.pp.variant == VAR_BLUE) chain.blue = true;
            cp.last = i + 1 == passes.size();
            if (cp.last) {
                cp.fwd_scale = it.scale ? (float)norm : 1.0f;
                cp.inv_scale = it.scale ? 1.0f : (float)norm;
            }
            // tables from the axis length before Bluestein padding
            rc = build_tables(tb, cp.pp, cp.ap, d.shape[cp.pp.axis]);
            if (rc != RK_OK) return rc;
            chain.passes.push_back(cp);
        }
        return RK_OK;
    }
    return rc;
}

rk_u32 batch_poll_sleep_max() {
    static const rk_u32 v = [] {
        const char* e = getenv("RK_BATCH_POLL_MAX");
        const long x = e && *e ? atol(e) : 16;
        return (rk_u32)std::min<long>(std::max<long>(x, 1), 1024);
    }();
    return v;
}

// Tile groups of a counted pass of n tiles: about sqrt(n), so that the
// group counters and the finished-group counter see similar numbers of
// atomics.
static rk_u32 tile_groups(uint64_t n) {
    rk_u32 g = 1;
    while (g < RK_BATCH_GROUPS_MAX && (uint64_t)(g + 1) * (g + 1) <= n) ++g;
    return g;
}

// Appends the real-to-complex combine pass to the chain of the half-length
// complex FFT of an RK_BATCH_R2C item (N real points, `count` transforms,
// output at out_offset + j * out_dist) and its A / B tables (FP64, rounded
// once).  The chain's own output (BUF_OUT) is the half-length spectrum Z,
// `count` transforms of N/2 points back to back; the combine pass reads it
// from offset 0 of that region.
int append_r2c_combine(int64_t n, int64_t count, int64_t out_offset, int64_t out_dist, bool scale, TableBuilder& tb,
                       Chain& chain) {
    const int64_t half = n / 2;
    std::vector<float2> ta((size_t)half), tbv((size_t)half);
    for (int64_t k = 0; k < half; ++k) {
        // W = exp(-2 pi i k / N), k < N/2: A = (1 - iW)/2, B = (1 + iW)/2
        const double ang = -2.0 * M_PI * (double)k / (double)n;
        const double c = std::cos(ang), sn = std::sin(ang);
        ta[(size_t)k] = make_float2((float)(0.5 * (1.0 + sn)), (float)(-0.5 * c));
        tbv[(size_t)k] = make_float2((float)(0.5 * (1.0 - sn)), (float)(0.5 * c));
    }
    ChainPass cp;
    memset(&cp.pp.args, 0, sizeof(cp.pp.args));
    rk_pass_args& a = cp.pp.args;
    const rk_u64 chunks = (rk_u64)((half + 1 + RK_BATCH_R2C_SPAN - 1) / RK_BATCH_R2C_SPAN);
    a.N = (rk_u64)half;
    a.n_inst = (rk_u64)count;
    a.dim_I = chunks;
    a.n_tiles = (rk_u64)count * chunks;
    a.in_off = 0;
    a.in_so = half;
    a.in_sa = 1;
    a.out_off = out_offset;
    a.out_so = out_dist;
    a.out_sa = 1;
    a.in_sign = a.out_sign = a.out_scale = 1.0f;
    a.n_valid_in = a.n_valid_out = (rk_u64)half;
    cp.pp.src = BUF_OUT;   // the chain's output region (Z)
    cp.pp.dst = BUF_OUT;
    cp.pp.L = 1;
    cp.pp.C = RK_BATCH_R2C_SPAN;
    cp.pp.variant = VAR_R2C;
    cp.pp.threads = RK_BATCH_THREADS;
    const std::string key = "r2c:" + std::to_string(n);
    cp.pp.chirp_off = tb.add(key + ":A", ta.data(), ta.size() * sizeof(float2));
    cp.pp.bspec_off = tb.add(key + ":B", tbv.data(), tbv.size() * sizeof(float2));
    cp.first = false;
    cp.last = true;
    cp.fwd_scale = scale ? (float)(1.0 / (double)n) : 1.0f;
    cp.inv_scale = 1.0f;
    if (chain.passes.empty()) return fail(RK_ERR_INVALID, "batch: r2c item without a complex chain");
    ChainPass& prev = chain.passes.back();
    prev.last = false;
    prev.fwd_scale = prev.inv_scale = 1.0f;
    chain.passes.push_back(cp);
    return RK_OK;
}

uint64_t assemble_queue(const std::vector<QueuedPass>& in, int order_mode, int flavour,
                        std::vector<rk_batch_pass>& out, std::vector<rk_u32>& tile_pass, int& levels,
                        size_t& lines) {
    // chain length of every pass's item (the passes of an item are consecutive)
    std::vector<int> rank(in.size());
    int longest = 0;
    for (size_t i = 0; i < in.size();) {
        size_t j = i;
        while (j < in.size() && in[j].item == in[i].item) ++j;
        longest = std::max<int>(longest, (int)(j - i));
        i = j;
    }
    for (size_t i = 0; i < in.size();) {
        size_t j = i;
        while (j < in.size() && in[j].item == in[i].item) ++j;
        const int shift = order_mode == 0 ? longest - (int)(j - i) : 0;
        for (size_t k = i; k < j; ++k) rank[k] = in[k].level + shift;
        i = j;
    }
    std::vector<size_t> order(in.size());
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
        const QueuedPass &x = in[a], &y = in[b];
        if (rank[a] != rank[b]) return rank[a] < rank[b];
        if (order_mode == 1) return x.item < y.item;
        const int64_t tx = (int64_t)x.cp->pp.C * x.cp->pp.L, ty = (int64_t)y.cp->pp.C * y.cp->pp.L;
        if (tx != ty) return tx > ty;
        return x.item < y.item;
    });
    std::vector<int> pos(in.size());
    for (size_t q = 0; q < order.size(); ++q) pos[order[q]] = (int)q;
    uint64_t total = 0;
    levels = 0;
    lines = 0;
    out.resize(in.size());
    for (size_t q = 0; q < order.size(); ++q) {
        const size_t i = order[q];
        const QueuedPass& qp = in[i];
        const ChainPass& cp = *qp.cp;
        rk_batch_pass& bp = out[q];
        memset(&bp, 0, sizeof(bp));
        bp.a = qp.a;
        bp.first = total;
        if (cp.pp.variant == VAR_R2C) {
            bp.kcase = RK_BATCH_CASE_R2C;
        } else {
            const BatchCase* bc = find_case(cp.pp.L, cp.pp.variant, flavour);
            bp.kcase = bc ? (rk_u32)bc->id : 0u;   // build_chain checked the case exists in the blue flavour
        }
        bp.src = qp.src;
        bp.dst = qp.dst;
        // the item's previous pass (earlier in the queue: it has a lower level)
        // is the only one it waits for, and it is the only pass that waits for
        // that one
        bp.wait = cp.first ? -1 : pos[i - 1];
        bp.awaited = cp.last ? 0u : 1u;
        if (!(cp.first && cp.last)) {
            bp.groups = tile_groups(qp.a.n_tiles);
            bp.line = (rk_u32)lines;
            lines += 1 + bp.groups;
        }
        bp.item_first = cp.first ? 1u : 0u;
        bp.item_last = cp.last ? 1u : 0u;
        bp.fwd_scale = cp.fwd_scale;
        bp.inv_scale = cp.inv_scale;
        bp.inverse = qp.inverse;
        total += qp.a.n_tiles;
        levels = std::max(levels, qp.level + 1);
    }
    for (rk_batch_pass& bp : out)
        if (bp.wait >= 0) {
            bp.wait_line = out[bp.wait].line;
            bp.wait_groups = out[bp.wait].groups;
        }
    tile_pass.resize(total);
    for (size_t q = 0; q < out.size(); ++q)
        for (rk_u64 t = 0; t < out[q].a.n_tiles; ++t) tile_pass[out[q].first + t] = (rk_u32)q;
    return total;
}

void resolve_tables(rk_pass_args& a, const PassPlan& pp, const char* tables) {
    auto tab = [&](int64_t off) -> const void* { return off < 0 || !tables ? nullptr : tables + off; };
    a.twl = (const float2*)tab(pp.twl_off);
    a.tw_hi = (const double2*)tab(pp.twhi_off);
    a.tw_lo = (const double2*)tab(pp.twlo_off);
    a.chirp = (const float2*)tab(pp.chirp_off);
    a.bspec = (const float2*)tab(pp.bspec_off);
    a.tile_base = 0;
    a.xcd_group = 0;
}

}  // namespace rk

struct rk_batch_plan {
    rk_batch_options opt;
    // grouped mode
    std::vector<rk_fft_plan*> plans;
    // queue mode
    std::vector<rk::Chain> chains;
    std::vector<rk_batch_pass> passes;   // queue order
    std::vector<rk_u32> tile_pass;
    std::vector<rk::CopyArgs> copies;    // length-1 items
    void* dev = nullptr;                 // tables | passes | tile_pass | counters | state
    rk_batch_work work;
    int flavour = rk::BATCH_LIGHT;
    int64_t n_items = 0;
    int levels = 0;
    size_t ws_bytes = 0;
    int launches = 0;
    bool dry = false;
    bool r2c = false;                    // has real-to-complex items (forward only)
    size_t counter_bytes = 0;            // dependency counters (work.counters)
};

extern "C" int rk_batch_plan_destroy(rk_batch_plan* plan) {
    if (!plan) return RK_OK;
    for (auto* p : plan->plans) rk_fft_plan_destroy(p);
    if (plan->dev) hipFree(plan->dev);
    delete plan;
    return RK_OK;
}

static int create_grouped(const rk_batch_item* items, int n_items, rk_batch_plan* plan) {
    for (int k = 0; k < n_items; ++k) {
        const rk_batch_item& it = items[k];
        if (it.count == 0) continue;
        if (it.flags & RK_BATCH_R2C)
            return fail(RK_ERR_UNSUPPORTED, "batch item %d: real-to-complex items run in queue mode only", k);
        rk_fft_desc d;
        memset(&d, 0, sizeof(d));
        d.dtype = RK_C64;
        d.ndim = 2;
        d.shape[0] = it.count;
        d.shape[1] = it.length;
        d.in_strides[0] = it.in_dist;
        d.in_strides[1] = 1;
        d.out_strides[0] = it.out_dist;
        d.out_strides[1] = 1;
        d.in_offset = it.in_offset;
        d.out_offset = it.out_offset;
        d.naxes = 1;
        d.axes[0] = 1;
        d.flags = ((it.flags & RK_BATCH_SCALE) ? RK_FFT_NORM_FORWARD : 0u) |
                  ((plan->opt.flags & RK_BATCH_DRY_RUN) ? RK_FFT_DRY_RUN : 0u);
        rk_fft_plan* p = nullptr;
        int rc = rk_fft_plan_create(&d, &p);
        if (rc != RK_OK) return rc;
        plan->plans.push_back(p);
        size_t ws = 0;
        rk_fft_plan_workspace_size(p, &ws);
        plan->ws_bytes = std::max(plan->ws_bytes, ws);
        int np = 0;
        rk_fft_plan_num_passes(p, &np);
        plan->launches += std::max(np, 1);
        plan->levels = std::max(plan->levels, np);
    }
    return RK_OK;
}

static int create_queue(const rk_batch_item* items, int n_items, rk_batch_plan* plan) {
    const rk::ChainLimits lim = rk::batch_limits(&plan->opt);
    rk::TableBuilder tb;
    std::vector<rk::QueuedPass> queued;
    size_t ws_elems = 0;
    plan->chains.reserve(n_items);   // QueuedPass keeps pointers into the chains
    for (int k = 0; k < n_items; ++k) {
        const rk_batch_item& it = items[k];
        if (it.count == 0) continue;
        plan->n_items += it.count;
        if (it.length == 1) {   // identity: a strided copy launched beside the queue
            rk::CopyArgs c;
            memset(&c, 0, sizeof(c));
            c.words = 1;
            c.ndim = 1;
            c.shape[0] = it.count;
            c.in_str[0] = it.in_dist;
            c.out_str[0] = it.out_dist;
            c.in_off = it.in_offset;
            c.out_off = it.out_offset;
            c.total = (rk_u64)it.count;
            plan->copies.push_back(c);
            continue;
        }
        // a real-to-complex item is the complex FFT of its N/2 pairs into a
        // workspace region Z, then the combine pass from Z to `out`
        const bool r2c = (it.flags & RK_BATCH_R2C) != 0;
        const int64_t n = r2c ? it.length / 2 : it.length;
        rk::ItemShape sh;
        sh.nd = 1;
        sh.n[0] = n;
        sh.count = it.count;
        sh.in_offset = it.in_offset;
        sh.in_dist = it.count > 1 ? it.in_dist : n;
        sh.out_offset = r2c ? 0 : it.out_offset;
        sh.out_dist = r2c ? n : it.count > 1 ? it.out_dist : n;
        sh.scale = !r2c && (it.flags & RK_BATCH_SCALE) != 0;
        plan->chains.emplace_back();
        rk::Chain& ch = plan->chains.back();
        int rc = rk::build_chain(sh, lim, tb, ch);
        if (rc != RK_OK) return rc;
        if (r2c) {
            rc = rk::append_r2c_combine(it.length, it.count, it.out_offset, it.count > 1 ? it.out_dist : n + 1,
                                        (it.flags & RK_BATCH_SCALE) != 0, tb, ch);
            if (rc != RK_OK) return rc;
            plan->r2c = true;
        }
        // this item's workspace regions: [ws_elems, + ws_total), then Z
        const int64_t base0 = (int64_t)ws_elems, base1 = (int64_t)(ws_elems + ch.ws_elems);
        const int64_t base_z = (int64_t)(ws_elems + ((ch.ws_total() + 31) & ~size_t(31)));
        for (size_t i = 0; i < ch.passes.size(); ++i) {
            const rk::ChainPass& cp = ch.passes[i];
            rk::QueuedPass qp;
            qp.cp = &cp;
            qp.a = cp.pp.args;
            qp.item = k;
            qp.level = (int)i;
            qp.inverse = 0;
            auto map = [&](int buf, rk_i64& off) -> rk_u32 {
                if (buf == rk::BUF_IN) return RK_BATCH_BUF_IN;
                if (buf == rk::BUF_OUT && !r2c) return RK_BATCH_BUF_OUT;
                off += buf == rk::BUF_WS0 ? base0 : buf == rk::BUF_WS1 ? base1 : base_z;
                return RK_BATCH_BUF_WS;
            };
            if (cp.pp.variant == rk::VAR_R2C) {
                qp.a.in_off += base_z;
                qp.src = RK_BATCH_BUF_WS;
                qp.dst = RK_BATCH_BUF_OUT;
            } else {
                qp.src = map(cp.pp.src, qp.a.in_off);
                qp.dst = map(cp.pp.dst, qp.a.out_off);
            }
            queued.push_back(qp);
        }
        ws_elems += (ch.ws_total() + 31) & ~size_t(31);   // 256-byte aligned regions
        if (r2c) ws_elems += ((size_t)(n * it.count) + 31) & ~size_t(31);
    }

    uint64_t total = 0;
    for (const auto& qp : queued) total += qp.a.n_tiles;
    plan->flavour = rk::BATCH_LIGHT;
    for (const auto& ch : plan->chains)
        if (ch.blue) plan->flavour = rk::BATCH_BLUE;
    if (total >= (uint64_t(1) << 31))
        return fail(RK_ERR_UNSUPPORTED, "batch has too many tiles (%llu)", (unsigned long long)total);
    plan->ws_bytes = ws_elems * sizeof(float2);
    plan->launches = (total ? 1 : 0) + (int)plan->copies.size();
    memset(&plan->work, 0, sizeof(plan->work));
    plan->work.n_passes = (rk_u32)queued.size();
    plan->work.total = (rk_u32)total;
    plan->work.spin_limit = kSpinLimit;
    if (const char* e = getenv("RK_BATCH_SPIN_LIMIT"))   // tests of the failure path
        if (*e && atol(e) > 0) plan->work.spin_limit = (uint32_t)atol(e);
    plan->work.poll_sleep_max = rk::batch_poll_sleep_max();
    size_t lines = 0;
    if (plan->dry) {
        rk::assemble_queue(queued, plan->opt.order, plan->flavour, plan->passes, plan->tile_pass, plan->levels, lines);
        return RK_OK;
    }

    // one device allocation: tables | passes | tile_pass | counters | state
    auto align = [](size_t x) { return (x + 255) & ~size_t(255); };
    const size_t off_pass = align(tb.bytes.size());
    const size_t off_tp = off_pass + align(queued.size() * sizeof(rk_batch_pass));
    const size_t off_cnt = off_tp + align((size_t)total * sizeof(rk_u32));
    for (auto& qp : queued) rk::resolve_tables(qp.a, qp.cp->pp, nullptr);
    rk::assemble_queue(queued, plan->opt.order, plan->flavour, plan->passes, plan->tile_pass, plan->levels, lines);
    const size_t off_state = off_cnt + align(lines * RK_BATCH_LINE * sizeof(rk_u32));
    plan->counter_bytes = off_state - off_cnt;
    const size_t bytes = off_state + 256;
    RK_HIP(hipMalloc(&plan->dev, bytes));
    char* base = (char*)plan->dev;
    for (auto& qp : queued) rk::resolve_tables(qp.a, qp.cp->pp, base);
    rk::assemble_queue(queued, plan->opt.order, plan->flavour, plan->passes, plan->tile_pass, plan->levels, lines);
    if (!tb.bytes.empty()) RK_HIP(hipMemcpy(base, tb.bytes.data(), tb.bytes.size(), hipMemcpyHostToDevice));
    if (!plan->passes.empty())
        RK_HIP(hipMemcpy(base + off_pass, plan->passes.data(), plan->passes.size() * sizeof(rk_batch_pass),
                         hipMemcpyHostToDevice));
    if (!plan->tile_pass.empty())
        RK_HIP(hipMemcpy(base + off_tp, plan->tile_pass.data(), plan->tile_pass.size() * sizeof(rk_u32),
                         hipMemcpyHostToDevice));
    RK_HIP(hipMemset(base + off_cnt, 0, bytes - off_cnt));
    // the plan executes on callers' streams, which do not wait for the null
    // stream: the uploads above must have landed before this returns
    RK_HIP(hipDeviceSynchronize());
    plan->work.passes = (const rk_batch_pass*)(base + off_pass);
    plan->work.tile_pass = (const rk_u32*)(base + off_tp);
    plan->work.counters = (rk_u32*)(base + off_cnt);
    plan->work.state = (rk_u32*)(base + off_state);
    return RK_OK;
}

extern "C" int rk_batch_plan_create_ex(const rk_batch_item* items, int n_items, const rk_batch_options* options,
                                       rk_batch_plan** out) {
    if (!out || n_items < 0 || (n_items > 0 && !items)) return fail(RK_ERR_INVALID, "rk_batch_plan_create: bad arguments");
    *out = nullptr;
    for (int k = 0; k < n_items; ++k) {
        const rk_batch_item& it = items[k];
        if (it.length < 1 || it.count < 0) return fail(RK_ERR_INVALID, "batch item %d: bad length/count", k);
        if (!(it.flags & RK_BATCH_R2C) && it.count > 1 && (it.in_dist < it.length || it.out_dist < it.length))
            return fail(RK_ERR_INVALID, "batch item %d: distance smaller than the length", k);
        if (it.in_offset < 0 || it.out_offset < 0) return fail(RK_ERR_INVALID, "batch item %d: negative offset", k);
        if (it.flags & ~(RK_BATCH_SCALE | RK_BATCH_R2C)) return fail(RK_ERR_INVALID, "batch item %d: unknown flags", k);
        if (it.flags & RK_BATCH_R2C) {
            if (it.length < 4 || it.length % 2)
                return fail(RK_ERR_INVALID, "batch item %d: a real-to-complex length must be even and >= 4", k);
            if (it.count > 1 && (it.in_dist < it.length / 2 || it.out_dist < it.length / 2 + 1))
                return fail(RK_ERR_INVALID, "batch item %d: distances smaller than N/2 (in) or N/2 + 1 (out)", k);
        }
    }
    std::unique_ptr<rk_batch_plan, int (*)(rk_batch_plan*)> plan(new rk_batch_plan(), rk_batch_plan_destroy);
    if (options) {
        plan->opt = *options;
    } else {
        memset(&plan->opt, 0, sizeof(plan->opt));
    }
    if (plan->opt.mode != RK_BATCH_MODE_QUEUE && plan->opt.mode != RK_BATCH_MODE_GROUPED)
        return fail(RK_ERR_INVALID, "rk_batch_plan_create: unknown mode %u", plan->opt.mode);
    plan->dry = (plan->opt.flags & RK_BATCH_DRY_RUN) != 0;
    int rc = plan->opt.mode == RK_BATCH_MODE_GROUPED ? create_grouped(items, n_items, plan.get())
                                                     : create_queue(items, n_items, plan.get());
    if (rc != RK_OK) return rc;
    if (plan->opt.mode == RK_BATCH_MODE_GROUPED)
        for (int k = 0; k < n_items; ++k) plan->n_items += items[k].count;
    *out = plan.release();
    return RK_OK;
}

extern "C" int rk_batch_plan_create(const rk_batch_item* items, int n_items, rk_batch_plan** out) {
    return rk_batch_plan_create_ex(items, n_items, nullptr, out);
}

extern "C" int rk_batch_plan_workspace_size(const rk_batch_plan* plan, size_t* bytes) {
    if (!plan || !bytes) return fail(RK_ERR_INVALID, "rk_batch_plan_workspace_size: null argument");
    *bytes = plan->ws_bytes;
    return RK_OK;
}

extern "C" int rk_batch_execute(const rk_batch_plan* plan, void* out, const void* in, int inverse, void* workspace,
                                void* stream) {
    if (!plan) return fail(RK_ERR_INVALID, "rk_batch_execute: null plan");
    if (plan->dry) return fail(RK_ERR_INVALID, "rk_batch_execute: a dry-run plan cannot execute");
    if (!out || !in) return fail(RK_ERR_INVALID, "rk_batch_execute: null in/out pointer");
    if (plan->ws_bytes && !workspace)
        return fail(RK_ERR_INVALID, "rk_batch_execute: plan needs %zu bytes of workspace", plan->ws_bytes);
    if (inverse && plan->r2c)
        return fail(RK_ERR_UNSUPPORTED, "rk_batch_execute: real-to-complex items are forward only");
    hipStream_t s = (hipStream_t)stream;
    if (plan->opt.mode == RK_BATCH_MODE_GROUPED) {
        for (auto* p : plan->plans) {
            int rc = rk_fft_execute(p, out, in, inverse, workspace, stream);
            if (rc != RK_OK) return rc;
        }
        return RK_OK;
    }
    for (const rk::CopyArgs& c0 : plan->copies) {
        rk::CopyArgs c = c0;
        c.in = (const float2*)in;
        c.out = (float2*)out;
        int rc = rk::launch_copy(c, s);
        if (rc != RK_OK) return rc;
    }
    if (plan->work.total == 0) return RK_OK;
    rk_batch_work w = plan->work;
    w.in = (const float2*)in;
    w.out = (float2*)out;
    w.ws = (float2*)workspace;
    w.inverse = inverse ? 1u : 0u;
    void* params[] = {&w};
    RK_HIP(hipLaunchKernel(rk::batch_kernel(plan->flavour), dim3(w.total), dim3(RK_BATCH_THREADS), params, 0, s));
    return RK_OK;
}

extern "C" int rk_batch_plan_stats(const rk_batch_plan* plan, int64_t* work_items, int* launches) {
    if (!plan) return fail(RK_ERR_INVALID, "rk_batch_plan_stats: null plan");
    if (work_items) *work_items = plan->n_items;
    if (launches) *launches = plan->launches;
    return RK_OK;
}

extern "C" int rk_batch_plan_info(const rk_batch_plan* plan, rk_batch_info* info) {
    if (!plan || !info) return fail(RK_ERR_INVALID, "rk_batch_plan_info: null argument");
    memset(info, 0, sizeof(*info));
    info->items = plan->n_items;
    info->passes = plan->opt.mode == RK_BATCH_MODE_GROUPED ? plan->launches : (int64_t)plan->passes.size();
    info->tiles = plan->work.total;
    info->workspace_bytes = (int64_t)plan->ws_bytes;
    info->levels = plan->levels;
    info->launches = plan->launches;
    info->mode = (int32_t)plan->opt.mode;
    info->lds_bytes = rk::batch_lds_bytes(plan->flavour);
    info->flavour = plan->flavour;
    return RK_OK;
}

extern "C" int rk_batch_plan_check(rk_batch_plan* plan, int* hit) {
    if (!plan || !hit) return fail(RK_ERR_INVALID, "rk_batch_plan_check: null argument");
    *hit = 0;
    if (plan->dry || !plan->work.state) return RK_OK;
    rk_u32 flag = 0;
    RK_HIP(hipDeviceSynchronize());
    RK_HIP(hipMemcpy(&flag, plan->work.state + 2, sizeof(flag), hipMemcpyDeviceToHost));
    if (flag) {
        // the counters of the failed execute may not have returned to zero:
        // clear them with the flag (the device is idle after the sync above)
        RK_HIP(hipMemset(plan->work.counters, 0, plan->counter_bytes));
        RK_HIP(hipMemset(plan->work.state, 0, 256));
        RK_HIP(hipDeviceSynchronize());
    }
    *hit = flag ? 1 : 0;
    return RK_OK;
}

// ---------------------------------------------------------------------------
// HIP graphs
// ---------------------------------------------------------------------------
struct rk_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

extern "C" int rk_graph_begin(void* stream) {
    RK_HIP(hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal));
    return RK_OK;
}

extern "C" int rk_graph_end(void* stream, rk_graph** out) {
    if (!out) return fail(RK_ERR_INVALID, "rk_graph_end: null argument");
    *out = nullptr;
    std::unique_ptr<rk_graph> g(new rk_graph());
    RK_HIP(hipStreamEndCapture((hipStream_t)stream, &g->graph));
    RK_HIP(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
    *out = g.release();
    return RK_OK;
}

extern "C" int rk_graph_launch(rk_graph* g, void* stream) {
    if (!g) return fail(RK_ERR_INVALID, "rk_graph_launch: null graph");
    RK_HIP(hipGraphLaunch(g->exec, (hipStream_t)stream));
    return RK_OK;
}

extern "C" int rk_graph_destroy(rk_graph* g) {
    if (!g) return RK_OK;
    if (g->exec) hipGraphExecDestroy(g->exec);
    if (g->graph) hipGraphDestroy(g->graph);
    delete g;
    return RK_OK;
}
