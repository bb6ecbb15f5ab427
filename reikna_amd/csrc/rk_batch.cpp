// rk_batch.cpp — batched transforms of uneven lengths behind
// rk_batch_plan_create / rk_batch_execute.
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
#include <cstring>
#include <memory>
#include <numeric>

#include "rk_internal.hpp"
#include "fft_batch.h"

using rk::fail;

namespace {

constexpr int64_t kBatchSingleMax = 4096;
constexpr int64_t kBatchFactorMax = 256;
constexpr int64_t kBatchBlueMax = 2048;
constexpr uint32_t kSpinLimit = 1u << 24;   // ~2 s of 256-cycle sleeps

const rk::BatchCase* find_case(int L, int variant) {
    for (const auto& c : rk::batch_cases())
        if (c.L == L && c.variant == variant) return &c;
    return nullptr;
}

bool batch_length_ok(int64_t L) { return L <= kBatchSingleMax && find_case((int)L, rk::VAR_PLAIN) != nullptr; }

int batch_tile_width(int L, bool) {
    const rk::BatchCase* c = find_case(L, rk::VAR_PLAIN);
    return c ? c->C : 1;
}

struct ItemPass {
    rk::PassPlan pp;
    rk::AxisPass ap;
    int item = 0, level = 0, kcase = -1;
    bool first = false, last = false;
    float fwd_scale = 1.0f, inv_scale = 1.0f;
};

}  // namespace

struct rk_batch_plan {
    rk_batch_options opt;
    // grouped mode
    std::vector<rk_fft_plan*> plans;
    // queue mode
    std::vector<rk_batch_pass> passes;   // queue order
    std::vector<rk_u32> tile_pass;
    std::vector<rk::CopyArgs> copies;    // length-1 items
    void* dev = nullptr;                 // tables | passes | tile_pass | done | state
    rk_batch_work work;
    int64_t n_items = 0;
    int levels = 0;
    size_t ws_bytes = 0;
    int launches = 0;
    bool dry = false;
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
    const rk_batch_options& o = plan->opt;
    rk::ChainLimits lim;
    lim.single_max = o.single_max > 0 ? std::min<int64_t>(o.single_max, kBatchSingleMax) : kBatchSingleMax;
    lim.factor_max = o.factor_max > 0 ? std::min<int64_t>(o.factor_max, kBatchFactorMax) : kBatchFactorMax;
    lim.blue_single_max = o.blue_max > 0 ? std::min<int64_t>(o.blue_max, kBatchBlueMax) : kBatchBlueMax;
    lim.pow2_bluestein = true;
    lim.length_ok = batch_length_ok;

    rk::TableBuilder tb;
    std::vector<ItemPass> all;
    size_t ws_elems = 0;
    for (int k = 0; k < n_items; ++k) {
        const rk_batch_item& it = items[k];
        if (it.count == 0) continue;
        if (it.length == 1) {   // identity: a strided copy launched beside the queue
            rk::CopyArgs c;
            memset(&c, 0, sizeof(c));
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
        // the item's chain, over the tile cases the batch kernel has; a smooth
        // length whose passes have no case goes through Bluestein instead
        std::vector<rk::AxisPass> aps;
        std::vector<ItemPass> chain;
        for (uint32_t force : {0u, (uint32_t)RK_FFT_FORCE_BLUESTEIN}) {
            chain.clear();
            int rc = rk::axis_passes(it.length, force, lim, aps);
            if (rc != RK_OK) {
                if (force) return rc;
                continue;
            }
            bool ok = true;
            const int64_t one = it.count;
            size_t item_ws = 0;
            for (size_t pi = 0; pi < aps.size() && ok; ++pi) {
                const rk::AxisPass& ap = aps[pi];
                ItemPass ip;
                ip.ap = ap;
                ip.item = k;
                ip.level = (int)pi;
                ip.first = pi == 0;
                ip.last = pi + 1 == aps.size();
                // src / dst: in -> ws0 -> ws1 -> ws0 ... -> out
                const int64_t shape_in[2] = {one, ap.in_len}, shape_out[2] = {one, ap.out_len};
                rk::Layout ls, ld;
                if (ip.first) {
                    ls.ndim = 2;
                    ls.shape[0] = one;
                    ls.shape[1] = ap.in_len;
                    ls.strides[0] = it.in_dist;
                    ls.strides[1] = 1;
                    ls.offset = it.in_offset;
                    ip.pp.src = rk::BUF_IN;
                } else {
                    ls = rk::contiguous_layout(2, shape_in);
                    ls.offset = (int64_t)ws_elems + (int64_t)(((pi - 1) % 2) * (size_t)(one * ap.Nax));
                    ip.pp.src = rk::BUF_WS0;
                }
                if (ip.last) {
                    ld.ndim = 2;
                    ld.shape[0] = one;
                    ld.shape[1] = ap.out_len;
                    ld.strides[0] = it.out_dist;
                    ld.strides[1] = 1;
                    ld.offset = it.out_offset;
                    ip.pp.dst = rk::BUF_OUT;
                } else {
                    ld = rk::contiguous_layout(2, shape_out);
                    ld.offset = (int64_t)ws_elems + (int64_t)((pi % 2) * (size_t)(one * ap.Nax));
                    ip.pp.dst = rk::BUF_WS0;
                    item_ws = std::max(item_ws, (size_t)((pi % 2) + 1) * (size_t)(one * ap.Nax));
                }
                int rc2 = rk::describe_pass(ap, ls, ld, 1, batch_tile_width, ip.pp);
                if (rc2 != RK_OK) return rc2;
                const rk::BatchCase* bc = find_case(ap.L, ip.pp.variant);
                if (!bc || bc->C != ip.pp.C || !ip.pp.loops.empty()) {
                    ok = false;
                    break;
                }
                ip.kcase = bc->id;
                chain.push_back(ip);
            }
            if (!ok) {
                if (force) return fail(RK_ERR_NO_KERNEL, "batch item %d (length %lld): no tile case for its passes", k,
                                       (long long)it.length);
                continue;
            }
            ws_elems += (item_ws + 31) & ~size_t(31);   // 256-byte aligned regions
            break;
        }
        const float scale = (float)(1.0 / (double)it.length);
        const bool norm_fwd = (it.flags & RK_BATCH_SCALE) != 0;
        for (ItemPass& ip : chain) {
            if (ip.last) {
                ip.fwd_scale = norm_fwd ? scale : 1.0f;
                ip.inv_scale = norm_fwd ? 1.0f : scale;
            }
            int rc = rk::build_tables(tb, ip.pp, ip.ap, it.length);
            if (rc != RK_OK) return rc;
            all.push_back(ip);
        }
        plan->n_items += it.count;
    }
    for (const auto& c : plan->copies) plan->n_items += (int64_t)c.total;

    // queue order: level by level; inside a level the largest tiles first
    // (longest-processing-time first), or item order
    std::vector<size_t> order(all.size());
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
        const ItemPass &x = all[a], &y = all[b];
        if (x.level != y.level) return x.level < y.level;
        if (o.order == 1) return x.item < y.item;
        const int64_t tx = (int64_t)x.pp.C * x.pp.L, ty = (int64_t)y.pp.C * y.pp.L;
        if (tx != ty) return tx > ty;
        return x.item < y.item;
    });
    std::vector<int> pos(all.size());
    for (size_t q = 0; q < order.size(); ++q) pos[order[q]] = (int)q;

    uint64_t total = 0;
    plan->passes.resize(all.size());
    for (size_t q = 0; q < order.size(); ++q) {
        const size_t i = order[q];
        const ItemPass& ip = all[i];
        rk_batch_pass& bp = plan->passes[q];
        memset(&bp, 0, sizeof(bp));
        bp.a = ip.pp.args;
        bp.first = total;
        bp.kcase = (rk_u32)ip.kcase;
        bp.src = ip.pp.src == rk::BUF_IN ? RK_BATCH_BUF_IN : RK_BATCH_BUF_WS;
        bp.dst = ip.pp.dst == rk::BUF_OUT ? RK_BATCH_BUF_OUT : RK_BATCH_BUF_WS;
        bp.wait = ip.first ? -1 : pos[i - 1];   // the item's previous pass
        bp.wait_tiles = ip.first ? 0u : (rk_u32)all[i - 1].pp.args.n_tiles;
        bp.signal = ip.last ? 0u : 1u;
        bp.item_first = ip.first ? 1u : 0u;
        bp.item_last = ip.last ? 1u : 0u;
        bp.fwd_scale = ip.fwd_scale;
        bp.inv_scale = ip.inv_scale;
        total += ip.pp.args.n_tiles;
        plan->levels = std::max(plan->levels, ip.level + 1);
    }
    if (total >= (uint64_t(1) << 31)) return fail(RK_ERR_UNSUPPORTED, "batch has too many tiles (%llu)", (unsigned long long)total);
    plan->tile_pass.resize(total);
    for (size_t q = 0; q < plan->passes.size(); ++q) {
        const rk_batch_pass& bp = plan->passes[q];
        for (rk_u64 t = 0; t < bp.a.n_tiles; ++t) plan->tile_pass[bp.first + t] = (rk_u32)q;
    }
    plan->ws_bytes = ws_elems * sizeof(float2);
    plan->launches = (total ? 1 : 0) + (int)plan->copies.size();

    memset(&plan->work, 0, sizeof(plan->work));
    plan->work.n_passes = (rk_u32)plan->passes.size();
    plan->work.total = (rk_u32)total;
    plan->work.spin_limit = kSpinLimit;
    if (plan->dry) return RK_OK;

    // one device allocation: tables | passes | tile_pass | done | state
    auto align = [](size_t x) { return (x + 255) & ~size_t(255); };
    const size_t off_pass = align(tb.bytes.size());
    const size_t off_tp = off_pass + align(plan->passes.size() * sizeof(rk_batch_pass));
    const size_t off_done = off_tp + align(plan->tile_pass.size() * sizeof(rk_u32));
    const size_t off_state = off_done + align(plan->passes.size() * sizeof(rk_u32));
    const size_t bytes = off_state + 256;
    RK_HIP(hipMalloc(&plan->dev, bytes));
    char* base = (char*)plan->dev;
    // table pointers (build_tables recorded offsets into the table block)
    for (size_t q = 0; q < order.size(); ++q) {
        const rk::PassPlan& pp = all[order[q]].pp;
        rk_pass_args& a = plan->passes[q].a;
        auto tab = [&](int64_t off) -> const void* { return off < 0 ? nullptr : base + off; };
        a.twl = (const float2*)tab(pp.twl_off);
        a.tw_hi = (const double2*)tab(pp.twhi_off);
        a.tw_lo = (const double2*)tab(pp.twlo_off);
        a.chirp = (const float2*)tab(pp.chirp_off);
        a.bspec = (const float2*)tab(pp.bspec_off);
        a.tile_base = 0;
        a.xcd_group = 0;
    }
    if (!tb.bytes.empty()) RK_HIP(hipMemcpy(base, tb.bytes.data(), tb.bytes.size(), hipMemcpyHostToDevice));
    if (!plan->passes.empty())
        RK_HIP(hipMemcpy(base + off_pass, plan->passes.data(), plan->passes.size() * sizeof(rk_batch_pass),
                         hipMemcpyHostToDevice));
    if (!plan->tile_pass.empty())
        RK_HIP(hipMemcpy(base + off_tp, plan->tile_pass.data(), plan->tile_pass.size() * sizeof(rk_u32),
                         hipMemcpyHostToDevice));
    RK_HIP(hipMemset(base + off_done, 0, bytes - off_done));
    plan->work.passes = (const rk_batch_pass*)(base + off_pass);
    plan->work.tile_pass = (const rk_u32*)(base + off_tp);
    plan->work.done = (rk_u32*)(base + off_done);
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
        if (it.count > 1 && (it.in_dist < it.length || it.out_dist < it.length))
            return fail(RK_ERR_INVALID, "batch item %d: distance smaller than the length", k);
        if (it.in_offset < 0 || it.out_offset < 0) return fail(RK_ERR_INVALID, "batch item %d: negative offset", k);
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
    RK_HIP(hipLaunchKernel(rk::batch_kernel(), dim3(w.total), dim3(RK_BATCH_THREADS), params, 0, s));
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
    info->lds_bytes = rk::batch_lds_bytes();
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
        const rk_u32 zero = 0;
        RK_HIP(hipMemcpy(plan->work.state + 2, &zero, sizeof(zero), hipMemcpyHostToDevice));
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
