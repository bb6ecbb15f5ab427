// rk_batch.cpp — batched transforms of uneven lengths behind
// rk_batch_plan_create / rk_batch_execute.
//
// No reference counterpart: reikna compiles one computation per shape.  The
// batch plan groups the items by length; each group is one native FFT plan
// over a (count, length) array with the item's distances, so a batch of
// 512 uneven transforms costs one launch per pass per distinct length instead
// of one computation call per transform.
#include <algorithm>
#include <cstring>
#include <map>
#include <memory>

#include "rk_internal.hpp"

using rk::fail;

struct rk_batch_plan {
    struct Group {
        rk_fft_plan* plan = nullptr;
        size_t ws = 0;
    };
    std::vector<Group> groups;
    std::vector<rk_batch_item> items;
    size_t ws_bytes = 0;
    int64_t work_items = 0;
    int launches = 0;
};

extern "C" int rk_batch_plan_destroy(rk_batch_plan* plan) {
    if (!plan) return RK_OK;
    for (auto& g : plan->groups) rk_fft_plan_destroy(g.plan);
    delete plan;
    return RK_OK;
}

extern "C" int rk_batch_plan_create(const rk_batch_item* items, int n_items, rk_batch_plan** out) {
    if (!items || !out || n_items < 0) return fail(RK_ERR_INVALID, "rk_batch_plan_create: bad arguments");
    *out = nullptr;
    std::unique_ptr<rk_batch_plan, int (*)(rk_batch_plan*)> plan(new rk_batch_plan(), rk_batch_plan_destroy);
    for (int k = 0; k < n_items; ++k) {
        const rk_batch_item& it = items[k];
        if (it.length < 1 || it.count < 0) return fail(RK_ERR_INVALID, "batch item %d: bad length/count", k);
        if (it.count == 0) continue;
        if (it.count > 1 && (it.in_dist < it.length || it.out_dist < it.length))
            return fail(RK_ERR_INVALID, "batch item %d: distance smaller than the length", k);
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
        d.flags = (it.flags & RK_BATCH_SCALE) ? RK_FFT_NORM_FORWARD : 0u;
        rk_batch_plan::Group g;
        int rc = rk_fft_plan_create(&d, &g.plan);
        if (rc != RK_OK) return rc;
        rk_fft_plan_workspace_size(g.plan, &g.ws);
        plan->ws_bytes = std::max(plan->ws_bytes, g.ws);
        int np = 0;
        rk_fft_plan_num_passes(g.plan, &np);
        plan->launches += std::max(np, 1);
        plan->work_items += it.count;
        plan->groups.push_back(g);
        plan->items.push_back(it);
    }
    *out = plan.release();
    return RK_OK;
}

extern "C" int rk_batch_plan_workspace_size(const rk_batch_plan* plan, size_t* bytes) {
    if (!plan || !bytes) return fail(RK_ERR_INVALID, "rk_batch_plan_workspace_size: null argument");
    *bytes = plan->ws_bytes;
    return RK_OK;
}

extern "C" int rk_batch_execute(const rk_batch_plan* plan, void* out, const void* in, int inverse, void* workspace,
                                void* stream) {
    if (!plan) return fail(RK_ERR_INVALID, "rk_batch_execute: null plan");
    for (const auto& g : plan->groups) {
        int rc = rk_fft_execute(g.plan, out, in, inverse, workspace, stream);
        if (rc != RK_OK) return rc;
    }
    return RK_OK;
}

extern "C" int rk_batch_plan_stats(const rk_batch_plan* plan, int64_t* work_items, int* launches) {
    if (!plan) return fail(RK_ERR_INVALID, "rk_batch_plan_stats: null plan");
    if (work_items) *work_items = plan->work_items;
    if (launches) *launches = plan->launches;
    return RK_OK;
}
