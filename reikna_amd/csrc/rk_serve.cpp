// rk_serve.cpp — a continuous-batching FFT request server behind
// rk_server_create / rk_server_submit / rk_server_progress.
//
// No reference counterpart (reikna has no serving layer).  Requests are 1-D
// or 2-D complex64 transforms (a batch of `count` each) whose input is
// already in device memory.  The server
//
//   * keeps one chain template per request shape resident: the pass chain of
//     the batch kernel (rk_batch_internal.hpp) with its twiddle / chirp /
//     Bluestein tables in device memory, built the first time the shape is
//     seen;
//   * admits queued requests in arrival order while their output and
//     workspace fit in the HBM budget (a request that can never fit is
//     refused at submit);
//   * launches every admitted request of a step together, as one launch of
//     the batch kernel, in the first free slot (a stream with its own queue
//     state); up to `slots` launches are in flight, so requests that arrive
//     while others run are admitted as soon as memory and a slot free up, and
//     their work overlaps on the device;
//   * reaps finished launches by their events, releases workspaces, and
//     reports each request's completion time (outputs stay allocated until
//     rk_server_release).
//
// Mode RK_SERVER_PLAIN runs the same requests the plain way instead: one
// single-shape plan per request shape (cached), each request launched by
// itself (its plan's passes) in the slot's stream, with the same admission.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <tuple>
#include <unordered_map>
#include <unordered_set>

#include "rk_batch_internal.hpp"

using rk::fail;

namespace {

constexpr uint32_t kSpinLimit = 1u << 24;
constexpr size_t kAlign = 256;

uint64_t now_ns() {
    return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

// First-fit allocator over [0, size) in kAlign units (host bookkeeping of one
// device block).
class Arena {
public:
    void init(size_t size) {
        size_ = size;
        free_.clear();
        free_[0] = size;
        used_ = 0;
    }
    bool alloc(size_t bytes, size_t* off) {
        bytes = (bytes + kAlign - 1) & ~(kAlign - 1);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                *off = it->first;
                const size_t rest = it->second - bytes, start = it->first + bytes;
                free_.erase(it);
                if (rest) free_[start] = rest;
                used_ += bytes;
                return true;
            }
        }
        return false;
    }
    void release(size_t off, size_t bytes) {
        bytes = (bytes + kAlign - 1) & ~(kAlign - 1);
        used_ -= bytes;
        auto next = free_.lower_bound(off);
        size_t start = off, len = bytes;
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                start = prev->first;
                len += prev->second;
                free_.erase(prev);
            }
        }
        if (next != free_.end() && off + bytes == next->first) {
            len += next->second;
            free_.erase(next);
        }
        free_[start] = len;
    }
    size_t used() const { return used_; }
    size_t size() const { return size_; }

private:
    size_t size_ = 0, used_ = 0;
    std::map<size_t, size_t> free_;
};

struct ShapeKey {
    int nd;
    int64_t n0, n1, count;
    bool operator<(const ShapeKey& o) const {
        return std::tie(nd, n0, n1, count) < std::tie(o.nd, o.n0, o.n1, o.count);
    }
};

struct Template {
    rk::Chain chain;
    void* tables = nullptr;      // device copy of the chain's tables
    rk_fft_plan* plain = nullptr;   // RK_SERVER_PLAIN: the single-shape plan
    size_t plain_ws = 0;
    uint64_t tiles = 0;
};

struct Request {
    rk_request r;
    const Template* t = nullptr;
    size_t out_off = 0, out_bytes = 0, ws_off = 0, ws_bytes = 0;
    uint64_t admitted_ns = 0;
    bool plain = false;   // runs by itself with its single-shape plan
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    rk_u32* flag = nullptr;          // pinned host copy of the launch's error flag
    void* dev = nullptr;             // passes | tile_pass | counters | state
    size_t cap_passes = 0, cap_tiles = 0, cap_lines = 0, bytes = 0;
    std::vector<rk_batch_pass> passes;
    std::vector<rk_u32> tile_pass;
    std::vector<Request> reqs;       // in flight
    bool busy = false;
    bool plain = false;              // the launch group runs single-shape plans
};

}  // namespace

struct rk_server {
    rk_server_config cfg;
    void* arena_base = nullptr;
    Arena arena;
    std::map<ShapeKey, std::unique_ptr<Template>> templates;
    std::deque<Request> pending;
    std::vector<Slot> slots;
    std::unordered_map<uint64_t, std::pair<size_t, size_t>> outputs;   // id -> (offset, bytes)
    std::unordered_set<uint64_t> live;   // ids submitted and not yet released
    std::vector<rk_completion> finished;
    uint64_t launches = 0, completed = 0, failed = 0;
};

static int destroy_server(rk_server* s) {
    if (!s) return RK_OK;
    for (auto& sl : s->slots) {
        if (sl.stream) hipStreamSynchronize(sl.stream);
        if (sl.done) hipEventDestroy(sl.done);
        if (sl.stream) hipStreamDestroy(sl.stream);
        if (sl.dev) hipFree(sl.dev);
        if (sl.flag) hipHostFree(sl.flag);
    }
    for (auto& kv : s->templates) {
        if (kv.second->tables) hipFree(kv.second->tables);
        if (kv.second->plain) rk_fft_plan_destroy(kv.second->plain);
    }
    if (s->arena_base) hipFree(s->arena_base);
    delete s;
    return RK_OK;
}

extern "C" int rk_server_destroy(rk_server* s) { return destroy_server(s); }

extern "C" int rk_server_create(const rk_server_config* cfg, rk_server** out) {
    if (!cfg || !out) return fail(RK_ERR_INVALID, "rk_server_create: null argument");
    *out = nullptr;
    if (cfg->budget_bytes < kAlign) return fail(RK_ERR_INVALID, "rk_server_create: budget too small");
    if (cfg->slots < 1 || cfg->slots > 16) return fail(RK_ERR_INVALID, "rk_server_create: slots must be 1..16");
    if (cfg->mode != RK_SERVER_BATCHED && cfg->mode != RK_SERVER_PLAIN)
        return fail(RK_ERR_INVALID, "rk_server_create: unknown mode");
    std::unique_ptr<rk_server, int (*)(rk_server*)> s(new rk_server(), destroy_server);
    s->cfg = *cfg;
    if (s->cfg.max_requests <= 0) s->cfg.max_requests = 1024;
    RK_HIP(hipMalloc(&s->arena_base, cfg->budget_bytes));
    s->arena.init(cfg->budget_bytes);
    s->slots.resize(cfg->slots);
    for (auto& sl : s->slots) {
        RK_HIP(hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking));
        RK_HIP(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
        RK_HIP(hipHostMalloc((void**)&sl.flag, 64, hipHostMallocDefault));
        *sl.flag = 0;
    }
    *out = s.release();
    return RK_OK;
}

static size_t request_elems(const rk_request& r) {
    return (size_t)r.count * (size_t)r.shape[0] * (size_t)(r.ndim == 2 ? r.shape[1] : 1);
}

// Requests that run by themselves with their single-shape plan.
static bool runs_plain(const rk_server* s, const rk_request& r) {
    return s->cfg.mode == RK_SERVER_PLAIN ||
           (s->cfg.plain_min_elements > 0 && (int64_t)request_elems(r) >= s->cfg.plain_min_elements);
}

static int get_template(rk_server* s, const rk_request& r, const Template** out) {
    ShapeKey key{r.ndim, r.shape[0], r.ndim == 2 ? r.shape[1] : 1, r.count};
    auto it = s->templates.find(key);
    if (it != s->templates.end()) {
        *out = it->second.get();
        return RK_OK;
    }
    std::unique_ptr<Template> t(new Template());
    if (runs_plain(s, r)) {
        rk_fft_desc d;
        memset(&d, 0, sizeof(d));
        d.dtype = RK_C64;
        d.ndim = 1 + r.ndim;
        d.shape[0] = r.count;
        int64_t inner = 1;
        for (int k = r.ndim; k >= 1; --k) {
            d.shape[k] = r.shape[k - 1];
            d.in_strides[k] = d.out_strides[k] = inner;
            inner *= r.shape[k - 1];
        }
        d.in_strides[0] = d.out_strides[0] = inner;
        d.naxes = r.ndim;
        for (int k = 0; k < r.ndim; ++k) d.axes[k] = k + 1;
        int rc = rk_fft_plan_create(&d, &t->plain);
        if (rc != RK_OK) return rc;
        rk_fft_plan_workspace_size(t->plain, &t->plain_ws);
    } else {
        rk::ItemShape sh;
        sh.nd = r.ndim;
        sh.n[0] = r.shape[0];
        sh.n[1] = r.ndim == 2 ? r.shape[1] : 1;
        sh.count = r.count;
        sh.in_dist = sh.out_dist = sh.elements();
        rk::TableBuilder tb;
        int rc = rk::build_chain(sh, rk::batch_limits(nullptr), tb, t->chain);
        if (rc != RK_OK) return rc;
        if (!tb.bytes.empty()) {
            RK_HIP(hipMalloc(&t->tables, tb.bytes.size()));
            RK_HIP(hipMemcpy(t->tables, tb.bytes.data(), tb.bytes.size(), hipMemcpyHostToDevice));
            // a copy from pageable memory may return before its DMA lands, and
            // the slots' streams do not wait for the null stream
            RK_HIP(hipDeviceSynchronize());
        }
        for (auto& cp : t->chain.passes) {
            rk::resolve_tables(cp.pp.args, cp.pp, (const char*)t->tables);
            t->tiles += cp.pp.args.n_tiles;
        }
    }
    *out = t.get();
    s->templates[key] = std::move(t);
    return RK_OK;
}

extern "C" int rk_server_submit(rk_server* s, const rk_request* reqs, int n) {
    if (!s || (n > 0 && !reqs) || n < 0) return fail(RK_ERR_INVALID, "rk_server_submit: bad arguments");
    // every request is checked (and its template built) before any is
    // queued: a batch is accepted whole or not at all
    std::vector<Request> batch;
    batch.reserve((size_t)n);
    std::unordered_set<uint64_t> ids;
    for (int i = 0; i < n; ++i) {
        const rk_request& r = reqs[i];
        if ((r.ndim != 1 && r.ndim != 2) || r.count < 1 || r.shape[0] < 2 || (r.ndim == 2 && r.shape[1] < 2) || !r.in)
            return fail(RK_ERR_INVALID, "rk_server_submit: request %d: bad shape/count/input", i);
        if (s->live.count(r.id))
            return fail(RK_ERR_INVALID, "rk_server_submit: request %d: id %llu is still in use", i,
                        (unsigned long long)r.id);
        if (!ids.insert(r.id).second)
            return fail(RK_ERR_INVALID, "rk_server_submit: request %d: id %llu appears twice in the batch", i,
                        (unsigned long long)r.id);
        Request q;
        q.r = r;
        int rc = get_template(s, r, &q.t);
        if (rc != RK_OK) return rc;
        q.out_bytes = request_elems(r) * sizeof(float2);
        q.plain = runs_plain(s, r);
        q.ws_bytes = q.plain ? q.t->plain_ws : q.t->chain.ws_total() * sizeof(float2);
        const size_t need = ((q.out_bytes + kAlign - 1) & ~(kAlign - 1)) + ((q.ws_bytes + kAlign - 1) & ~(kAlign - 1));
        if (need > s->arena.size())
            return fail(RK_ERR_BUDGET, "rk_server_submit: request %d needs %zu bytes, the budget is %zu", i, need,
                        s->arena.size());
        batch.push_back(q);
    }
    const uint64_t t = now_ns();
    for (Request& q : batch) {
        if (q.r.t_arrival_ns == 0) q.r.t_arrival_ns = t;
        s->live.insert(q.r.id);
        s->pending.push_back(q);
    }
    return RK_OK;
}

static void finish_slot(rk_server* s, Slot& sl, int status) {
    const uint64_t t = now_ns();
    for (Request& q : sl.reqs) {
        if (q.ws_bytes) s->arena.release(q.ws_off, q.ws_bytes);
        rk_completion c;
        memset(&c, 0, sizeof(c));
        c.id = q.r.id;
        c.out = (char*)s->arena_base + q.out_off;
        c.t_arrival_ns = q.r.t_arrival_ns;
        c.t_admitted_ns = q.admitted_ns;
        c.t_done_ns = t;
        c.status = status;
        s->finished.push_back(c);
        s->completed++;
    }
    sl.reqs.clear();
    sl.busy = false;
}

static int launch_batched(rk_server* s, Slot& sl) {
    std::vector<rk::QueuedPass> queued;
    for (size_t k = 0; k < sl.reqs.size(); ++k) {
        const Request& q = sl.reqs[k];
        const rk::Chain& ch = q.t->chain;
        char* arena = (char*)s->arena_base;
        float2* out = (float2*)(arena + q.out_off);
        float2* ws = (float2*)(arena + q.ws_off);
        for (size_t i = 0; i < ch.passes.size(); ++i) {
            const rk::ChainPass& cp = ch.passes[i];
            rk::QueuedPass qp;
            qp.cp = &cp;
            qp.a = cp.pp.args;
            qp.item = (int)k;
            qp.level = (int)i;
            qp.inverse = q.r.inverse ? 1u : 0u;
            qp.src = qp.dst = RK_BATCH_BUF_ABS;
            auto ptr = [&](int buf, rk_i64& off) -> float2* {
                if (buf == rk::BUF_IN) return (float2*)q.r.in;
                if (buf == rk::BUF_OUT) return out;
                if (buf == rk::BUF_WS1) off += (rk_i64)ch.ws_elems;
                return ws;
            };
            qp.a.in = ptr(cp.pp.src, qp.a.in_off);
            qp.a.out = ptr(cp.pp.dst, qp.a.out_off);
            queued.push_back(qp);
        }
    }
    int levels = 0, flavour = rk::BATCH_LIGHT;
    for (const Request& q : sl.reqs)
        if (q.t->chain.blue) flavour = rk::BATCH_BLUE;
    size_t lines = 0;
    const uint64_t total = rk::assemble_queue(queued, 0, flavour, sl.passes, sl.tile_pass, levels, lines);
    if (total == 0) return RK_OK;
    // device buffers of the slot (grown as needed; the kernel leaves its
    // counters and state zero at its end)
    const size_t line_bytes = RK_BATCH_LINE * sizeof(rk_u32);
    if (sl.passes.size() > sl.cap_passes || total > sl.cap_tiles || lines > sl.cap_lines) {
        RK_HIP(hipStreamSynchronize(sl.stream));
        const size_t cap_passes = std::max<size_t>(sl.passes.size() * 2, 1024);
        const size_t cap_tiles = std::max<size_t>(total * 2, 16384);
        const size_t cap_lines = std::max<size_t>(lines * 2, 1024);
        const size_t bytes = cap_passes * sizeof(rk_batch_pass) + cap_tiles * 4 + cap_lines * line_bytes + 256;
        // the new block is allocated and cleared before the old one goes, and
        // the capacities change only with it: a failure leaves the slot as it
        // was (its old block, if any, still valid for its old capacities)
        void* fresh = nullptr;
        if (s->cfg.fault_inject & RK_SERVER_FAULT_SLOT_ALLOC) {
            s->cfg.fault_inject &= ~RK_SERVER_FAULT_SLOT_ALLOC;
            return fail(RK_ERR_HIP, "rk_server: slot buffer of %zu bytes: allocation failed (injected)", bytes);
        }
        RK_HIP(hipMalloc(&fresh, bytes));
        // on the slot's stream: ordered before the descriptor uploads below
        // (the slot's stream does not wait for the null stream)
        const hipError_t me = hipMemsetAsync(fresh, 0, bytes, sl.stream);
        if (me != hipSuccess) {
            hipFree(fresh);
            return fail(RK_ERR_HIP, "rk_server: clearing a slot buffer: %s", hipGetErrorString(me));
        }
        if (sl.dev) hipFree(sl.dev);   // idle: the stream was synchronised above
        sl.dev = fresh;
        sl.bytes = bytes;
        sl.cap_passes = cap_passes;
        sl.cap_tiles = cap_tiles;
        sl.cap_lines = cap_lines;
    }
    char* base = (char*)sl.dev;
    rk_batch_work w;
    memset(&w, 0, sizeof(w));
    w.passes = (const rk_batch_pass*)base;
    w.tile_pass = (const rk_u32*)(base + sl.cap_passes * sizeof(rk_batch_pass));
    char* counters = base + sl.cap_passes * sizeof(rk_batch_pass) + sl.cap_tiles * 4;
    w.counters = (rk_u32*)counters;
    w.state = (rk_u32*)(counters + sl.cap_lines * line_bytes);
    w.n_passes = (rk_u32)sl.passes.size();
    w.total = (rk_u32)total;
    w.spin_limit = s->cfg.spin_limit ? s->cfg.spin_limit : kSpinLimit;
    w.poll_sleep_max = rk::batch_poll_sleep_max();
    RK_HIP(hipMemcpyAsync((void*)w.passes, sl.passes.data(), sl.passes.size() * sizeof(rk_batch_pass),
                          hipMemcpyHostToDevice, sl.stream));
    RK_HIP(hipMemcpyAsync((void*)w.tile_pass, sl.tile_pass.data(), sl.tile_pass.size() * 4, hipMemcpyHostToDevice,
                          sl.stream));
    void* params[] = {&w};
    RK_HIP(hipLaunchKernel(rk::batch_kernel(flavour), dim3(w.total), dim3(RK_BATCH_THREADS), params, 0, sl.stream));
    // the launch's error flag comes back with it (read when its event completes)
    *sl.flag = 0;
    RK_HIP(hipMemcpyAsync(sl.flag, w.state + 2, sizeof(rk_u32), hipMemcpyDeviceToHost, sl.stream));
    RK_HIP(hipEventRecord(sl.done, sl.stream));
    s->launches++;
    return RK_OK;
}

static int launch_plain(rk_server* s, Slot& sl) {
    char* arena = (char*)s->arena_base;
    for (const Request& q : sl.reqs) {
        int rc = rk_fft_execute(q.t->plain, arena + q.out_off, q.r.in, q.r.inverse, q.ws_bytes ? arena + q.ws_off : nullptr,
                                sl.stream);
        if (rc != RK_OK) return rc;
        s->launches++;
    }
    RK_HIP(hipEventRecord(sl.done, sl.stream));
    return RK_OK;
}

extern "C" int rk_server_progress(rk_server* s, rk_completion* done, int max_done, int* n_done) {
    if (!s || !n_done || (max_done > 0 && !done)) return fail(RK_ERR_INVALID, "rk_server_progress: bad arguments");
    // reap finished launches
    for (auto& sl : s->slots) {
        if (!sl.busy) continue;
        hipError_t e = hipEventQuery(sl.done);
        if (e == hipSuccess) {
            int status = 0;
            if (!sl.plain && *sl.flag) {
                // a dependency wait gave up: the launch's results are not valid
                // and its counters may not have returned to zero; clear the
                // slot's counters and state before its next launch
                status = 1;
                s->failed++;
                RK_HIP(hipMemsetAsync(sl.dev, 0, sl.bytes, sl.stream));
                *sl.flag = 0;
            }
            finish_slot(s, sl, status);
        } else if (e != hipErrorNotReady) {
            return fail(RK_ERR_HIP, "rk_server_progress: %s", hipGetErrorString(e));
        }
    }
    // admit and launch: each free slot takes requests from the head of the
    // queue, all of one kind (batched or single-shape plans), while their
    // output and workspace fit
    for (auto& sl : s->slots) {
        if (sl.busy || s->pending.empty()) continue;
        const uint64_t t = now_ns();
        sl.plain = s->pending.front().plain;
        while (!s->pending.empty() && (int)sl.reqs.size() < s->cfg.max_requests) {
            Request& q = s->pending.front();
            if (q.plain != sl.plain) break;
            size_t out_off = 0, ws_off = 0;
            if (!s->arena.alloc(q.out_bytes, &out_off)) break;
            if (q.ws_bytes && !s->arena.alloc(q.ws_bytes, &ws_off)) {
                s->arena.release(out_off, q.out_bytes);
                break;
            }
            q.out_off = out_off;
            q.ws_off = ws_off;
            q.admitted_ns = t;
            s->outputs[q.r.id] = {out_off, q.out_bytes};
            sl.reqs.push_back(q);
            s->pending.pop_front();
        }
        if (sl.reqs.empty()) break;   // memory is full: wait for a launch to finish
        sl.busy = true;
        int rc = sl.plain ? launch_plain(s, sl) : launch_batched(s, sl);
        if (rc != RK_OK) {
            // whatever part of the launch was enqueued ends before the
            // requests' memory is reused; the requests complete with status
            // 2 (not run), their workspaces are released and the slot is free
            hipStreamSynchronize(sl.stream);
            s->failed++;
            finish_slot(s, sl, 2);
            return rc;
        }
    }
    const int n = std::min<int>(max_done, (int)s->finished.size());
    for (int i = 0; i < n; ++i) done[i] = s->finished[i];
    s->finished.erase(s->finished.begin(), s->finished.begin() + n);
    *n_done = n;
    return RK_OK;
}

extern "C" int rk_server_release(rk_server* s, uint64_t id) {
    if (!s) return fail(RK_ERR_INVALID, "rk_server_release: null server");
    auto it = s->outputs.find(id);
    if (it == s->outputs.end()) return fail(RK_ERR_INVALID, "rk_server_release: unknown request %llu", (unsigned long long)id);
    s->arena.release(it->second.first, it->second.second);
    s->outputs.erase(it);
    s->live.erase(id);
    return RK_OK;
}

extern "C" int rk_server_stats(const rk_server* s, rk_server_stats_t* st) {
    if (!s || !st) return fail(RK_ERR_INVALID, "rk_server_stats: null argument");
    memset(st, 0, sizeof(*st));
    st->pending = (int64_t)s->pending.size();
    for (const auto& sl : s->slots)
        if (sl.busy) {
            st->in_flight_launches++;
            st->in_flight_requests += (int64_t)sl.reqs.size();
        }
    st->completed = (int64_t)s->completed;
    st->launches = (int64_t)s->launches;
    st->budget_used = s->arena.used();
    st->budget_bytes = s->arena.size();
    st->templates = (int64_t)s->templates.size();
    st->failed_launches = (int64_t)s->failed;
    return RK_OK;
}

extern "C" uint64_t rk_server_now_ns(void) { return now_ns(); }

// Open-loop replay of a request stream: each request is submitted when the
// clock reaches its t_arrival_ns, the server is progressed in between, and
// done[k] receives the k-th completion.  Outputs are released as they
// complete unless keep_outputs.  Gives up (RK_ERR_HIP) if nothing completes
// for `stall_ms` while requests are outstanding.
extern "C" int rk_server_replay(rk_server* s, const rk_request* reqs, int n, rk_completion* done, int keep_outputs,
                                int stall_ms) {
    if (!s || n < 0 || (n > 0 && (!reqs || !done))) return fail(RK_ERR_INVALID, "rk_server_replay: bad arguments");
    // every completion it sees must be one of `reqs`
    bool idle = s->pending.empty() && s->finished.empty();
    for (const auto& sl : s->slots) idle = idle && !sl.busy;
    if (!idle) return fail(RK_ERR_INVALID, "rk_server_replay: the server has requests of its own in flight");
    int next = 0, completed = 0;
    std::vector<rk_completion> buf(1024);
    uint64_t last_progress = now_ns();
    const uint64_t stall_ns = (uint64_t)(stall_ms > 0 ? stall_ms : 10000) * 1000000ull;
    while (completed < n) {
        const uint64_t t = now_ns();
        int k = next;
        while (k < n && reqs[k].t_arrival_ns <= t) ++k;
        if (k > next) {
            int rc = rk_server_submit(s, reqs + next, k - next);
            if (rc != RK_OK) return rc;
            next = k;
        }
        int got = 0;
        int rc = rk_server_progress(s, buf.data(), (int)buf.size(), &got);
        if (rc != RK_OK) return rc;
        for (int i = 0; i < got; ++i) {
            done[completed++] = buf[i];
            if (!keep_outputs) {
                rc = rk_server_release(s, buf[i].id);
                if (rc != RK_OK) return rc;
            }
        }
        if (got) {
            last_progress = t;
        } else if (next > completed && t - last_progress > stall_ns) {
            return fail(RK_ERR_HIP, "rk_server_replay: no request completed for %d ms", stall_ms > 0 ? stall_ms : 10000);
        }
    }
    return RK_OK;
}
