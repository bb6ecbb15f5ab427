/*
 * reikna_amd.h — C-ABI of the MI355X (gfx950) FFT library.
 *
 * The reference (reikna) has no native code: its FFT is a Python planner that
 * renders Mako templates and enqueues them through grunnur.  The native
 * boundary it would bind is the kernel-level call site
 *
 *     plan.kernel_call(TEMPLATE.get_def("fft_local"|"fft_global"),
 *                      [out, in, (kweights), inverse], ...)
 *        reikna/fft/fft.py:704-711, executed by StaticKernel.__call__
 *        at reikna/core/computation.py:795
 *
 * together with grunnur's device runtime (Array / Queue / Buffer:
 * reikna/core/computation.py:7-16,401,444,683-693).  Every entry point
 * below replaces one of those pieces; the comment on each cites it.
 *
 * Conventions
 *   - Every function returns an int status: 0 (RK_OK) on success, a negative
 *     RK_ERR_* code otherwise; rk_last_error() then returns a thread-local
 *     message.  Nothing here throws across the boundary.
 *   - Strides and offsets are in ELEMENTS (complex64 = 8 bytes, complex128 = 16), not bytes.
 *   - Streams are hipStream_t passed as void* (NULL = the default stream).
 *   - Plans are immutable after creation: concurrent rk_fft_execute calls on
 *     different streams with different workspaces are safe.
 */
#ifndef REIKNA_AMD_H
#define REIKNA_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RK_MAX_DIMS 8

enum rk_status {
    RK_OK = 0,
    RK_ERR_INVALID = -1,     /* bad argument (Python maps it to ValueError) */
    RK_ERR_HIP = -2,         /* a HIP runtime call failed */
    RK_ERR_NO_KERNEL = -3,   /* no compiled kernel for a pass and JIT failed */
    RK_ERR_UNSUPPORTED = -4, /* valid request this build does not implement */
    RK_ERR_BUDGET = -5       /* request exceeds an HBM budget */
};

enum rk_dtype { RK_C64 = 0, RK_C128 = 1 };   /* complex64, complex128 (FFT plans only) */

/* ------------------------------------------------------------------ */
/* Errors                                                              */
/* ------------------------------------------------------------------ */

/* Thread-local text of the last failure on the calling thread. */
const char* rk_last_error(void);

/* Library version string, e.g. "reikna_amd 0.1 gfx950". */
const char* rk_version(void);

/* ------------------------------------------------------------------ */
/* Device runtime — replaces grunnur Array/Queue/Buffer                */
/* (reikna/core/computation.py:401 Array.from_host, :689 Array.empty_like, */
/*  :594-597 Queue.synchronize, :686 Buffer.allocate)                  */
/* ------------------------------------------------------------------ */

typedef struct {
    char name[256];
    char arch[64];              /* e.g. "gfx950" */
    int compute_units;
    int wavefront_size;
    int max_threads_per_block;
    size_t lds_per_block;       /* bytes of LDS a work-group may use */
    size_t total_global_mem;
    int l2_cache_bytes;
    int clock_khz;
    int memory_clock_khz;
    int memory_bus_width;
} rk_device_info;

int rk_device_count(int* count);
int rk_set_device(int device);
int rk_get_device_info(int device, rk_device_info* info);

int rk_malloc(void** ptr, size_t bytes);
int rk_free(void* ptr);
int rk_memcpy_htod(void* dst, const void* src, size_t bytes, void* stream);
int rk_memcpy_dtoh(void* dst, const void* src, size_t bytes, void* stream);
int rk_memcpy_dtod(void* dst, const void* src, size_t bytes, void* stream);
int rk_memset(void* dst, int value, size_t bytes, void* stream);
int rk_mem_info(size_t* free_bytes, size_t* total_bytes);

int rk_stream_create(void** stream);
int rk_stream_destroy(void* stream);
int rk_stream_synchronize(void* stream);
int rk_device_synchronize(void);

int rk_event_create(void** event);
int rk_event_destroy(void* event);
int rk_event_record(void* event, void* stream);
int rk_event_synchronize(void* event);
int rk_event_query(void* event, int* done);
int rk_event_elapsed_ms(void* start, void* stop, float* ms);
int rk_stream_wait_event(void* stream, void* event);

/* ------------------------------------------------------------------ */
/* Run-time compiled modules — replaces grunnur StaticKernel's driver JIT */
/* (reikna/core/computation.py:579-591)                                */
/* ------------------------------------------------------------------ */

/* Compile HIP C++ source for the current device with hiprtc.  The generated
   code object is cached on disk (RK_CACHE_DIR, default ~/.cache/reikna_amd),
   keyed by source, options and target.  include_dir is added as -I. */
int rk_module_compile(const char* source, const char* options,
                      void** module, char* log, size_t log_size);
/* Compile only (no device needed): returns RK_OK if the source compiles for
   gfx950, else RK_ERR_NO_KERNEL with the compiler log in `log`. */
int rk_module_check(const char* source, const char* options, char* log, size_t log_size);
int rk_module_get_function(void* module, const char* name, void** function);
int rk_module_unload(void* module);
/* Launch a run-time compiled kernel: args is a packed argument buffer laid out
   exactly as the kernel's parameter list (HIP_LAUNCH_PARAM_BUFFER_POINTER). */
int rk_launch(void* function, const uint32_t grid[3], const uint32_t block[3],
              uint32_t shared_bytes, void* stream, const void* args,
              size_t args_bytes);

/* ------------------------------------------------------------------ */
/* FFT plans — replace FFT._build_plan + the fft_local/fft_global chain */
/* (reikna/fft/fft.py:532-584 get_fft_kernels, :664-721 buffer routing, */
/*  :723-756 _build_plan, fft.mako:750-814 fft_local, :816-1038 fft_global) */
/* ------------------------------------------------------------------ */

typedef struct rk_fft_plan rk_fft_plan;

typedef struct {
    int dtype;                          /* RK_C64 or RK_C128 */
    int ndim;                           /* 1..RK_MAX_DIMS */
    int64_t shape[RK_MAX_DIMS];
    int64_t in_strides[RK_MAX_DIMS];    /* elements */
    int64_t out_strides[RK_MAX_DIMS];   /* elements */
    int64_t in_offset;                  /* elements */
    int64_t out_offset;                 /* elements */
    int naxes;                          /* 0 = transform over all axes */
    int axes[RK_MAX_DIMS];              /* non-negative; order as given */
    uint32_t flags;                     /* RK_FFT_* below */
} rk_fft_desc;

/* flags */
#define RK_FFT_FORCE_BLUESTEIN 0x1u     /* use chirp-z even for smooth lengths */
#define RK_FFT_FORCE_MULTIPASS 0x2u     /* split every axis into >= 2 passes */
#define RK_FFT_NO_JIT          0x4u     /* never run hiprtc for missing lengths */
#define RK_FFT_NORM_FORWARD    0x8u     /* scale the forward result by 1/prod(N), not the inverse */
#define RK_FFT_OUT_LAST_ONLY   0x10u    /* only the last pass touches `out`: earlier axes finish in workspace */
#define RK_FFT_DRY_RUN         0x20u    /* plan only: no device tables, no run-time compilation;
                                           the plan can be inspected but not executed */

/* Build a plan: radix factorisation, pass chain, buffer routing, and all
   device tables (twiddles in FP64 rounded once to FP32, Bluestein chirp from
   exact integer n^2 mod 2N, Bluestein kernel spectrum by an FP64 host FFT).
   Replaces FFT.__init__ + _build_plan (fft.py:630-645, :723-756) and the
   kweights persistent array (fft.py:216-236, :691-692). */
int rk_fft_plan_create(const rk_fft_desc* desc, rk_fft_plan** plan);

/* Planner knobs of one plan (rk_fft_plan_create_tuned).  Every field has a
   "planner's choice" value; a knob the plan cannot honour (a first factor that
   does not divide an axis, a tile wider than 1024 threads or 159 KB of LDS)
   is ignored.  Replaces reikna's fixed local-size limits and their halving
   retry (reikna/fft/fft.py:700-756) with knobs a measured search
   (reikna_amd.fft.autotune.autotune_fft) chooses. */
typedef struct {
    int32_t factor_max;     /* longest pass of a multi-pass chain (0: 256) */
    int32_t first_factor;   /* length of the first pass of every multi-pass axis (0: the balanced split) */
    int32_t col_tile;       /* instances per tile of column-mapped passes (0: planner's) */
    int32_t row_tile;       /* instances per tile of row-mapped passes (0: planner's) */
    int64_t chunk_bytes;    /* chunk size of chained passes (-1: planner's, 0: no chunks) */
    int32_t fixed_modes;    /* kernels with load/store modes compiled in (-1: planner's, 0: off, 1: on) */
    int32_t reserved;       /* 0 */
} rk_fft_tuning;

/* rk_fft_plan_create with planner knobs (NULL: the planner's choices).  The
   RK_ROW_TILE / RK_COL_TILE / RK_FACTOR_MAX / RK_CHUNK_BYTES /
   RK_FIXED_MODES environment overrides apply only where a knob is left at
   the planner's choice. */
int rk_fft_plan_create_tuned(const rk_fft_desc* desc, const rk_fft_tuning* tuning, rk_fft_plan** plan);

/* Bytes of scratch device memory rk_fft_execute needs (0 if none).
   Replaces the plan's temp_array allocations (computation.py:404-434,
   fft.py:681-689). */
int rk_fft_plan_workspace_size(const rk_fft_plan* plan, size_t* bytes);

/* Enqueue the transform on `stream`.  inverse = 0: forward, exp(-2 pi i kn/N);
   inverse != 0: backward with 1/prod(N) normalisation, so that
   IFFT(FFT(x)) = x (fft.mako:741-747).  in == out is allowed.
   `workspace` must hold rk_fft_plan_workspace_size bytes (may be NULL when
   that is 0).  Asynchronous.  Replaces ComputationCallable.__call__ over the
   FFT kernel chain (computation.py:766-801). */
int rk_fft_execute(const rk_fft_plan* plan, void* out, const void* in,
                   int inverse, void* workspace, void* stream);

/* rk_fft_execute for callers that pass buffer sizes: refuses, before any
   launch, an input or output buffer shorter than the extent its layout in the
   plan's descriptor reaches (shape, strides, offset), a layout reaching before
   its buffer's start, or a workspace shorter than rk_fft_plan_workspace_size.
   Replaces the same call (reikna/core/computation.py:795), where grunnur's
   typed Array arguments carry their sizes. */
int rk_fft_execute_checked(const rk_fft_plan* plan, void* out, size_t out_bytes,
                           const void* in, size_t in_bytes, int inverse,
                           void* workspace, size_t workspace_bytes, void* stream);

/* Same as rk_fft_execute, with the library's own device scratch array as the
   workspace: it calls no hipMalloc/hipFree and allocates nothing.  No
   counterpart in the reference: for callers that cannot allocate (code under
   HIP graph capture, tight serving loops, harnesses that forbid allocation).
   Limit: plans whose workspace (rk_fft_plan_workspace_size) exceeds
   RK_SCRATCH_BYTES are refused with RK_ERR_INVALID.  The array (one per GPU,
   held until the process ends) is shared by every call on that GPU and keeps
   the last call's contents; calls whose work overlaps on one GPU race. */
#define RK_SCRATCH_BYTES (64ull << 20)
int rk_fft_execute_scratch(const rk_fft_plan* plan, void* out, const void* in,
                           int inverse, void* stream);

/* Same, for plans whose first/last pass kernels were replaced by fused
   (transformation-carrying) kernels with rk_fft_plan_set_pass_kernel: the
   byte blobs are copied into the user-data slots of the first pass
   (load side) and the last pass (store side).  Each is at most
   RK_FUSED_BLOB_BYTES.  Replaces the load_combined_idx/store_combined_idx
   splicing of transformations (fft.mako:220-221,871-872,1000-1002). */
#define RK_FUSED_BLOB_BYTES 512
int rk_fft_execute_fused(const rk_fft_plan* plan, void* out, const void* in,
                         int inverse, void* workspace, void* stream,
                         const void* load_blob, size_t load_bytes,
                         const void* store_blob, size_t store_bytes);

int rk_fft_plan_destroy(rk_fft_plan* plan);

/* Plan introspection (for tests, benches and the fused-kernel generator). */
typedef struct {
    int L;                /* DFT length computed by this pass */
    int C;                /* transforms per work-group tile */
    int threads;          /* work-group size */
    int kind;             /* 0 = Stockham pass, 1 = single-kernel Bluestein, 2 = copy */
    int load_mode;        /* 0 direct, 1 column-staged, 2 flat-staged */
    int store_mode;
    int src;              /* 0 = in, 1 = out, 2 = ws0, 3 = ws1 */
    int dst;
    int axis;             /* array axis this pass belongs to */
    uint32_t pre_ops, post_ops;
    uint64_t n_inst;      /* transforms (instances) in the pass */
    uint64_t grid;        /* work-groups launched */
    uint64_t axis_len;    /* N of the axis (M for Bluestein passes) */
    uint64_t Ns;          /* product of the lengths of earlier passes of this axis */
    int lds_bytes;
    int jit;              /* 1 if the kernel came from the run-time compiler */
    int variant;          /* compile-time kernel variant (pre-op | post-op << 2 | bluestein << 4) */
    int64_t in_dims[3];   /* folded (O, axis length read, I) of the pass input */
    int stream;           /* 1 if the kernel loads and stores with the nontemporal hint */
    int fixed_modes;      /* 1 if the kernel has load_mode/store_mode compiled in */
    int nt_load;          /* 1 if the kernel loads with the nontemporal hint */
    int nt_store;         /* 1 if the kernel stores with the nontemporal hint */
    int wave;             /* 1 if each wave holds whole instances and the exchanges
                             between stages stay inside it (row mapping) */
    int pipe;             /* 1 if the kernel is persistent and loads its next tile
                             while it computes the current one */
    int occ;              /* 1 if the kernel's launch bounds ask for as many resident
                             work-groups per CU as its LDS allows */
} rk_fft_pass_info;

int rk_fft_plan_num_passes(const rk_fft_plan* plan, int* n);
/* Chunked execution: every pass runs over *chunk_outer outer-batch indices
   before the next chunk starts (of *n_outer); 0 when the plan is not chunked. */
int rk_fft_plan_chunking(const rk_fft_plan* plan, int64_t* chunk_outer, int64_t* n_outer);
int rk_fft_plan_pass_info(const rk_fft_plan* plan, int pass, rk_fft_pass_info* info);

/* Replace the kernel of one pass with a run-time compiled one whose
   parameter list is the same pass-argument struct (fused transformations). */
int rk_fft_plan_set_pass_kernel(rk_fft_plan* plan, int pass, void* function);

/* Device-side duration of the last rk_fft_execute of `pass` measured with
   hipEvents when RK_FFT_TIMING was set at plan creation (ms), else -1. */
int rk_fft_plan_set_timing(rk_fft_plan* plan, int enable);
int rk_fft_plan_pass_time_ms(const rk_fft_plan* plan, int pass, float* ms);

/* ------------------------------------------------------------------ */
/* Batched uneven transforms — the whole batch in one kernel launch     */
/* (no reference counterpart: reikna compiles and launches one          */
/* computation per shape, reikna/fft/fft.py:723-756).                   */
/* ------------------------------------------------------------------ */

typedef struct rk_batch_plan rk_batch_plan;

typedef struct {
    int64_t length;      /* 1-D transform length N (any N >= 1) */
    int64_t count;       /* number of transforms of this length */
    int64_t in_offset;   /* element offset of the first transform in `in` */
    int64_t out_offset;  /* element offset of the first transform in `out` */
    int64_t in_dist;     /* elements between consecutive transforms (>= N) */
    int64_t out_dist;
    uint32_t flags;      /* RK_BATCH_SCALE: multiply result by 1/N;
                            RK_BATCH_R2C: real-to-complex item (below) */
} rk_batch_item;

#define RK_BATCH_SCALE 0x1u
/* Real-to-complex item (numpy.fft.rfft; the r2c packing of
   examples/demo_specialized_fft.py:156-331 run inside the batch kernel):
   `length` N is even and >= 4; transform j reads N float32 values starting
   at complex64 element in_offset + j * in_dist of `in` (read as N/2 complex
   values, so in_dist >= N/2) and writes the N/2 + 1 non-negative frequencies
   at out_offset + j * out_dist (out_dist >= N/2 + 1).  Forward only, queue
   mode only. */
#define RK_BATCH_R2C 0x2u

/* Execution modes: one launch of the work-queue kernel for the whole batch
   (default), or one single-shape plan per item (the per-request path). */
#define RK_BATCH_MODE_QUEUE 0u
#define RK_BATCH_MODE_GROUPED 1u

#define RK_BATCH_DRY_RUN 0x1u   /* plan only: no device memory, cannot execute */

typedef struct {
    uint32_t mode;        /* RK_BATCH_MODE_* */
    uint32_t flags;       /* RK_BATCH_DRY_RUN */
    int32_t single_max;   /* longest single-pass length (0: 4096, at most 4096) */
    int32_t factor_max;   /* longest pass of a multi-pass chain (0: 512, at most 1024) */
    int32_t blue_max;     /* reserved, 0 (Bluestein runs as two chains in the batch kernel) */
    int32_t order;        /* queue order: 0 critical path (longest chains first, all chains
                             ending together), 1 level by level in item order,
                             2 level by level, largest tiles first */
} rk_batch_options;

typedef struct {
    int64_t items;        /* transforms in the batch (sum of counts) */
    int64_t passes;       /* passes over all items */
    int64_t tiles;        /* work-groups of the queue launch */
    int64_t workspace_bytes;
    int32_t levels;       /* longest pass chain */
    int32_t launches;     /* kernel launches per execute */
    int32_t mode;
    int32_t lds_bytes;    /* LDS per work-group of the queue kernel */
    int32_t flavour;      /* 0: light kernel, 1: kernel with single-kernel Bluestein cases */
    int32_t _reserved;
} rk_batch_info;

int rk_batch_plan_create(const rk_batch_item* items, int n_items,
                         rk_batch_plan** plan);
int rk_batch_plan_create_ex(const rk_batch_item* items, int n_items,
                            const rk_batch_options* options, rk_batch_plan** plan);
int rk_batch_plan_workspace_size(const rk_batch_plan* plan, size_t* bytes);
/* Asynchronous on `stream`; a plan's queue state is shared by its
   executions, so executions of one plan must not overlap. */
int rk_batch_execute(const rk_batch_plan* plan, void* out, const void* in,
                     int inverse, void* workspace, void* stream);
int rk_batch_plan_destroy(rk_batch_plan* plan);
/* Number of work items / launches the batch plan uses (introspection). */
int rk_batch_plan_stats(const rk_batch_plan* plan, int64_t* work_items,
                        int* launches);
int rk_batch_plan_info(const rk_batch_plan* plan, rk_batch_info* info);
/* Synchronous: 1 in *hit if a dependency wait of an earlier execute gave up
   after its poll limit (the results of that execute are not valid); clears
   the flag and the plan's dependency counters. */
int rk_batch_plan_check(rk_batch_plan* plan, int* hit);

/* ------------------------------------------------------------------ */
/* Request server — continuous batching of FFT requests (no reference   */
/* counterpart: reikna has no serving layer).                           */
/* ------------------------------------------------------------------ */
typedef struct rk_server rk_server;

#define RK_SERVER_BATCHED 0u   /* admitted requests run together in one batch-kernel launch */
#define RK_SERVER_PLAIN 1u     /* each request runs by itself (its single-shape plan) */

typedef struct {
    size_t budget_bytes;   /* HBM for request outputs and workspaces */
    int32_t slots;         /* launches in flight at once (one stream each), 1..16 */
    int32_t max_requests;  /* requests per launch (0: 1024) */
    uint32_t mode;         /* RK_SERVER_* */
    int64_t plain_min_elements;  /* RK_SERVER_BATCHED: requests of at least this many
                                    elements run by themselves with their single-shape
                                    plan (0: never) */
    uint32_t spin_limit;   /* polls before a dependency wait of the batch kernel gives
                              up and flags its launch (0: 2^24; small values are for
                              tests of the failure path) */
    uint32_t fault_inject; /* 0; RK_SERVER_FAULT_* bits make the next such step fail
                              (tests of the failure path) */
} rk_server_config;

#define RK_SERVER_FAULT_SLOT_ALLOC 1u   /* the next growth of a slot's buffer fails */

typedef struct {
    uint64_t id;           /* caller's tag, unique among unreleased requests */
    int32_t ndim;          /* 1 or 2: the rank of each transform */
    int32_t inverse;
    int64_t shape[2];      /* (n) or (n0, n1), row-major, last axis contiguous */
    int64_t count;         /* transforms in the request, back to back */
    const void* in;        /* device input, count * prod(shape) complex64 */
    uint64_t t_arrival_ns; /* rk_server_now_ns() clock; 0: time of submit */
} rk_request;

typedef struct {
    uint64_t id;
    void* out;             /* device output, valid until rk_server_release(id) */
    uint64_t t_arrival_ns, t_admitted_ns, t_done_ns;
    int32_t status;        /* 0; 1 if a dependency wait of its launch gave up (the
                              output is not valid; the server has reset its state);
                              2 if its launch could not be enqueued (not run; the
                              rk_server_progress call that tried returned the error) */
    int32_t _reserved;
} rk_completion;

typedef struct {
    int64_t pending, in_flight_requests, in_flight_launches, completed, launches, templates;
    size_t budget_used, budget_bytes;
    int64_t failed_launches;   /* launches whose completions carry status 1 or 2 */
} rk_server_stats_t;

int rk_server_create(const rk_server_config* config, rk_server** server);
/* Queue requests (they wait for memory and a slot); RK_ERR_BUDGET for a
   request that can never fit in the budget, RK_ERR_INVALID for an id that is
   still in use (submitted and not released) or given twice.  All requests are
   checked first: on an error none of them is queued. */
int rk_server_submit(rk_server* server, const rk_request* requests, int n);
/* Reap finished launches, admit and launch queued requests, and return up to
   max_done completions (never blocks). */
int rk_server_progress(rk_server* server, rk_completion* done, int max_done, int* n_done);
int rk_server_release(rk_server* server, uint64_t id);
int rk_server_stats(const rk_server* server, rk_server_stats_t* stats);
uint64_t rk_server_now_ns(void);
/* Open-loop replay: submit each request when rk_server_now_ns() reaches its
   t_arrival_ns, progress the server meanwhile, and write the n completions to
   done (in completion order); outputs are released as they complete unless
   keep_outputs.  Fails if nothing completes for stall_ms (0: 10 s), and
   refuses to start while the server has other requests queued or in flight. */
int rk_server_replay(rk_server* server, const rk_request* requests, int n,
                     rk_completion* done, int keep_outputs, int stall_ms);
int rk_server_destroy(rk_server* server);

/* ------------------------------------------------------------------ */
/* HIP graphs: capture a stream's work once, replay it with one call.   */
/* ------------------------------------------------------------------ */
typedef struct rk_graph rk_graph;
int rk_graph_begin(void* stream);
int rk_graph_end(void* stream, rk_graph** graph);
int rk_graph_launch(rk_graph* graph, void* stream);
int rk_graph_destroy(rk_graph* graph);

#ifdef __cplusplus
}
#endif

#endif /* REIKNA_AMD_H */
