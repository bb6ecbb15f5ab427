// fft_batch.h — descriptors of the persistent batch kernel (rk_fft_batch).
//
// A batch plan turns every transform of a batch into a chain of passes
// (rk_pass_args, see fft_args.h) and every pass into tiles.  All tiles of all
// passes form one queue, ordered level by level (every first pass, then every
// second pass, ...).  Work-group b of the launch runs queue position b; a
// tile of a pass that reads the output of an earlier pass waits until all
// tiles of that pass are counted done.  Every pass a tile can wait for sits
// earlier in the queue, and work-groups are dispatched in index order, so the
// awaited tiles belong to work-groups that are running or finished and every
// wait ends (a poll limit turns a wait that does not end into an error flag
// instead of a hang).
//
// Counting: the tiles of a counted pass are split into `groups` groups (tile
// t in group t % groups), each with a tile counter on a 128-byte line of its
// own, plus one line that counts finished groups; a waiter polls that line
// alone.  Same-address device atomics serialise (~11 ns each), so a pass of n
// tiles finishes its counting after about n / groups + groups atomics instead
// of n.  Each pass is awaited by at most one later pass (the next pass of its
// transform).  The tile that completes the last group of a pass clears the
// counters of the pass it waited for (all of whose readers have finished) and,
// when nothing waits for its own pass, its own counters: every counter is
// zero again when the launch ends.
//
// Shared by the host planner and device code, so it uses only builtin types.
#ifndef RK_FFT_BATCH_H
#define RK_FFT_BATCH_H

#include "fft_args.h"

#define RK_BATCH_THREADS 256

#define RK_BATCH_BUF_IN 0u
#define RK_BATCH_BUF_OUT 1u
#define RK_BATCH_BUF_WS 2u
#define RK_BATCH_BUF_ABS 3u   // the pointer already in the pass arguments

// Tile case of the real-to-complex combine pass (not an FFT tile): tile t
// of the pass computes outputs k = (t % a.dim_I) * RK_BATCH_R2C_SPAN + ... of
// transform t / a.dim_I, X[k] = Z[k] A[k] + conj(Z[N/2 - k]) B[k] from the
// half-length spectrum Z (a.in, a.N = N/2 points per transform, a.in_so
// apart) with A = a.chirp, B = a.bspec, and X[N/2] = Re Z[0] - Im Z[0].
#define RK_BATCH_CASE_R2C 0xffffu
#define RK_BATCH_R2C_SPAN (4 * RK_BATCH_THREADS)

#define RK_BATCH_LINE 32u         // u32 per counter line (128 bytes)
#define RK_BATCH_GROUPS_MAX 16u   // tile groups of one pass

struct rk_batch_pass {
    rk_pass_args a;          // in/out are taken from the launch (src/dst below)
    rk_u64 first;            // queue position of tile 0 of this pass
    rk_u32 kcase;            // tile case of the persistent kernel
    rk_u32 src, dst;         // RK_BATCH_BUF_*
    rk_i32 wait;             // pass whose tiles must all be done first, -1: none
    rk_u32 wait_line;        // first counter line of that pass
    rk_u32 wait_groups;      // its tile groups (done when all are counted)
    rk_u32 line;             // first counter line of this pass: groups done, then one per group
    rk_u32 groups;           // tile groups of this pass (0: its tiles are not counted)
    rk_u32 awaited;          // 1: a later pass waits for this one
    rk_u32 item_first;       // first pass of its transform (inverse: conjugate input)
    rk_u32 item_last;        // last pass of its transform (inverse: conjugate output)
    float fwd_scale;         // output scale of the forward transform
    float inv_scale;         // output scale of the inverse transform
    rk_u32 inverse;          // 1: this pass belongs to an inverse transform
};

struct rk_batch_work {
    const rk_batch_pass* passes;
    const rk_u32* tile_pass;  // pass of each queue position
    rk_u32* counters;         // RK_BATCH_LINE u32 per line (zero between launches)
    rk_u32* state;            // [2]: set when a wait gave up after spin_limit polls
    const float2* in;
    float2* out;
    float2* ws;
    rk_u32 n_passes;
    rk_u32 total;             // queue positions: tiles of all passes
    rk_u32 spin_limit;
    rk_u32 inverse;
    rk_u32 poll_sleep_max;    // longest sleep between polls, in units of 2 x 64 clocks
    rk_u32 _pad;
};

#endif
