// fft_args.h — the argument block of one FFT pass launch.
//
// Shared verbatim by the host planner (hipcc) and by device code compiled
// either ahead of time or at run time with hiprtc, so it uses only builtin
// types.  One pass computes n_inst independent DFTs of length L ("instances")
// drawn from an axis of length N:
//
//   instance index  inst = ((o * H + h) * NS + l) * I + i      (i fastest)
//   input element p (0 <= p < L) sits at axis position
//        m    = (h * NS + l) + p * (N / L)
//   output element k sits at axis position
//        kabs = h * NS * L + l + k * NS
//
// which is one radix-L step of a Stockham autosort FFT over the whole axis
// (NS = product of the lengths of the earlier passes of this axis).  A
// single-pass transform is the case NS = H = 1, L = N.  Inputs of a pass with
// NS > 1 are multiplied by the inter-pass twiddle W_{NS*L}^{p*l}.
#ifndef RK_FFT_ARGS_H
#define RK_FFT_ARGS_H

typedef unsigned long long rk_u64;
typedef long long rk_i64;
typedef unsigned int rk_u32;
typedef int rk_i32;

// Compute mapping of threads to (instance c, thread-in-instance t):
//   row mapping    (cmap = 0): c = tid / T, t = tid % T  (lanes along an instance)
//   column mapping (cmap = 1): c = tid % C, t = tid / C  (lanes across instances)
// load/store modes: how global memory is walked
#define RK_MODE_DIRECT 0u   // each thread touches its own elements in the compute mapping
#define RK_MODE_COLUMN 1u   // staged through LDS, lanes across adjacent instances
#define RK_MODE_FLAT 2u     // staged through LDS, the tile is one contiguous block of C*L elements
#define RK_MODE_ROW 3u      // staged through LDS, lanes along each instance's elements

// pre-ops, applied to each loaded element (in this order)
#define RK_PRE_CHIRP 0x1u     // Bluestein: x[m] * chirp[m], zero for m >= n_valid_in
#define RK_PRE_TWIDDLE 0x2u   // inter-pass twiddle W_N^{p*l*tw_mult} (FP64 tables)

// post-ops, applied to each stored element (in this order)
#define RK_POST_KMUL_CONJ 0x1u  // Bluestein: conj(y * bspec[kabs])
#define RK_POST_CONJ 0x2u       // conj(y)
#define RK_POST_CHIRP 0x4u      // y * chirp[kabs], nothing stored for kabs >= n_valid_out

struct rk_pass_args {
    const float2* in;
    float2* out;
    const float2* twl;       // W_L^e, e in [0, L): in-pass stage twiddles
    const double2* tw_hi;    // W_N^{(e >> tw_shift) << tw_shift}
    const double2* tw_lo;    // W_N^{e & tw_mask}
    const float2* chirp;     // exp(-i pi (n^2 mod 2Nr) / Nr), n < Nr
    const float2* bspec;     // FFT_N(conj chirp, wrapped) / N
    rk_u64 n_inst;
    rk_u64 n_tiles;
    rk_u64 tile_base;        // first tile of this launch (grids are split on the host)
    rk_u32 xcd_group;        // >1: consecutive tiles in groups of this size run on one XCD
    rk_u32 n_xcd;            // XCDs the dispatcher round-robins work-groups over
    rk_u64 dim_I, dim_NS, dim_H;
    rk_i64 in_off, in_so, in_si, in_sa;
    rk_i64 out_off, out_so, out_si, out_sa;
    rk_u64 N;                // axis length this pass works on
    rk_u64 pos_mult;         // N / L
    rk_u64 tw_mult;          // N / (NS * L)
    rk_u64 tw_mask;
    rk_u32 tw_shift;
    rk_u32 load_mode, store_mode;
    rk_u32 pre_ops, post_ops;
    rk_u32 cmap;             // 0 row mapping, 1 column mapping
    rk_u64 n_valid_in;       // positions m >= n_valid_in read as zero
    rk_u64 n_valid_out;      // positions kabs >= n_valid_out are not stored
    float in_sign;           // -1 conjugates the input (inverse transform)
    float out_sign;          // -1 conjugates the output
    float out_scale;         // 1/N on the last pass of an inverse transform
    rk_u32 div_magic;        // 1: the mag_/sh_ divisors below are set (n_inst < 2^31)
    // x / dim = (mulhi(x, mag) + x) >> sh for every x < 2^31 (round-up
    // multipliers computed on the host, sh = ceil(log2 dim))
    rk_u32 mag_I, mag_NS, mag_H;
    rk_u32 sh_I, sh_NS, sh_H;
    double out_scale64;      // out_scale in FP64 (complex128 passes)
};

// Arguments of fused (transformation-carrying) pass kernels: the leaf
// pointers and scalars of the attached transformations, packed by the host in
// the order and alignment of the generated struct that reinterprets it.
#define RK_USER_BLOB_BYTES 512
struct rk_user_blob {
    alignas(16) unsigned char b[RK_USER_BLOB_BYTES];
};

#endif
