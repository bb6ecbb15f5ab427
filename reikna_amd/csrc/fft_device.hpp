// fft_device.hpp — device building blocks of the gfx950 FFT.
//
// One kernel template, fft_pass_body<L, C, V, IO>, computes a tile of C
// independent DFTs of compile-time length L inside one work-group (V is the
// compile-time variant: pre-op, post-op, Bluestein; see Variant below):
//
//   * each DFT ("instance") is worked on by T = L / R0 threads, R0 being the
//     largest radix of the compile-time radix plan of L (16, 8, 4, 2, 7, 5, 3);
//   * the stages are Stockham autosort steps: a thread reads R values at
//     j + r*L/R from LDS, multiplies by the stage twiddles W_{Ns*R}^{r*(j%Ns)}
//     (a table built in FP64 and rounded once), does an in-register DFT_R and
//     writes them to (j/Ns)*Ns*R + j%Ns + r*Ns, so the result comes out in
//     natural order with no bit reversal;
//   * global memory is walked in one of three ways per side (rk_pass_args
//     load_mode/store_mode): DIRECT (each thread its own elements: coalesced
//     when the axis is contiguous), COLUMN (lanes across adjacent instances,
//     staged through LDS: coalesced when instances are adjacent, i.e. strided
//     axes and the inner passes of a multi-pass transform), FLAT (the tile is
//     one contiguous block, streamed linearly through LDS);
//   * the Bluestein variant runs the whole chirp-z convolution in one tile:
//     chirp on load, forward DFT_L, multiply by the kernel spectrum and
//     conjugate, forward DFT_L again, conjugate, chirp and truncate on store.
//
// The inverse transform is the forward one on conjugated data
// (IFFT(x) = conj(FFT(conj x)) / N), so the direction is the run-time
// in_sign / out_sign / out_scale of the pass arguments, not a kernel variant.
//
// The IO policy supplies the element load/store so that fused input/output
// transformations (generated at run time) reuse the same kernel body.
#ifndef RK_FFT_DEVICE_HPP
#define RK_FFT_DEVICE_HPP

#if !defined(__HIPCC_RTC__)
#include <hip/hip_runtime.h>
#endif

#include "fft_args.h"

#define RK_HD __host__ __device__ __forceinline__

namespace rkfft {

// ------------------------------------------------------------------------
// complex arithmetic (x = re, y = im) on complex64 (float2) and complex128
// (double2) values; a pass's element type is its IO policy's value_type
// ------------------------------------------------------------------------
template <class X> struct RealOf;
template <> struct RealOf<float2> { typedef float type; };
template <> struct RealOf<double2> { typedef double type; };
template <class X> using real_t = typename RealOf<X>::type;

template <class X> RK_HD X cx(real_t<X> re, real_t<X> im);
template <> RK_HD float2 cx<float2>(float re, float im) { return make_float2(re, im); }
template <> RK_HD double2 cx<double2>(double re, double im) { return make_double2(re, im); }

template <class X> RK_HD X c_add(X a, X b) { return cx<X>(a.x + b.x, a.y + b.y); }
template <class X> RK_HD X c_sub(X a, X b) { return cx<X>(a.x - b.x, a.y - b.y); }
template <class X> RK_HD X c_mul(X a, X b) { return cx<X>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
template <class X> RK_HD X c_conj(X a) { return cx<X>(a.x, -a.y); }
template <class X> RK_HD X c_mul_mi(X a) { return cx<X>(a.y, -a.x); }   // a * (-i)
template <class X> RK_HD X c_scale(X a, real_t<X> s) { return cx<X>(a.x * s, a.y * s); }
// cond ? a : b part by part (a select of the whole 16-byte complex128 value
// is lowered through scratch memory)
template <class X> RK_HD X c_select(bool cond, X a, X b) { return cx<X>(cond ? a.x : b.x, cond ? a.y : b.y); }

// Butterfly constants: double literals, rounded once to the element's real
// type where they are used ((F)RK_C8 and so on).
#define RK_C8 0.7071067811865476      // cos(pi/4)
#define RK_C16 0.9238795325112867     // cos(pi/8)
#define RK_S16 0.3826834323650898     // sin(pi/8)
#define RK_S3 0.8660254037844387      // sin(2pi/3)
#define RK_C5_1 0.30901699437494745   // cos(2pi/5)
#define RK_C5_2 -0.8090169943749473   // cos(4pi/5)
#define RK_S5_1 0.9510565162951535    // sin(2pi/5)
#define RK_S5_2 0.5877852522924732    // sin(4pi/5)
#define RK_C7_1 0.6234898018587336
#define RK_C7_2 -0.22252093395631434
#define RK_C7_3 -0.900968867902419
#define RK_S7_1 0.7818314824680298
#define RK_S7_2 0.9749279121818236
#define RK_S7_3 0.43388373911755823

// ------------------------------------------------------------------------
// in-register forward DFTs, y_k = sum_n v_n exp(-2 pi i n k / R), in place
// ------------------------------------------------------------------------
template <class X> RK_HD void dft2(X& a, X& b) {
    X t = a;
    a = c_add(t, b);
    b = c_sub(t, b);
}

template <class X> RK_HD void dft4(X& a0, X& a1, X& a2, X& a3) {
    X t0 = c_add(a0, a2), t1 = c_sub(a0, a2);
    X t2 = c_add(a1, a3), t3 = c_mul_mi(c_sub(a1, a3));
    a0 = c_add(t0, t2);
    a2 = c_sub(t0, t2);
    a1 = c_add(t1, t3);
    a3 = c_sub(t1, t3);
}

// multiply by W8^1 = (1 - i)/sqrt2 and W8^3 = (-1 - i)/sqrt2
template <class X> RK_HD X c_w8_1(X a) {
    using F = real_t<X>;
    return cx<X>((F)RK_C8 * (a.x + a.y), (F)RK_C8 * (a.y - a.x));
}
template <class X> RK_HD X c_w8_3(X a) {
    using F = real_t<X>;
    return cx<X>((F)RK_C8 * (a.y - a.x), -(F)RK_C8 * (a.x + a.y));
}

template <class X> RK_HD void dft8(X* v) {
    X e0 = v[0], e1 = v[2], e2 = v[4], e3 = v[6];
    X o0 = v[1], o1 = v[3], o2 = v[5], o3 = v[7];
    dft4(e0, e1, e2, e3);
    dft4(o0, o1, o2, o3);
    o1 = c_w8_1(o1);
    o2 = c_mul_mi(o2);
    o3 = c_w8_3(o3);
    v[0] = c_add(e0, o0); v[4] = c_sub(e0, o0);
    v[1] = c_add(e1, o1); v[5] = c_sub(e1, o1);
    v[2] = c_add(e2, o2); v[6] = c_sub(e2, o2);
    v[3] = c_add(e3, o3); v[7] = c_sub(e3, o3);
}

// 16 = 4 x 4: DFT4 over n2 for each n1 (x[n1 + 4 n2]), twiddle W16^{n1 k2},
// DFT4 over n1 for each k2, output y[k2 + 4 k1].
template <class X> RK_HD void dft16(X* v) {
    using F = real_t<X>;
    dft4(v[0], v[4], v[8], v[12]);
    dft4(v[1], v[5], v[9], v[13]);
    dft4(v[2], v[6], v[10], v[14]);
    dft4(v[3], v[7], v[11], v[15]);
    const X w1 = cx<X>((F)RK_C16, -(F)RK_S16);
    const X w3 = cx<X>((F)RK_S16, -(F)RK_C16);
    const X w9 = cx<X>(-(F)RK_C16, (F)RK_S16);
    v[5] = c_mul(v[5], w1);
    v[9] = c_w8_1(v[9]);
    v[13] = c_mul(v[13], w3);
    v[6] = c_w8_1(v[6]);
    v[10] = c_mul_mi(v[10]);
    v[14] = c_w8_3(v[14]);
    v[7] = c_mul(v[7], w3);
    v[11] = c_w8_3(v[11]);
    v[15] = c_mul(v[15], w9);
    dft4(v[0], v[1], v[2], v[3]);
    dft4(v[4], v[5], v[6], v[7]);
    dft4(v[8], v[9], v[10], v[11]);
    dft4(v[12], v[13], v[14], v[15]);
    // v[4 k2 + k1] holds y[k2 + 4 k1]: transpose the 4x4 block
    X t;
    t = v[1]; v[1] = v[4]; v[4] = t;
    t = v[2]; v[2] = v[8]; v[8] = t;
    t = v[3]; v[3] = v[12]; v[12] = t;
    t = v[6]; v[6] = v[9]; v[9] = t;
    t = v[7]; v[7] = v[13]; v[13] = t;
    t = v[11]; v[11] = v[14]; v[14] = t;
}

template <class X> RK_HD void dft3(X* v) {
    using F = real_t<X>;
    X t = c_add(v[1], v[2]);
    X m = cx<X>(v[0].x - (F)0.5 * t.x, v[0].y - (F)0.5 * t.y);
    X d = c_sub(v[1], v[2]);
    X s = cx<X>((F)RK_S3 * d.y, -(F)RK_S3 * d.x);   // -i sin(2pi/3) d
    v[0] = c_add(v[0], t);
    v[1] = c_add(m, s);
    v[2] = c_sub(m, s);
}

template <class X> RK_HD void dft5(X* v) {
    using F = real_t<X>;
    const F c1 = (F)RK_C5_1, c2 = (F)RK_C5_2, s1 = (F)RK_S5_1, s2 = (F)RK_S5_2;
    X t1 = c_add(v[1], v[4]), t2 = c_add(v[2], v[3]);
    X t3 = c_sub(v[1], v[4]), t4 = c_sub(v[2], v[3]);
    X a0 = v[0];
    X b1 = cx<X>(a0.x + c1 * t1.x + c2 * t2.x, a0.y + c1 * t1.y + c2 * t2.y);
    X b2 = cx<X>(a0.x + c2 * t1.x + c1 * t2.x, a0.y + c2 * t1.y + c1 * t2.y);
    // u1 = s1 t3 + s2 t4, u2 = s2 t3 - s1 t4; y1 = b1 - i u1, y2 = b2 - i u2
    X u1 = cx<X>(s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y);
    X u2 = cx<X>(s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y);
    v[0] = cx<X>(a0.x + t1.x + t2.x, a0.y + t1.y + t2.y);
    v[1] = c_add(b1, c_mul_mi(u1));
    v[4] = c_sub(b1, c_mul_mi(u1));
    v[2] = c_add(b2, c_mul_mi(u2));
    v[3] = c_sub(b2, c_mul_mi(u2));
}

// Odd prime radix 7: y_k = A_k - i B_k, y_{7-k} = A_k + i B_k with
// A_k = a0 + sum_n cos(2 pi n k/7)(a_n + a_{7-n}), B_k = sum_n sin(2 pi n k/7)(a_n - a_{7-n}).
template <class X> RK_HD void dft7(X* v) {
    using F = real_t<X>;
    const F c[4] = {(F)1, (F)RK_C7_1, (F)RK_C7_2, (F)RK_C7_3};
    const F s[4] = {(F)0, (F)RK_S7_1, (F)RK_S7_2, (F)RK_S7_3};
    X p[4], m[4];
    for (int n = 1; n <= 3; ++n) {
        p[n] = c_add(v[n], v[7 - n]);
        m[n] = c_sub(v[n], v[7 - n]);
    }
    X a0 = v[0];
    v[0] = cx<X>(a0.x + p[1].x + p[2].x + p[3].x, a0.y + p[1].y + p[2].y + p[3].y);
    for (int k = 1; k <= 3; ++k) {
        X A = a0, B = cx<X>((F)0, (F)0);
        for (int n = 1; n <= 3; ++n) {
            int nk = (n * k) % 7;
            F cc = nk <= 3 ? c[nk] : c[7 - nk];
            F ss = nk <= 3 ? s[nk] : -s[7 - nk];
            A.x += cc * p[n].x; A.y += cc * p[n].y;
            B.x += ss * m[n].x; B.y += ss * m[n].y;
        }
        v[k] = c_add(A, c_mul_mi(B));
        v[7 - k] = c_sub(A, c_mul_mi(B));
    }
}

template <int R, class X> RK_HD void dft(X* v) {
    if constexpr (R == 2) dft2(v[0], v[1]);
    else if constexpr (R == 3) dft3(v);
    else if constexpr (R == 4) dft4(v[0], v[1], v[2], v[3]);
    else if constexpr (R == 5) dft5(v);
    else if constexpr (R == 7) dft7(v);
    else if constexpr (R == 8) dft8(v);
    else if constexpr (R == 16) dft16(v);
    else static_assert(R == 1, "radix without a butterfly");
}

// ------------------------------------------------------------------------
// compile-time radix plan of a length L = 2^a 3^b 5^c 7^d
// ------------------------------------------------------------------------
struct RadixList {
    int n;
    int r[24];
};

__host__ __device__ constexpr RadixList make_radix_list(int L) {
    RadixList s = {0, {0}};
    int x = L, p2 = 0, n3 = 0, n5 = 0, n7 = 0;
    while (x % 2 == 0) { x /= 2; ++p2; }
    while (x % 3 == 0) { x /= 3; ++n3; }
    while (x % 5 == 0) { x /= 5; ++n5; }
    while (x % 7 == 0) { x /= 7; ++n7; }
    // largest radix first: it fixes T = L / r[0] threads per instance
    while (p2 >= 4) { s.r[s.n++] = 16; p2 -= 4; }
    if (p2 > 0) s.r[s.n++] = 1 << p2;
    for (int k = 0; k < n7; ++k) s.r[s.n++] = 7;
    for (int k = 0; k < n5; ++k) s.r[s.n++] = 5;
    for (int k = 0; k < n3; ++k) s.r[s.n++] = 3;
    if (s.n == 0) s.r[s.n++] = 1;
    // move the largest radix to the front (e.g. [4, 5, 5] -> [5, 4, 5])
    int imax = 0;
    for (int k = 1; k < s.n; ++k)
        if (s.r[k] > s.r[imax]) imax = k;
    int t = s.r[0]; s.r[0] = s.r[imax]; s.r[imax] = t;
    return s;
}

__host__ __device__ constexpr bool is_supported_length(int L) {
    int x = L;
    if (x < 1) return false;
    while (x % 2 == 0) x /= 2;
    while (x % 3 == 0) x /= 3;
    while (x % 5 == 0) x /= 5;
    while (x % 7 == 0) x /= 7;
    return x == 1;
}

template <int L> struct Plan {
    static constexpr RadixList RL = make_radix_list(L);
    static constexpr int NSTAGE = RL.n;
    static constexpr int R0 = RL.r[0];
    static constexpr int T = L / R0;
    static constexpr int radix(int s) { return RL.r[s]; }
    static constexpr int ns(int s) {
        int p = 1;
        for (int k = 0; k < s; ++k) p *= RL.r[k];
        return p;
    }
    static constexpr int nb(int s) { return L / RL.r[s]; }
    static constexpr int q(int s) { return (L / RL.r[s] + T - 1) / T; }
    static constexpr int vmax() {
        int m = 0;
        for (int s = 0; s < RL.n; ++s)
            if (q(s) * RL.r[s] > m) m = q(s) * RL.r[s];
        return m;
    }
};

// LDS layout of one instance: one padding slot every 16 complex values.
RK_HD constexpr int lds_pad(int p) { return p + (p >> 4); }
// lds_pad(b + c) for a run-time b >= 0 with b % 16 < BM and a c >= 0 that is a
// constant once the caller's loops are unrolled.  When b % 16 + c % 16 < 16
// no padding slot falls between b and b + c, and the padded position is
// lds_pad(b) + lds_pad(c): one address register per exchange, the constant in
// the LDS instruction's offset field, where the plain form costs a register
// and three instructions per value (the compiler does not split a shift of a
// sum whose low bits it cannot see).
template <int BM> RK_HD constexpr int lds_pad_add(int b, int c) {
    return (BM - 1) + (c & 15) < 16 ? lds_pad(b) + lds_pad(c) : lds_pad(b + c);
}
// Base row pitch of the instances of a tile: odd, so that lanes walking
// adjacent instances hit distinct banks.
RK_HD constexpr int lds_pitch_base(int L) {
    int s = lds_pad(L - 1) + 1;
    return (s % 2 == 0) ? s + 1 : s;
}
// A tile's exchanges go through halves (real parts, then imaginary parts)
// when it needs more than 80 KB at the base pitch; ES is the element size
// (8 bytes complex64, 16 complex128).
RK_HD constexpr bool lds_split(int L, int C, int ES = 8) { return C * lds_pitch_base(L) * ES > 80 * 1024; }
// Row pitch of a tile of C instances of L points (T threads each).  Under the
// column mapping a wave covers C instances x 64/C threads.  For C a power of
// two from 2 to 16 (32 when split) and T >= 16, an odd pitch stacks the C
// instances' accesses on the same banks; a pitch of m/C modulo m (m = 32
// slots of 64-bit accesses, or 64 banks of split 32-bit accesses) gives each
// instance its own share of the banks.  (A bank model of every exchange of
// every catalogue tile: 1.2-3x fewer conflict cycles where this applies,
// averaged over load/store modes and both mappings; the other tiles keep the
// odd pitch, which the model favours for them.)  Complex128 tiles keep the
// odd pitch (the rule is for 64-bit accesses and their 32-bit halves).
RK_HD constexpr int lds_pitch_for(int L, int C, int T, bool force_split = false, int ES = 8) {
    if (ES != 8) return lds_pitch_base(L);
    const bool split = force_split || lds_split(L, C);
    const bool pow2 = C > 0 && (C & (C - 1)) == 0;
    if (!pow2 || C < 2 || T < 16 || C > (split ? 32 : 16)) return lds_pitch_base(L);
    const int m = split ? 64 : 32;
    const int target = m / C;
    int p = lds_pad(L - 1) + 1;
    while (p % m != target) ++p;
    return p;
}

// ------------------------------------------------------------------------
// per-instance coordinates
// ------------------------------------------------------------------------
struct Inst {
    rk_u64 o, i, l;
    rk_u64 jpass;    // h*NS + l: axis position of input element 0
    rk_u64 kbase;    // h*NS*L + l: axis position of output element 0
    rk_i64 in_base, out_base;
    bool valid;
};

// x / d for x < 2^31 with the host's round-up multiplier m and shift s.
__device__ __forceinline__ rk_u32 magic_div(rk_u32 x, rk_u32 m, rk_u32 s) { return (__umulhi(x, m) + x) >> s; }

template <int L>
__device__ __forceinline__ Inst decompose(const rk_pass_args& a, rk_u64 inst) {
    Inst r;
    r.valid = inst < a.n_inst;
    rk_u64 x = r.valid ? inst : 0;
    rk_u64 i = 0, l = 0, h = 0;
    if (a.div_magic) {   // a multiply and a shift per division instead of a division sequence
        rk_u32 y = (rk_u32)x, q;
        if (a.dim_I > 1) { q = magic_div(y, a.mag_I, a.sh_I); i = y - q * (rk_u32)a.dim_I; y = q; }
        if (a.dim_NS > 1) { q = magic_div(y, a.mag_NS, a.sh_NS); l = y - q * (rk_u32)a.dim_NS; y = q; }
        if (a.dim_H > 1) { q = magic_div(y, a.mag_H, a.sh_H); h = y - q * (rk_u32)a.dim_H; y = q; }
        x = y;
    } else {
        if (a.dim_I > 1) { i = x % a.dim_I; x /= a.dim_I; }
        if (a.dim_NS > 1) { l = x % a.dim_NS; x /= a.dim_NS; }
        if (a.dim_H > 1) { h = x % a.dim_H; x /= a.dim_H; }
    }
    r.o = x;
    r.i = i;
    r.l = l;
    r.jpass = h * a.dim_NS + l;
    r.kbase = h * a.dim_NS * (rk_u64)L + l;
    r.in_base = a.in_off + (rk_i64)x * a.in_so + (rk_i64)i * a.in_si;
    r.out_base = a.out_off + (rk_i64)x * a.out_so + (rk_i64)i * a.out_si;
    return r;
}

// W_N^e for 0 <= e < N from two FP64 tables; the product is formed in FP64
// and rounded once to the element's real type.
__device__ __forceinline__ double2 d_mul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ __forceinline__ double2 twiddle_n64(const rk_pass_args& a, rk_u64 e) {
    return d_mul(a.tw_hi[e >> a.tw_shift], a.tw_lo[e & a.tw_mask]);
}

// The inter-pass twiddles of one thread, W_N^{p*l*tw_mult} for p = p0, p0+dp,
// p0+2dp, ...: two table lookups for the first one and the step, then an FP64
// recurrence; every value is rounded once to the element's real type.
struct TwiddleWalk {
    double2 w, s;
    __device__ __forceinline__ void init(const rk_pass_args& a, rk_u64 l, rk_u64 p0, rk_u64 dp) {
        const rk_u64 lm = l * a.tw_mult;
        w = twiddle_n64(a, p0 * lm);
        s = twiddle_n64(a, dp * lm);
    }
    template <class X> __device__ __forceinline__ X next() {
        const X r = cx<X>((real_t<X>)w.x, (real_t<X>)w.y);
        w = d_mul(w, s);
        return r;
    }
};

// The inter-pass twiddles of one thread's values x[r] at positions p0 + r*dp
// of instance l.  (An exact FP32 table of W_{NS*L}^j read at index l*p was
// measured instead: on 64 x 2^20 the 128-point middle pass went from 181 to
// 226 us, its lanes' table reads scattered over many cache lines per load,
// while the FP64 walk costs VALU only.)
template <int NV, class X>
__device__ __forceinline__ void apply_pass_twiddles(const rk_pass_args& a, rk_u64 l, int p0, int dp, X* x) {
    TwiddleWalk tw;
    tw.init(a, l, (rk_u64)p0, (rk_u64)dp);
#pragma unroll
    for (int r = 0; r < NV; ++r) x[r] = c_mul(x[r], tw.template next<X>());
}

// The table pointers of the pass arguments, as the element type X of a pass
// (complex64 or complex128 tables; the host builds the ones its plan needs).
template <class X> __device__ __forceinline__ const X* table(const float2* p) { return reinterpret_cast<const X*>(p); }

// The output scale (1/N on the last pass of an inverse transform) in the
// element's real type: complex128 passes take the FP64 value.
template <class X> __device__ __forceinline__ real_t<X> out_scale_of(const rk_pass_args& a) {
    if constexpr (sizeof(real_t<X>) == 8) return a.out_scale64;
    else return a.out_scale;
}

// Plain strided element access: the identity IO policy.
// An IO policy is an object with load(a, o, m, i, addr) and
// store(a, o, k, i, addr, v): o / i are the flat indices over the dims before /
// after the transformed axis, m / k the axis position, addr the plain strided
// element address.  Fused transformations (generated at run time) use o, m, i.
// kBranchFree: a lane with nothing to load loads a valid element instead (its
// instance's first position, or the tile's first instance: valid indices o,
// m, i and a valid address) and discards it, so every load of a thread is
// issued without a branch around it (a branch per load serialises them on
// their completion).
// Streaming variant (NT = true): every load and store carries the
// nontemporal hint, so a pass whose input and output far exceed the 256 MB
// Infinity Cache does not fill it and the L2 with lines nothing reads again
// (measured on 64 x 2^20: 0.640 -> 0.599 ms; a hand-written copy of 512 MiB
// runs at 5.8 TB/s plain and 6.0 TB/s with the hint).  Passes whose data fit
// in that cache keep the plain accesses: the hint costs the 256^3 transform,
// whose in-place passes re-read what the previous pass just wrote, 22 %.
// The hint is chosen per side (NTL loads, NTS stores): a chunked chain whose
// intermediate stays in that cache between passes streams only its first
// pass's loads and its last pass's stores.  X is the element type
// (complex64 float2, complex128 double2); every IO policy names it value_type.
template <class F> using rk_v2 = F __attribute__((ext_vector_type(2)));
template <bool NTL, bool NTS = NTL, class X = float2> struct GlobalIO {
    using value_type = X;
    static constexpr bool kBranchFree = true;
    __device__ __forceinline__ X load(const rk_pass_args& a, rk_u64 /*o*/, rk_u64 /*m*/, rk_u64 /*i*/,
                                      rk_i64 addr) const {
        const X* p = reinterpret_cast<const X*>(a.in) + addr;
        if constexpr (NTL) {
            const rk_v2<real_t<X>> x = __builtin_nontemporal_load(reinterpret_cast<const rk_v2<real_t<X>>*>(p));
            return cx<X>(x.x, x.y);
        } else {
            return *p;
        }
    }
    __device__ __forceinline__ void store(const rk_pass_args& a, rk_u64 /*o*/, rk_u64 /*k*/, rk_u64 /*i*/,
                                          rk_i64 addr, X v) const {
        X* p = reinterpret_cast<X*>(a.out) + addr;
        if constexpr (NTS) {
            rk_v2<real_t<X>> x;
            x.x = v.x;
            x.y = v.y;
            __builtin_nontemporal_store(x, reinterpret_cast<rk_v2<real_t<X>>*>(p));
        } else {
            *p = v;
        }
    }
};
using PlainIO = GlobalIO<false>;
using StreamIO = GlobalIO<true>;

// Plain strided IO whose loads and stores are coherent across the whole
// device (relaxed agent-scope atomics: they bypass the caches that are not
// shared by all CUs).  Used where one kernel launch hands data from one
// work-group to another: after the producer's stores have completed, a
// counter in memory is enough to publish them, with no cache write-back or
// invalidation in between.  (Choosing cached or coherent accesses per pass
// with a run-time branch was measured: the second code path raised the
// batch kernel from 160 to 200 VGPRs and cost more than it saved.)
struct CoherentIO {
    using value_type = float2;
    static constexpr bool kBranchFree = true;
    __device__ __forceinline__ float2 load(const rk_pass_args& a, rk_u64 /*o*/, rk_u64 /*m*/, rk_u64 /*i*/,
                                           rk_i64 addr) const {
        const rk_u64 bits = __hip_atomic_load(reinterpret_cast<rk_u64*>(const_cast<float2*>(a.in + addr)),
                                              __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return make_float2(__uint_as_float((unsigned)bits), __uint_as_float((unsigned)(bits >> 32)));
    }
    __device__ __forceinline__ void store(const rk_pass_args& a, rk_u64 /*o*/, rk_u64 /*k*/, rk_u64 /*i*/,
                                          rk_i64 addr, float2 v) const {
        const rk_u64 bits = (rk_u64)__float_as_uint(v.x) | ((rk_u64)__float_as_uint(v.y) << 32);
        __hip_atomic_store(reinterpret_cast<rk_u64*>(a.out + addr), bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
};

// Compile-time pass variant: which pre-/post-ops the pass applies and whether
// it is the single-kernel Bluestein convolution.  Keeping them out of the
// run-time arguments keeps plain passes free of table loads (and registers).
#define RK_VAR_PRE_NONE 0
#define RK_VAR_PRE_TWIDDLE 1
#define RK_VAR_PRE_CHIRP 2
#define RK_VAR_POST_NONE 0
#define RK_VAR_POST_KMUL_CONJ 1
#define RK_VAR_POST_CONJ_CHIRP 2
#define RK_VARIANT(pre, post, blue) ((pre) | ((post) << 2) | ((blue) << 4))
// Bit of a catalogue kernel's variant number (not of Variant<V>): the kernel
// uses StreamIO instead of PlainIO.
#define RK_VAR_STREAM 128
// Bits of a catalogue kernel's variant number: the kernel has the pass's load
// and store modes compiled in (DIRECT/DIRECT, DIRECT/FLAT, FLAT/FLAT) instead of reading
// them from the pass arguments.  The run-time switch keeps the code of every
// mode alive in one kernel: a 96-point tile holds 146 VGPRs with it and 75-82
// with its modes fixed.
#define RK_VAR_MODE_DD 256
#define RK_VAR_MODE_DF 512
#define RK_VAR_MODE_FF 1024
// Bits of a catalogue kernel's variant number: nontemporal loads only / stores
// only (GlobalIO<true, false> / GlobalIO<false, true>; RK_VAR_STREAM is both).
#define RK_VAR_NT_LOAD 2048
#define RK_VAR_NT_STORE 4096
// Bit of a catalogue kernel's variant number: the row mapping is compiled in
// and each wave holds whole instances (64 % T == 0), so the exchanges between
// Stockham stages stay inside one wave and need no work-group barrier; only
// the staged global sides (COLUMN, FLAT) exchange across waves.
#define RK_VAR_WAVE 8192
// Bit of a catalogue kernel's variant number: the persistent, software-
// pipelined body (fft_pass_body_pipe; DIRECT loads only).
#define RK_VAR_PIPE 16384
// Bit of a catalogue kernel's variant number: the tile's exchanges go through
// re/im halves even when the whole tile would fit (half the LDS per
// work-group, twice the resident work-groups).
#define RK_VAR_HALF 32768
// Bit of a catalogue kernel's variant number: the launch bounds ask for as
// many resident work-groups per CU as the tile's LDS allows, so the register
// budget does not cut the residency to fewer.
#define RK_VAR_OCC 65536
// Bit of a kernel's variant number: complex128 elements (FP64 butterflies,
// twiddles and tables); such kernels come from the run-time compiler only.
#define RK_VAR_F64 131072
// MODES template argument of fft_pass_tile: 0 reads the modes from the pass
// arguments, RK_MODES(load, store) compiles them in.
#define RK_MODES(load, store) (16 | (int)(load) | ((int)(store) << 2))
template <int V> struct Variant {
    static constexpr int PRE = V & 3;
    static constexpr int POST = (V >> 2) & 3;
    static constexpr bool BLUE = ((V >> 4) & 1) != 0;
};

// Padded sides: only a Bluestein chirp pass reads positions beyond its data
// (the input beyond n_valid_in reads as zero) and only the last pass of a
// Bluestein chain leaves positions unwritten (beyond n_valid_out); every
// other pass reads and writes all N positions of its axis (the planners check
// this), so its elements need no per-element bounds.
template <int V> __host__ __device__ constexpr bool pad_in() { return Variant<V>::PRE == RK_VAR_PRE_CHIRP; }
template <int V> __host__ __device__ constexpr bool pad_out() { return Variant<V>::POST == RK_VAR_POST_CONJ_CHIRP; }

// One thread's walk over the positions p0 + c of one instance on one side
// (c a compile-time offset): axis position and element address advance by
// uniform steps computed once per thread, so an element costs one add and no
// multiply.
struct Walk {
    rk_u64 m0, mstep;   // axis position of position p0, and per unit of c
    rk_i64 a0, astep;   // element address of position p0, and per unit of c
};

__device__ __forceinline__ Walk in_walk(const rk_pass_args& a, const Inst& ic, int p0) {
    Walk w;
    w.m0 = ic.jpass + (rk_u64)p0 * a.pos_mult;
    w.mstep = a.pos_mult;
    w.a0 = ic.in_base + (rk_i64)w.m0 * a.in_sa;
    w.astep = (rk_i64)a.pos_mult * a.in_sa;
    return w;
}

// output position k of an instance sits at axis position kbase + k * dim_NS
__device__ __forceinline__ Walk out_walk(const rk_pass_args& a, const Inst& ic, int k0) {
    Walk w;
    w.m0 = ic.kbase + (rk_u64)k0 * a.dim_NS;
    w.mstep = a.dim_NS;
    w.a0 = ic.out_base + (rk_i64)w.m0 * a.out_sa;
    w.astep = (rk_i64)a.dim_NS * a.out_sa;
    return w;
}

// Loads position p0 + c of the walk's instance (raw: no sign, no pre-op).  An
// invalid instance (past n_inst) was decomposed as instance 0, so its
// addresses are valid: it reads them and its results are never stored.
template <int V, class IO>
__device__ __forceinline__ typename IO::value_type load_at(const rk_pass_args& a, const IO& io, const Inst& ic,
                                                           const Walk& w, int c) {
    using X = typename IO::value_type;
    const rk_u64 m = w.m0 + (rk_u64)c * w.mstep;
    const rk_i64 addr = w.a0 + (rk_i64)c * w.astep;
    const X zero = cx<X>((real_t<X>)0, (real_t<X>)0);
    if constexpr (pad_in<V>()) {
        const bool ok = ic.valid && m < a.n_valid_in;
        if constexpr (IO::kBranchFree) {
            // a lane with nothing to load reads its instance's first position
            // (a valid address, different per instance) and discards it: one
            // shared clamp address would be a hot spot for loads that bypass
            // the caches
            const rk_u64 spare_m = ic.jpass < a.n_valid_in ? ic.jpass : 0;
            const rk_i64 spare = ic.in_base + (rk_i64)spare_m * a.in_sa;
            const X y = io.load(a, ic.o, ok ? m : spare_m, ic.i, ok ? addr : spare);
            return c_select(ok, y, zero);
        } else {
            return ok ? io.load(a, ic.o, m, ic.i, addr) : zero;
        }
    } else {
        if constexpr (IO::kBranchFree) {
            return io.load(a, ic.o, m, ic.i, addr);
        } else {
            return ic.valid ? io.load(a, ic.o, m, ic.i, addr) : zero;
        }
    }
}

// Input-side pre-processing of n loaded values at positions p0 + c_i
// (c_i = cpos(i)): conjugation for the inverse direction (a uniform branch),
// then the Bluestein chirp (compile-time).  The inter-pass twiddle (PRE_TWIDDLE)
// is applied by the caller's TwiddleWalk.
template <int V, int NV, class X, class FC>
__device__ __forceinline__ void pre_ops(const rk_pass_args& a, const Inst& ic, const Walk& w, X* x, FC cpos) {
    if (a.in_sign < 0.0f) {
#pragma unroll
        for (int i = 0; i < NV; ++i) x[i].y = -x[i].y;
    }
    if constexpr (Variant<V>::PRE == RK_VAR_PRE_CHIRP) {
#pragma unroll
        for (int i = 0; i < NV; ++i) {
            const rk_u64 m = w.m0 + (rk_u64)cpos(i) * w.mstep;
            const bool ok = ic.valid && m < a.n_valid_in;
            x[i] = c_mul(x[i], table<X>(a.chirp)[ok ? m : 0]);
        }
    }
}

// The output-side operations of one value at axis position kabs: the
// compile-time post-op, then conjugation and scale for the direction.
template <class X> __device__ __forceinline__ X adjust_out(const rk_pass_args& a, X y) {
    y.y *= (real_t<X>)a.out_sign;
    return c_scale(y, out_scale_of<X>(a));
}

template <int V, class X>
__device__ __forceinline__ X post_op(const rk_pass_args& a, rk_u64 kabs, X y, bool adjust) {
    constexpr int POST = Variant<V>::POST;
    if constexpr (POST == RK_VAR_POST_KMUL_CONJ) y = c_conj(c_mul(y, table<X>(a.bspec)[kabs]));
    if constexpr (POST == RK_VAR_POST_CONJ_CHIRP) y = c_mul(c_conj(y), table<X>(a.chirp)[kabs]);
    if (adjust) y = adjust_out(a, y);
    return y;
}

// Stores position k0 + c of the walk's instance (the caller skips invalid
// instances); a padded side skips positions at or beyond n_valid_out.
template <int V, class IO>
__device__ __forceinline__ void store_at(const rk_pass_args& a, const IO& io, const Inst& ic, const Walk& w, int c,
                                         typename IO::value_type y, bool adjust) {
    const rk_u64 kabs = w.m0 + (rk_u64)c * w.mstep;
    if constexpr (pad_out<V>())
        if (kabs >= a.n_valid_out) return;
    io.store(a, ic.o, kabs, ic.i, w.a0 + (rk_i64)c * w.astep, post_op<V>(a, kabs, y, adjust));
}

// ------------------------------------------------------------------------
// LDS exchange between two register layouts
// ------------------------------------------------------------------------
// Every value a thread holds in `w` is written to LDS at wpos(i) (skipped
// when negative) and every value of `r` is read from rpos(i).  With SPLIT the
// real and imaginary parts travel separately through a buffer of floats, so
// the tile needs half the LDS (two work-groups then fit on a CU); `w` and `r`
// may be the same array.
// Ordering point of an exchange.  Work-group scope: a barrier with
// work-group fences (__syncthreads).  Wave scope (every position a lane
// writes or reads belongs to instances its own wave holds): no barrier; the
// wave-scope fences keep the compiler from moving LDS accesses across the
// point, and the wait for the wave's outstanding LDS operations
// (lgkmcnt(0): s_waitcnt encoding 0xc07f leaves vmcnt and expcnt at their
// maxima) completes its writes before its reads.
template <bool WAVE> __device__ __forceinline__ void exchange_sync() {
    if constexpr (WAVE) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_s_waitcnt(0xc07f);
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
        __syncthreads();
    }
}

template <bool SPLIT, int NW, int NR, bool WAVE = false, class X, class FW, class FR>
__device__ __forceinline__ void lds_exchange(void* base, const X* w, X* r, FW wpos, FR rpos) {
    using F = real_t<X>;
    if constexpr (!SPLIT) {
        X* lds = reinterpret_cast<X*>(base);
        exchange_sync<WAVE>();
#pragma unroll
        for (int i = 0; i < NW; ++i) {
            const int p = wpos(i);
            if (p >= 0) lds[p] = w[i];
        }
        exchange_sync<WAVE>();
#pragma unroll
        for (int i = 0; i < NR; ++i) {
            const int p = rpos(i);
            if (p >= 0) r[i] = lds[p];
        }
    } else {
        F* f = reinterpret_cast<F*>(base);
        F wy[NW], rx[NR];
#pragma unroll
        for (int i = 0; i < NW; ++i) wy[i] = w[i].y;
        exchange_sync<WAVE>();
#pragma unroll
        for (int i = 0; i < NW; ++i) {
            const int p = wpos(i);
            if (p >= 0) f[p] = w[i].x;
        }
        exchange_sync<WAVE>();
#pragma unroll
        for (int i = 0; i < NR; ++i) {
            const int p = rpos(i);
            rx[i] = p >= 0 ? f[p] : (F)0;
        }
        exchange_sync<WAVE>();
#pragma unroll
        for (int i = 0; i < NW; ++i) {
            const int p = wpos(i);
            if (p >= 0) f[p] = wy[i];
        }
        exchange_sync<WAVE>();
#pragma unroll
        for (int i = 0; i < NR; ++i) {
            const int p = rpos(i);
            if (p >= 0) r[i] = cx<X>(rx[i], f[p]);
        }
    }
}

// ------------------------------------------------------------------------
// the Stockham stages of one tile
// ------------------------------------------------------------------------
// Stage s holds, in thread t, the butterflies j = t + q*T (q < Q_s, j < NB_s);
// v[q*R + r] is input r of butterfly j, at position j + r*NB_s of the
// instance.  The butterfly's outputs go to (j/Ns)*Ns*R + j%Ns + r*Ns, which is
// where stage s+1 reads them; after the last stage v[q*R + r] holds output
// k = j + r*NB.
template <int L, int S>
RK_HD constexpr bool stage_valid(int t, int q) {
    using P = Plan<L>;
    return (P::nb(S) % P::T) == 0 || (t + q * P::T) < P::nb(S);
}

template <int L, int S>
RK_HD constexpr int stage_in_pos(int t, int i) {   // input position of v[i]
    using P = Plan<L>;
    constexpr int R = P::radix(S);
    const int q = i / R, r = i % R;
    return (t + q * P::T) + r * P::nb(S);
}

template <int L, int S>
RK_HD constexpr int stage_out_pos(int t, int i) {  // output position of v[i]
    using P = Plan<L>;
    constexpr int R = P::radix(S);
    constexpr int NS = P::ns(S);
    const int q = i / R, r = i % R;
    const int j = t + q * P::T;
    return (j / NS) * NS * R + (j % NS) + r * NS;
}

// The padded LDS positions of stage_in_pos / stage_out_pos (lds_pad of them),
// written as a per-thread part plus a per-value constant where lds_pad_add
// allows.  Input: t + (q*T + r*NB) with t < T.  Output: when T % NS == 0,
// j / NS = t / NS + q*T/NS and j % NS = t % NS, so the position is
// (t/NS)*NS*R + t%NS + (q*T*R + r*NS), whose per-thread part is below NS
// modulo 16 when NS < 16 and NS*R % 16 == 0; when the stage has no more
// butterflies than NS (j < NS) it is t + (q*T + r*NS).
template <int L, int S>
RK_HD constexpr int stage_in_lds(int t, int i) {
    using P = Plan<L>;
    constexpr int R = P::radix(S);
    constexpr int T = P::T;
    const int q = i / R, r = i % R;
    return lds_pad_add<(T < 16 ? T : 16)>(t, q * T + r * P::nb(S));
}

template <int L, int S>
RK_HD constexpr int stage_out_lds(int t, int i) {
    using P = Plan<L>;
    constexpr int R = P::radix(S);
    constexpr int NS = P::ns(S);
    constexpr int T = P::T;
    const int q = i / R, r = i % R;
    if constexpr (T % NS == 0) {
        constexpr int BM = (NS < 16 && (NS * R) % 16 == 0) ? NS : 16;
        return lds_pad_add<BM>((t / NS) * NS * R + t % NS, q * T * R + r * NS);
    } else if constexpr (P::nb(S) <= NS) {
        return lds_pad_add<(T < 16 ? T : 16)>(t, q * T + r * NS);
    } else {
        return lds_pad(stage_out_pos<L, S>(t, i));
    }
}

// Padded LDS position, within a tile of pitch LP, of element it of thread tid
// on an instance-linear (FLAT) side: e = tid + it*NT is instance e / L,
// position e % L.  When NT is a multiple of L every thread stays on one
// position of its instances; when L is a multiple of NT, iteration it covers
// positions (it*NT) % L + tid of instance (it*NT) / L.
template <int L, int NT, int LP>
RK_HD constexpr int flat_lds(int tid, int it) {
    if constexpr (NT % L == 0) {
        return (tid / L) * LP + lds_pad(tid % L) + it * (NT / L) * LP;
    } else if constexpr (L % NT == 0) {
        return ((it * NT) / L) * LP + lds_pad_add<(NT < 16 ? NT : 16)>(tid, (it * NT) % L);
    } else {
        const int e = tid + it * NT;
        const int ce = e / L;
        return ce * LP + lds_pad(e - ce * L);
    }
}

// Padded LDS position of element it of a thread at position ps < T of an
// instance on a staged COLUMN / ROW side: ps + it*T.
template <int T> RK_HD constexpr int col_lds(int ps, int it) { return lds_pad_add<(T < 16 ? T : 16)>(ps, it * T); }

template <int L, int S, bool SPLIT, int VM, bool WAVE = false, class X>
__device__ __forceinline__ void run_stages(X (&v)[VM], void* my, int t, const X* twl) {
    using P = Plan<L>;
    constexpr int R = P::radix(S);
    constexpr int NS = P::ns(S);
    constexpr int Q = P::q(S);
    constexpr int T = P::T;
    if constexpr (S > 0) {
        // stage twiddles W_{NS*R}^{r*(j%NS)} from the FP64-rounded W_L table
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int j = t + q * T;
            if (stage_valid<L, S>(t, q)) {
                const int m = j % NS;
#pragma unroll
                for (int r = 1; r < R; ++r)
                    v[q * R + r] = c_mul(v[q * R + r], twl[(r * m) * (L / (NS * R))]);
            }
        }
    }
#pragma unroll
    for (int q = 0; q < Q; ++q)
        if (stage_valid<L, S>(t, q)) dft<R>(&v[q * R]);
    if constexpr (S + 1 < P::NSTAGE) {
        constexpr int R1 = P::radix(S + 1);
        constexpr int Q1 = P::q(S + 1);
        lds_exchange<SPLIT, Q * R, Q1 * R1, WAVE>(
            my, v, v,
            [&](int i) { return stage_valid<L, S>(t, i / R) ? stage_out_lds<L, S>(t, i) : -1; },
            [&](int i) { return stage_valid<L, S + 1>(t, i / R1) ? stage_in_lds<L, S + 1>(t, i) : -1; });
        run_stages<L, S + 1, SPLIT, VM, WAVE>(v, my, t, twl);
    }
}

// ------------------------------------------------------------------------
// the pass kernel body
// ------------------------------------------------------------------------
// Tile of this work-group.  The dispatcher hands work-group b to XCD
// b % n_xcd; with xcd_group = G > 1 the G tiles of each group of consecutive
// tiles go to work-groups on one XCD, so tiles that share cache lines (narrow
// column tiles) share that XCD's L2 instead of each fetching whole lines.
__device__ __forceinline__ rk_u64 xcd_tile(const rk_pass_args& a) {
    const rk_u32 b = blockIdx.x;
    const rk_u32 g = a.xcd_group, nx = a.n_xcd;
    if (g <= 1) return b;
    const rk_u32 span = g * nx;
    const rk_u32 full = (gridDim.x / span) * span;
    if (b >= full) return b;
    const rk_u32 x = b % nx, k = b / nx;
    return (rk_u64)(k / g) * span + x * g + (k % g);
}

// LDS bytes of a tile and whether its exchanges are split into halves.
// HALF forces the split exchange on a tile that would fit whole (half the LDS,
// twice the resident work-groups).
template <int L, int C, bool HALF = false, int ES = 8> __host__ __device__ constexpr bool tile_split() {
    return HALF || lds_split(L, C, ES);
}
template <int L, int C, bool HALF = false, int ES = 8> __host__ __device__ constexpr int tile_pitch() {
    return lds_pitch_for(L, C, Plan<L>::T, HALF, ES);
}
template <int L, int C, bool HALF = false, int ES = 8> __host__ __device__ constexpr int tile_lds_bytes() {
    return C * tile_pitch<L, C, HALF, ES>() * (tile_split<L, C, HALF, ES>() ? ES / 2 : ES);
}

// One tile (instances tile*C ... tile*C + C - 1) computed by threads
// tid = 0 ... C*T - 1 with `lds` of at least tile_lds_bytes<L, C, HALF, ES>()
// bytes (ES = sizeof the IO policy's element type).
// Every value a thread computes and every address it writes is a function of
// (a, tile, tid) alone, so a caller with more threads than C*T may give the
// surplus threads tid = C*T - 1: they repeat that thread's work, writing the
// same values to the same addresses, and take part in every barrier.
// With PRE, the tile's DIRECT loads were issued earlier (raw values in
// `pre`, see direct_raw_loads): the tile starts from them instead of loading.
template <int L, int C, int V, class IO, int MODES = 0, bool WAVE = false, bool PRE = false, bool HALF = false,
          bool TWLDS = false>
__device__ __forceinline__ void fft_pass_tile(const rk_pass_args& a, const IO& io, typename IO::value_type* lds,
                                              rk_u64 tile, int tid, const typename IO::value_type* pre = nullptr,
                                              const typename IO::value_type* twl_lds = nullptr) {
    using X = typename IO::value_type;
    using F = real_t<X>;
    constexpr int ES = (int)sizeof(X);
    using P = Plan<L>;
    static_assert(!WAVE || (64 % P::T == 0), "wave-local stages need whole instances per wave");
    constexpr int T = P::T;
    constexpr int NT = C * T;
    constexpr int R0 = P::R0;
    constexpr int LP = tile_pitch<L, C, HALF, ES>();
    constexpr int VM = P::vmax();
    constexpr int SL = P::NSTAGE - 1;
    constexpr int RL = P::radix(SL);
    constexpr int QL = P::q(SL);
    constexpr bool BLUE = Variant<V>::BLUE;
    constexpr bool TWID = Variant<V>::PRE == RK_VAR_PRE_TWIDDLE;
    constexpr bool SPLIT = tile_split<L, C, HALF, ES>();
    const rk_u64 inst0 = tile * (rk_u64)C;
    const rk_u32 load_mode = MODES ? (rk_u32)(MODES & 3) : a.load_mode;
    const rk_u32 store_mode = MODES ? (rk_u32)((MODES >> 2) & 3) : a.store_mode;

    // compute mapping (see fft_args.h): lanes along an instance or across
    // instances (WAVE: the row mapping, compiled in)
    const bool colmap = WAVE ? false : a.cmap != 0;
    const int c = colmap ? (tid % C) : (tid / T);
    int t = colmap ? (tid / C) : (tid % T);
    const int cbase = c * LP;
    // the other mapping, for a side staged through LDS
    const int cs = colmap ? (tid / T) : (tid % C);
    const int ps = colmap ? (tid % T) : (tid / C);

    X v[VM];
    X g[R0];   // staging registers (global side of a staged load/store)

    // ---- stage-0 inputs: element p = t + r*T of instance c ----
    const Inst ic = decompose<L>(a, inst0 + c);
    auto own_pos = [&](int i) { return cbase + stage_in_lds<L, 0>(t, i); };
    if (PRE || load_mode == RK_MODE_DIRECT) {
        const Walk w = in_walk(a, ic, t);
        if constexpr (PRE) {
#pragma unroll
            for (int r = 0; r < R0; ++r) v[r] = pre[r];
        } else {
#pragma unroll
            for (int r = 0; r < R0; ++r) v[r] = load_at<V, IO>(a, io, ic, w, r * T);
        }
        pre_ops<V, R0>(a, ic, w, v, [](int r) { return r * T; });
        if constexpr (TWID) apply_pass_twiddles<R0>(a, ic.l, t, T, v);
    } else if (load_mode == RK_MODE_FLAT) {
        // instance-linear block of C*L elements, no pre-ops
        const rk_i64 base = a.in_off + (rk_i64)inst0 * L;
#pragma unroll
        for (int it = 0; it < R0; ++it) {
            const int e = tid + it * NT;
            const int ce = e / L, p = e - ce * L;
            X x = cx<X>((F)0, (F)0);
            const bool ok = inst0 + ce < a.n_inst;
            if constexpr (IO::kBranchFree) {
                // a lane past the last instance loads the tile's first element
                const X y = io.load(a, ok ? inst0 + ce : inst0, ok ? p : 0, 0, ok ? base + e : base);
                x = c_select(ok, y, x);
            } else {
                if (ok) x = io.load(a, inst0 + ce, p, 0, base + e);
            }
            g[it] = x;
        }
        if (a.in_sign < 0.0f) {
#pragma unroll
            for (int it = 0; it < R0; ++it) g[it].y = -g[it].y;
        }
        lds_exchange<SPLIT, R0, R0>(
            lds, g, v,
            [&](int it) { return flat_lds<L, NT, LP>(tid, it); }, own_pos);
    } else {
        // COLUMN (under the row mapping) or ROW (under the column mapping)
        const Inst ics = decompose<L>(a, inst0 + cs);
        const Walk w = in_walk(a, ics, ps);
#pragma unroll
        for (int it = 0; it < R0; ++it) g[it] = load_at<V, IO>(a, io, ics, w, it * T);
        pre_ops<V, R0>(a, ics, w, g, [](int it) { return it * T; });
        if constexpr (TWID) apply_pass_twiddles<R0>(a, ics.l, ps, T, g);
        lds_exchange<SPLIT, R0, R0>(
            lds, g, v, [&](int it) { return cs * LP + col_lds<T>(ps, it); }, own_pos);
    }

    // LDS of this thread's instance (the stages index it relative to the instance)
    void* my = SPLIT ? (void*)(reinterpret_cast<F*>(lds) + cbase) : (void*)(lds + cbase);
    auto last_out = [&](int i) {
        return stage_valid<L, SL>(t, i / RL) ? cbase + stage_out_lds<L, SL>(t, i) : -1;
    };
    const X* twl;
    if constexpr (TWLDS) twl = twl_lds;
    else twl = table<X>(a.twl);
    run_stages<L, 0, SPLIT, VM, WAVE>(v, my, t, twl);

    if constexpr (BLUE) {
        // A = FFT(chirp * x): multiply by the kernel spectrum, conjugate, and
        // transform again (the inverse DFT as conj(FFT(conj .))).
#pragma unroll
        for (int i = 0; i < QL * RL; ++i)
            if (stage_valid<L, SL>(t, i / RL)) {
                const int k = stage_out_pos<L, SL>(t, i);
                v[i] = c_conj(c_mul(v[i], table<X>(a.bspec)[k]));
            }
        lds_exchange<SPLIT, QL * RL, R0, WAVE>(lds, v, v, last_out, own_pos);
        // the second transform and the stores recompute what depends on the
        // thread index: the index goes through an empty asm, so the first
        // transform's stage twiddles and the output chirp values are not held
        // in registers across both transforms (the 2048-point tile: 162 -> 124
        // VGPRs, four waves per SIMD instead of two)
        asm volatile("" : "+v"(t));
        if constexpr (TWLDS) run_stages<L, 0, SPLIT, VM, WAVE>(v, my, t, twl_lds);
        else run_stages<L, 0, SPLIT, VM, WAVE>(v, my, t, table<X>(a.twl));
    }

    // ---- store: output element k = stage_out_pos of the last stage ----
    const bool adjust = a.out_sign < 0.0f || a.out_scale != 1.0f;
    if (store_mode == RK_MODE_DIRECT) {
        // last stage: butterfly j = t + q*T < NS, so output k = t + stage_out_pos(0, i)
        static_assert(P::nb(SL) == P::ns(SL), "the last Stockham stage writes its outputs at j + r * NS");
        if (ic.valid) {
            const Walk w = out_walk(a, ic, t);
#pragma unroll
            for (int i = 0; i < QL * RL; ++i)
                if (stage_valid<L, SL>(t, i / RL)) store_at<V, IO>(a, io, ic, w, stage_out_pos<L, SL>(0, i), v[i], adjust);
        }
    } else if (store_mode == RK_MODE_FLAT) {
        // no post-ops beyond sign and scale
        lds_exchange<SPLIT, QL * RL, R0>(lds, v, g, last_out, [&](int it) { return flat_lds<L, NT, LP>(tid, it); });
        const rk_i64 base = a.out_off + (rk_i64)inst0 * L;
#pragma unroll
        for (int it = 0; it < R0; ++it) {
            const int e = tid + it * NT;
            const int ce = e / L, k = e - ce * L;
            if (inst0 + ce < a.n_inst) {
                X y = g[it];
                if (adjust) y = adjust_out(a, y);
                io.store(a, inst0 + ce, k, 0, base + e, y);
            }
        }
    } else {
        lds_exchange<SPLIT, QL * RL, R0>(lds, v, g, last_out, [&](int it) { return cs * LP + col_lds<T>(ps, it); });
        const Inst ics = decompose<L>(a, inst0 + cs);
        if (ics.valid) {
            const Walk w = out_walk(a, ics, ps);
#pragma unroll
            for (int it = 0; it < R0; ++it) store_at<V, IO>(a, io, ics, w, it * T, g[it], adjust);
        }
    }
}

// Threads per work-group of fft_pass_body<L, C, ...>.
template <int L, int C> __host__ __device__ constexpr int pass_threads() { return C * Plan<L>::T; }

// Tiles whose stage twiddles are read from an LDS copy of the W_L table
// instead of global memory (the table loads otherwise take a large share of
// the vector-memory instructions): complex64 tiles with at least one stage
// after the first and exchanges with work-group barriers (no wave-local
// stages), where the extra L entries leave the
// tile's LDS-limited residency per CU unchanged, or cost at most 16 KB and
// leave two work-groups per CU.  Measured on row tiles, whose residency the
// table does lower: 4096 x 1000 20.0 -> 17.6 us, 16384 x 2048 117 -> 102 us,
// 16384 x 1024 50.5 -> 47.7 us, the single-kernel Bluestein 4096 x 997 (both
// of its transforms read the table) 53 -> 44 us; 4096-point rows (a 32 KB
// table, half the residency) 64 -> 70 us, so they keep the global table.
template <int L, int C, int V, bool WAVE, bool HALF, int ES> __host__ __device__ constexpr bool twl_in_lds() {
    if (WAVE || ES != 8 || Plan<L>::NSTAGE < 2) return false;
    constexpr int base = tile_lds_bytes<L, C, HALF, ES>();
    constexpr int with = base + L * ES;
    return (160 * 1024) / with == (160 * 1024) / base || (L * ES <= 16 * 1024 && (160 * 1024) / with >= 2);
}

// One work-group of a pass launch: the tile this work-group owns, with the
// work-group's own LDS.
template <int L, int C, int V, class IO, int MODES = 0, bool WAVE = false, bool HALF = false>
__device__ __forceinline__ void fft_pass_body(const rk_pass_args& a, const IO& io) {
    using X = typename IO::value_type;
    constexpr int ES = (int)sizeof(X);
    constexpr int LP = tile_pitch<L, C, HALF, ES>();
    __shared__ X lds[tile_split<L, C, HALF, ES>() ? (C * LP + 1) / 2 : C * LP];
    const rk_u64 tile = xcd_tile(a) + a.tile_base;
    if (tile >= a.n_tiles) return;
    if constexpr (twl_in_lds<L, C, V, WAVE, HALF, ES>()) {
        // the stage twiddles W_L^e from LDS, published by the tile's first
        // exchange barrier (before stage 1 reads them)
        __shared__ X twl_s[L];
        for (int e = (int)threadIdx.x; e < L; e += pass_threads<L, C>()) twl_s[e] = table<X>(a.twl)[e];
        fft_pass_tile<L, C, V, IO, MODES, WAVE, false, HALF, true>(a, io, lds, tile, (int)threadIdx.x, nullptr, twl_s);
    } else {
        fft_pass_tile<L, C, V, IO, MODES, WAVE, false, HALF>(a, io, lds, tile, (int)threadIdx.x);
    }
}

// The raw DIRECT loads of one tile (what fft_pass_tile<..., PRE> starts from).
template <int L, int C, int V, class IO, bool WAVE>
__device__ __forceinline__ void direct_raw_loads(const rk_pass_args& a, const IO& io, rk_u64 tile, int tid,
                                                 typename IO::value_type* raw) {
    using P = Plan<L>;
    constexpr int T = P::T;
    const bool colmap = WAVE ? false : a.cmap != 0;
    const int c = colmap ? (tid % C) : (tid / T);
    const int t = colmap ? (tid / C) : (tid % T);
    const Inst ic = decompose<L>(a, tile * (rk_u64)C + c);
    const Walk w = in_walk(a, ic, t);
#pragma unroll
    for (int r = 0; r < P::R0; ++r) raw[r] = load_at<V, IO>(a, io, ic, w, r * T);
}

// Persistent, software-pipelined pass (DIRECT loads only): work-group b runs
// the tiles tile_base + b, + G, + 2G, ... below n_tiles (G = gridDim.x) and
// issues the global loads of its next tile before it computes and stores the
// current one, so one tile's loads are in flight while another is computed.
// The host launches about one work-group per resident slot.
template <int L, int C, int V, class IO, int MODES, bool WAVE, bool HALF>
__device__ __forceinline__ void fft_pass_body_pipe(const rk_pass_args& a, const IO& io) {
    using X = typename IO::value_type;
    constexpr int ES = (int)sizeof(X);
    constexpr int LP = tile_pitch<L, C, HALF, ES>();
    constexpr int R0 = Plan<L>::R0;
    __shared__ X lds[tile_split<L, C, HALF, ES>() ? (C * LP + 1) / 2 : C * LP];
    const int tid0 = (int)threadIdx.x;
    const rk_u64 G = gridDim.x;
    rk_u64 tile = a.tile_base + blockIdx.x;
    if (tile >= a.n_tiles) return;
    X nxt[R0];
    direct_raw_loads<L, C, V, IO, WAVE>(a, io, tile, tid0, nxt);
    for (;;) {
        // everything derived from the thread index (stage twiddle indices and
        // loads, LDS positions) is recomputed per tile instead of being
        // hoisted out of the loop and held in registers across tiles: the
        // empty asm makes the index opaque to the optimiser each iteration
        int tid = tid0;
        asm volatile("" : "+v"(tid));
        X cur[R0];
#pragma unroll
        for (int r = 0; r < R0; ++r) cur[r] = nxt[r];
        const rk_u64 next = tile + G;
        if (next < a.n_tiles) direct_raw_loads<L, C, V, IO, WAVE>(a, io, next, tid, nxt);
        fft_pass_tile<L, C, V, IO, MODES, WAVE, true, HALF>(a, io, lds, tile, tid, cur);
        if (next >= a.n_tiles) break;
        tile = next;
    }
}

// The element type of a variant number: complex128 with RK_VAR_F64.
template <bool F64> struct ElemOf { typedef float2 type; };
template <> struct ElemOf<true> { typedef double2 type; };
template <int V> using elem_t = typename ElemOf<(V & RK_VAR_F64) != 0>::type;

// The pass kernel of a catalogue variant number V: the low seven bits are the
// compile-time Variant (pre-op, post-op, Bluestein), the higher bits choose the
// IO policy (RK_VAR_STREAM, RK_VAR_NT_LOAD, RK_VAR_NT_STORE), compiled-in
// load/store modes (RK_VAR_MODE_*), wave-local stages (RK_VAR_WAVE) and the
// element type (RK_VAR_F64).  The ahead-of-time catalogue and the run-time
// compiler instantiate this same entry.
template <int L, int C, int V>
__device__ __forceinline__ void fft_pass_entry(const rk_pass_args& a) {
    using X = elem_t<V>;
    constexpr int VB = V & (RK_VAR_STREAM - 1);
    constexpr int M = (V & RK_VAR_MODE_DD) ? RK_MODES(RK_MODE_DIRECT, RK_MODE_DIRECT)
                    : (V & RK_VAR_MODE_DF) ? RK_MODES(RK_MODE_DIRECT, RK_MODE_FLAT)
                    : (V & RK_VAR_MODE_FF) ? RK_MODES(RK_MODE_FLAT, RK_MODE_FLAT) : 0;
    constexpr bool NTL = (V & (RK_VAR_STREAM | RK_VAR_NT_LOAD)) != 0;
    constexpr bool NTS = (V & (RK_VAR_STREAM | RK_VAR_NT_STORE)) != 0;
    constexpr bool WAVE = (V & RK_VAR_WAVE) != 0;
    constexpr bool HALF = (V & RK_VAR_HALF) != 0;
    if constexpr ((V & RK_VAR_PIPE) != 0)
        fft_pass_body_pipe<L, C, VB, GlobalIO<NTL, NTS, X>, M, WAVE, HALF>(a, GlobalIO<NTL, NTS, X>{});
    else
        fft_pass_body<L, C, VB, GlobalIO<NTL, NTS, X>, M, WAVE, HALF>(a, GlobalIO<NTL, NTS, X>{});
}

// Minimum waves per SIMD of a pass kernel's launch bounds: three for the
// single-kernel Bluestein tiles with DIRECT/DIRECT compiled in of at most 256
// threads (at most 168 VGPRs), else one (the tile's own thread count bounds
// its registers).
// An RK_VAR_OCC kernel asks for the waves per SIMD its LDS allows, so
// registers do not cut the occupancy; a persistent (RK_VAR_PIPE) kernel
// without it for one resident work-group per CU (the registers of its
// prefetched tile then come out of that work-group's share).
template <int L, int C, int V> __host__ __device__ constexpr int pass_min_waves() {
    if ((V & (RK_VAR_PIPE | RK_VAR_OCC)) != 0) {
        const int waves = (pass_threads<L, C>() + 63) / 64;
        const int slots = (V & RK_VAR_OCC) ? (160 * 1024) / tile_lds_bytes<L, C, (V & RK_VAR_HALF) != 0,
                                                                           (int)sizeof(elem_t<V>)>() : 1;
        const int w = (slots * waves) / 4;
        return w < 1 ? 1 : (w > 8 ? 8 : w);
    }
    return ((V & RK_VAR_MODE_DD) && (V & 16) && pass_threads<L, C>() <= 256) ? 3 : 1;
}

}   // namespace rkfft

#endif
