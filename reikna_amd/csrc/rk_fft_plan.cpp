// rk_fft_plan.cpp — the FFT planner and executor behind rk_fft_plan_create /
// rk_fft_execute.
//
// What it replaces in the reference:
//   get_fft_kernels / get_fft_1d_kernels (reikna/fft/fft.py:487-584): the
//     per-axis kernel chain.  Here an axis of smooth length N (2^a 3^b 5^c 7^d)
//     is one pass when N <= kSingleMax and otherwise two or more Stockham
//     passes whose lengths multiply to N; a non-smooth axis goes through
//     Bluestein's chirp-z transform over a smooth M >= 2N-1 (the reference uses
//     M = 2 * bounding_power_of_2(N), fft.py:551-561).
//   _build_limited_plan buffer routing (fft.py:664-721): the first pass reads
//     `in`, the last writes `out`, intermediate passes ping-pong between two
//     workspace buffers.
//   get_kweights (fft.py:216-236): the Bluestein kernel spectrum, here an FP64
//     host FFT rounded once to complex64.
//   xweight (fft.mako:196-202): the chirp exp(-i pi (n^2 mod 2N)/N) with the
//     exact integer n^2 mod 2N, here evaluated in FP64 and rounded once.
//   the run-time `inverse` flag and 1/N normalisation (fft.mako:741-747).
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>

#include "fft_device.hpp"
#include "rk_internal.hpp"

using rk::fail;

namespace rk {

// Knobs of the plan being built on this thread (rk_fft_plan_create_tuned;
// null: the planner's own choices), and its element bytes (8 complex64, 16
// complex128).  Set for the duration of one plan creation only.
static thread_local const rk_fft_tuning* t_tuning = nullptr;
static thread_local int t_elem_bytes = 8;

// Longest single-pass DFT of a smooth length, and longest single-kernel
// Bluestein convolution (RK_SINGLE_MAX overrides both, at most 16384).  Above
// 4096 points a single-pass tile is one instance of 512+ threads with 64+ KB
// of LDS, one or two work-groups per CU, and streams slower than two passes
// of short DFTs: 4096 x 12288 took 413 us in one pass and 345 us in two
// (96 x 128).  A single-kernel Bluestein convolution reads and writes its
// data once instead of running two chains through a workspace of twice the
// length, so it keeps a longer limit, up to the point where its tile of 640+
// threads runs out of registers: 4096 rows of 4093 points (M = 8192) took
// 350 us in one kernel and 453 us as two chains, 4096 x 5003 (M = 10240)
// 1085 us in one kernel and 634 us as two chains, 4096 x 7919 (M = 16384)
// 1177 and 891 us.
static int64_t env_single_max() {
    static const int64_t v = [] {
        const char* e = getenv("RK_SINGLE_MAX");
        return e && *e ? std::min<int64_t>(std::max<int64_t>(atoll(e), 2), 16384) : (int64_t)0;
    }();
    return v;
}
static int64_t single_max() { return env_single_max() ? env_single_max() : 4096; }
// Complex128 tiles hold twice the registers: their single-kernel Bluestein
// stops at M = 2048 (256 threads, 196 VGPRs), longer convolutions run as
// chains.
static int64_t blue_single_max() {
    if (env_single_max()) return env_single_max();
    return t_elem_bytes == 16 ? 2048 : 8192;
}
#define kSingleMax (single_max())

// Longest DFT of one pass in a multi-pass chain.  Column-style passes need
// 16 adjacent instances per tile for 128-byte accesses, so a pass of length L
// holds a 16 x L tile.  Complex64 chains take passes up to 1024 points: the
// 16 x 1024 tile (1024 threads, 68 KB of LDS through re/im halves) fits two
// work-groups per CU in 64 VGPRs (launch bounds, occ below), and a chain
// then needs one pass fewer from 2^17 to 2^20 points.  Measured per 512 MiB
// transform, three passes of <= 256 against two of <= 1024: 512 x 2^17 0.445
// against 0.347 ms, 256 x 2^18 0.449 / 0.376, 128 x 2^19 0.446 / 0.410,
// 64 x 2^20 0.508 / 0.448, 64 x 10^6 0.639 / 0.571; from 2^21 points both
// limits give three passes.  Complex128 tiles hold twice the registers and
// keep 256.  RK_FACTOR_MAX overrides.
static int64_t factor_max() {
    static const int64_t v = [] {
        const char* e = getenv("RK_FACTOR_MAX");
        return e && *e ? (int64_t)atoll(e) : (int64_t)0;
    }();
    if (v) return v;
    return t_elem_bytes == 16 ? 256 : 1024;
}

static bool length_ok(const ChainLimits& lim, int64_t n) { return lim.length_ok ? lim.length_ok(n) : is_smooth(n); }

// Split n into exactly p factors <= lim.factor_max that lim accepts, the
// first as close to n^(1/p) as possible (balanced factors keep every tile small).
static bool split_into(int64_t n, int p, const ChainLimits& lim, std::vector<int64_t>& out) {
    const int64_t fmax = lim.factor_max;
    if (p == 1) {
        if (n <= fmax && length_ok(lim, n)) {
            out.push_back(n);
            return true;
        }
        return false;
    }
    const double target = std::pow((double)n, 1.0 / p);
    std::vector<int64_t> cands;
    for (int64_t d = 2; d <= fmax && d <= n; ++d)
        if (n % d == 0 && length_ok(lim, d)) cands.push_back(d);
    std::sort(cands.begin(), cands.end(), [&](int64_t a, int64_t b) {
        return std::fabs(std::log((double)a / target)) < std::fabs(std::log((double)b / target));
    });
    for (int64_t d : cands) {
        const size_t keep = out.size();
        out.push_back(d);
        if (split_into(n / d, p - 1, lim, out)) return true;
        out.resize(keep);
    }
    return false;
}

ChainLimits default_limits() {
    ChainLimits l;
    l.single_max = single_max();
    l.factor_max = factor_max();
    l.pow2_bluestein = false;
    l.blue_single_max = blue_single_max();
    l.length_ok = nullptr;
    l.strided_single_max = 0;
    if (t_tuning && t_tuning->factor_max > 0) l.factor_max = t_tuning->factor_max;
    if (t_tuning && t_tuning->first_factor > 0) l.first_factor = t_tuning->first_factor;
    return l;
}

int64_t TableBuilder::add(const std::string& key, const void* data, size_t n) {
    auto it = index.find(key);
    if (it != index.end()) return it->second;
    size_t off = (bytes.size() + 255) & ~size_t(255);
    bytes.resize(off + n);
    memcpy(bytes.data() + off, data, n);
    index[key] = (int64_t)off;
    return (int64_t)off;
}

bool is_smooth(int64_t n) {
    if (n < 1) return false;
    for (int p : {2, 3, 5, 7})
        while (n % p == 0) n /= p;
    return n == 1;
}

// ---------------------------------------------------------------------------
// host mirrors of the compile-time geometry in fft_device.hpp
// ---------------------------------------------------------------------------
int threads_per_instance(int L) {
    rkfft::RadixList rl = rkfft::make_radix_list(L);
    return L / rl.r[0];
}

int lds_pitch_host(int L) { return rkfft::lds_pitch_base(L); }

int tile_lds_bytes_host(int L, int C, int es) {
    const int pitch = rkfft::lds_pitch_for(L, C, threads_per_instance(L), false, es);
    return C * pitch * (rkfft::lds_split(L, C, es) ? es / 2 : es);
}

// Tuning overrides of the tile widths for experiments (kernels outside the
// catalog come from the run-time compiler): RK_ROW_TILE / RK_COL_TILE force
// the instances per row / column tile when the tile stays within 1024
// threads and 159 KB of LDS.
// A requested tile width c of length-L instances, or 0 when that tile would
// exceed 1024 threads or 159 KB of LDS.
static int tile_if_fits(int c, int L) {
    if (c <= 0) return 0;
    if ((int64_t)c * threads_per_instance(L) > 1024 || (int64_t)c * lds_pitch_host(L) * t_elem_bytes > 159 * 1024)
        return 0;
    return c;
}

static int forced_tile(const char* name, int L) {
    const char* e = getenv(name);
    return tile_if_fits(e && *e ? atoi(e) : 0, L);
}

static int default_tile_c_row(int L) {
    int T = threads_per_instance(L);
    int C = std::max(1, 256 / T);
    while (C > 1 && (int64_t)C * lds_pitch_host(L) * t_elem_bytes > 64 * 1024) C /= 2;
    return C;
}

int tile_c_row(int L) {
    if (t_tuning && t_tuning->row_tile > 0)
        if (int f = tile_if_fits(t_tuning->row_tile, L)) return f;
    static const bool any = getenv("RK_ROW_TILE") != nullptr;
    if (any)
        if (int f = forced_tile("RK_ROW_TILE", L)) return f;
    return default_tile_c_row(L);
}

int tile_c_col(int L) {
    if (t_tuning && t_tuning->col_tile > 0)
        if (int f = tile_if_fits(t_tuning->col_tile, L)) return f;
    static const bool any = getenv("RK_COL_TILE") != nullptr;
    if (any)
        if (int f = forced_tile("RK_COL_TILE", L)) return f;
    int T = threads_per_instance(L);
    int C = std::max(16, (default_tile_c_row(L) / 16) * 16);
    while (C > 1 && ((int64_t)C * T > 1024 || (int64_t)C * lds_pitch_host(L) * t_elem_bytes > 159 * 1024)) C /= 2;
    // tuning override: an upper bound on the column tile width (kernels outside
    // the catalog come from the run-time compiler)
    static const int cap = [] {
        const char* e = getenv("RK_COL_TILE_MAX");
        return e && *e ? atoi(e) : 0;
    }();
    if (cap > 0 && C > cap) C = cap;
    return C;
}

// ---------------------------------------------------------------------------
// FP64 host math for the tables
// ---------------------------------------------------------------------------
using cd = std::complex<double>;

// exp(-2 pi i e / n), e reduced exactly modulo n first
static cd unit_root(int64_t e, int64_t n) {
    e %= n;
    if (e < 0) e += n;
    // use the symmetric argument (-n/2, n/2] for a smaller FP64 argument
    int64_t s = (2 * e > n) ? e - n : e;
    double ang = -2.0 * M_PI * (double)s / (double)n;
    return cd(std::cos(ang), std::sin(ang));
}

// One Stockham pass of radix R over n points (NS = product of the earlier
// radices): butterfly j = g*NS + k reads src[j + r*n/R], multiplies by
// W_{NS R}^{r k}, and writes its outputs to dst[g*NS*R + k + q*NS].
template <int R, class Root>
static void host_stage(const cd* src, cd* dst, int64_t n, int64_t ns, const Root& root) {
    // products written out (std::complex's operator* checks for infinities)
    auto mul = [](cd a, cd b) {
        return cd(a.real() * b.real() - a.imag() * b.imag(), a.real() * b.imag() + a.imag() * b.real());
    };
    const int64_t nb = n / R, tw_step = n / (ns * R);
    cd wr[R];   // W_R^q
    for (int q = 0; q < R; ++q) wr[q] = unit_root(q, R);
    for (int64_t g = 0; g < nb / ns; ++g) {
        for (int64_t k = 0; k < ns; ++k) {
            const int64_t j = g * ns + k;
            cd v[R], o[R];
            v[0] = src[j];
            for (int r = 1; r < R; ++r) v[r] = k ? mul(src[j + r * nb], root(r * k * tw_step)) : src[j + r * nb];
            if constexpr (R == 4) {
                const cd a = v[0] + v[2], b = v[0] - v[2], c = v[1] + v[3];
                const cd d(v[1].imag() - v[3].imag(), v[3].real() - v[1].real());   // -i (v1 - v3)
                o[0] = a + c; o[1] = b + d; o[2] = a - c; o[3] = b - d;
            } else if constexpr (R == 2) {
                o[0] = v[0] + v[1];
                o[1] = v[0] - v[1];
            } else {
                for (int q = 0; q < R; ++q) {
                    cd acc = v[0];
                    for (int r = 1; r < R; ++r) acc += mul(v[r], wr[(r * q) % R]);
                    o[q] = acc;
                }
            }
            const int64_t base = g * ns * R + k;
            for (int q = 0; q < R; ++q) dst[base + q * ns] = o[q];
        }
    }
}

// Forward DFT in FP64 of a length n = 2^a 3^b 5^c 7^d, used for the
// Bluestein kernel spectrum: iterative Stockham passes of radix 4, 2, 3, 5
// and 7 between two buffers (natural-order output, no bit reversal), with
// the twiddles W_n^e formed from two tables of exactly reduced roots
// (hi[e >> s] times lo[e & (2^s - 1)], 2^s >= sqrt(n)).
static void fft_host(std::vector<cd>& x) {
    const int64_t n = (int64_t)x.size();
    if (n <= 1) return;
    std::vector<int> radices;
    int64_t m = n;
    while (m % 4 == 0) { radices.push_back(4); m /= 4; }
    while (m % 2 == 0) { radices.push_back(2); m /= 2; }
    for (int p : {3, 5, 7})
        while (m % p == 0) { radices.push_back(p); m /= p; }
    if (m != 1) {   // not smooth: direct O(n^2) (the planner never asks for it)
        std::vector<cd> y(n);
        for (int64_t k = 0; k < n; ++k) {
            cd acc = 0;
            for (int64_t j = 0; j < n; ++j) acc += x[j] * unit_root(j * k % n, n);
            y[k] = acc;
        }
        x.swap(y);
        return;
    }
    int sh = 0;
    while ((int64_t(1) << (2 * sh)) < n) ++sh;
    const int64_t lo_n = int64_t(1) << sh, hi_n = (n + lo_n - 1) / lo_n;
    std::vector<cd> lo(lo_n), hi(hi_n);
    for (int64_t e = 0; e < lo_n; ++e) lo[e] = unit_root(e, n);
    for (int64_t e = 0; e < hi_n; ++e) hi[e] = unit_root(e * lo_n, n);
    auto root = [&](int64_t e) {   // W_n^e, e < n
        const cd a = hi[e >> sh], b = lo[e & (lo_n - 1)];
        return cd(a.real() * b.real() - a.imag() * b.imag(), a.real() * b.imag() + a.imag() * b.real());
    };
    std::vector<cd> y(n);
    cd* src = x.data();
    cd* dst = y.data();
    int64_t ns = 1;
    for (const int R : radices) {
        switch (R) {
            case 4: host_stage<4>(src, dst, n, ns, root); break;
            case 2: host_stage<2>(src, dst, n, ns, root); break;
            case 3: host_stage<3>(src, dst, n, ns, root); break;
            case 5: host_stage<5>(src, dst, n, ns, root); break;
            default: host_stage<7>(src, dst, n, ns, root); break;
        }
        std::swap(src, dst);
        ns *= R;
    }
    if (src != x.data()) std::copy(src, src + n, x.data());
}

// ---------------------------------------------------------------------------
// plan structures
// ---------------------------------------------------------------------------
Layout contiguous_layout(int ndim, const int64_t* shape) {
    Layout l;
    l.ndim = ndim;
    l.offset = 0;
    int64_t s = 1;
    for (int d = ndim - 1; d >= 0; --d) {
        l.shape[d] = shape[d];
        l.strides[d] = s;
        s *= shape[d];
    }
    return l;
}

}  // namespace rk

struct rk_fft_plan {
    rk_fft_desc desc;
    std::vector<rk::PassPlan> passes;
    void* tables = nullptr;
    size_t ws_bytes = 0, ws1_offset = 0;
    double norm = 1.0;
    bool trivial = false;
    int es = 8;   // element bytes: 8 complex64, 16 complex128
    rk::CopyArgs copy;
    bool timing = false;
    std::vector<hipEvent_t> events;   // 2 per pass when timing
    int device = 0;
    // Chunked execution of a multi-pass plan over an untransformed dim 0:
    // every pass runs on `chunk_outer` indices of dim 0 before the next chunk
    // starts, and the workspace holds one chunk, so a chunk's intermediate
    // can stay in the Infinity Cache between the passes.  0 = off.
    int64_t chunk_outer = 0;
    int64_t n_outer = 1;
};

namespace rk {

// Bytes from a buffer's start that a layout (shape, strides and offset in
// elements) reaches: one past its farthest element, or -1 when an element
// would sit before the buffer's start (a negative stride or offset).
static int64_t layout_extent_bytes(const rk_fft_desc& d, const int64_t* strides, int64_t offset, int es) {
    int64_t lo = offset, hi = offset;
    for (int k = 0; k < d.ndim; ++k) {
        const int64_t span = (d.shape[k] - 1) * strides[k];
        if (span < 0) lo += span; else hi += span;
    }
    if (lo < 0) return -1;
    return (hi + 1) * (int64_t)es;
}

}  // namespace rk

namespace rk {

// ---------------------------------------------------------------------------
// planning helpers
// ---------------------------------------------------------------------------

// Split a smooth n into pass lengths, each <= lim.factor_max and accepted by
// lim.length_ok, as balanced as possible.
static bool split_factors(int64_t n, const ChainLimits& lim, std::vector<int64_t>& out) {
    out.clear();
    const int64_t fmax = lim.factor_max;
    if (n <= fmax && length_ok(lim, n)) {
        out.push_back(n);
        return true;
    }
    // the fewest passes whose lengths can all stay <= fmax
    for (int p = 2; p <= 8; ++p) {
        double cap = 1.0;
        for (int k = 0; k < p; ++k) cap *= (double)fmax;
        if (cap < (double)n) continue;
        // a requested first pass length that divides n (else the balanced split)
        const int64_t ff = lim.first_factor;
        if (ff > 1 && ff <= fmax && n % ff == 0 && length_ok(lim, ff)) {
            out.push_back(ff);
            if (split_into(n / ff, p - 1, lim, out)) return true;
            out.clear();
        }
        if (split_into(n, p, lim, out)) return true;
        out.clear();
    }
    return false;
}

// Smallest M >= 2n-1 of the form 2^k, 3*2^k or 5*2^k (2^k only with pow2).
static int64_t bluestein_length(int64_t n, bool pow2 = false) {
    int64_t need = 2 * n - 1, best = 0;
    for (int64_t base : {1, 3, 5}) {
        if (pow2 && base != 1) continue;
        int64_t m = base;
        while (m < need) m *= 2;
        if (!best || m < best) best = m;
    }
    return best;
}

// The passes of one axis of length n.
int axis_passes(int64_t n, uint32_t flags, const ChainLimits& lim, std::vector<AxisPass>& out) {
    out.clear();
    const int64_t kSingle = lim.single_max;
    const bool smooth = is_smooth(n) && !(flags & RK_FFT_FORCE_BLUESTEIN);
    auto stockham_chain = [&](int64_t N, const std::vector<int64_t>& f, int64_t in_len, int64_t out_len,
                              uint32_t first_pre, uint32_t last_post, int64_t valid_in, int64_t valid_out) {
        int64_t ns = 1;
        for (size_t p = 0; p < f.size(); ++p) {
            AxisPass ap;
            ap.L = (int)f[p];
            ap.Nax = N;
            ap.NS = ns;
            ap.H = N / (ns * f[p]);
            ap.blue_single = false;
            ap.pre_ops = (p == 0 ? first_pre : 0u) | (ns > 1 ? RK_PRE_TWIDDLE : 0u);
            ap.post_ops = (p + 1 == f.size()) ? last_post : 0u;
            ap.n_valid_in = (p == 0) ? valid_in : N;
            ap.n_valid_out = (p + 1 == f.size()) ? valid_out : N;
            ap.in_len = (p == 0) ? in_len : N;
            ap.out_len = (p + 1 == f.size()) ? out_len : N;
            out.push_back(ap);
            ns *= f[p];
        }
    };
    std::vector<int64_t> f;
    const bool single_ok = n <= kSingle && !(flags & RK_FFT_FORCE_MULTIPASS) && length_ok(lim, n);
    if (smooth && (single_ok || split_factors(n, lim, f))) {
        if (single_ok) f.assign(1, n);
        if ((flags & RK_FFT_FORCE_MULTIPASS) && f.size() == 1 && n > 1) {
            // split a single-pass length into two passes for testing the chain
            int64_t d = 0;
            for (int64_t k = 2; k * k <= n; ++k)
                if (n % k == 0) d = k;
            if (d) f = {d, n / d};
        }
        stockham_chain(n, f, n, n, 0u, 0u, n, n);
        return RK_OK;
    }
    if (smooth && !lim.length_ok)
        return fail(RK_ERR_UNSUPPORTED, "cannot split axis length %lld into passes", (long long)n);
    const int64_t m = bluestein_length(n, lim.pow2_bluestein);
    const bool multi = m > lim.blue_single_max || (flags & RK_FFT_FORCE_MULTIPASS);
    if (!multi) {
        AxisPass ap;
        ap.L = (int)m;
        ap.Nax = m;
        ap.NS = 1;
        ap.H = 1;
        ap.blue_single = true;
        ap.pre_ops = RK_PRE_CHIRP;
        ap.post_ops = RK_POST_CONJ | RK_POST_CHIRP;
        ap.n_valid_in = n;
        ap.n_valid_out = n;
        ap.in_len = n;
        ap.out_len = n;
        out.push_back(ap);
        return RK_OK;
    }
    if (m <= kSingle && length_ok(lim, m) && !(flags & RK_FFT_FORCE_MULTIPASS)) {
        f.assign(1, m);   // two single-pass chains
    } else if (!split_factors(m, lim, f)) {
        return fail(RK_ERR_UNSUPPORTED, "cannot split Bluestein length %lld", (long long)m);
    }
    if (f.size() == 1 && (flags & RK_FFT_FORCE_MULTIPASS)) {   // forced multi-pass of a short M
        int64_t d = 0;
        for (int64_t k = 2; k * k <= m; ++k)
            if (m % k == 0) d = k;
        if (d) f = {d, m / d};
    }
    // forward M-point transform of chirp * x (zero padded), then multiply by
    // the kernel spectrum and conjugate ...
    stockham_chain(m, f, n, m, RK_PRE_CHIRP, RK_POST_KMUL_CONJ, n, m);
    // ... forward transform again, conjugate, chirp, truncate to n.
    std::vector<AxisPass> second;
    std::swap(second, out);
    stockham_chain(m, f, m, n, 0u, RK_POST_CONJ | RK_POST_CHIRP, m, n);
    second.insert(second.end(), out.begin(), out.end());
    out.swap(second);
    return RK_OK;
}

// Fold the batch dims (all dims but `axis`) of the src and dst layouts into
// O (outer, before the axis) and I (inner, after the axis) groups the kernel
// walks with one stride each; what does not fold becomes host loops.
static void fold_batch(const Layout& src, const Layout& dst, int axis, int64_t& O, int64_t& so_in, int64_t& so_out,
                       int64_t& I, int64_t& si_in, int64_t& si_out, std::vector<HostLoop>& loops) {
    auto fold = [&](int d0, int d1, std::vector<HostLoop>& groups) {
        // d0..d1-1, outermost first; returns groups innermost last
        for (int d = d0; d < d1; ++d) {
            if (src.shape[d] == 1) continue;
            HostLoop g{src.shape[d], src.strides[d], dst.strides[d]};
            if (!groups.empty()) {
                HostLoop& prev = groups.back();
                if (prev.in_str == g.in_str * g.n && prev.out_str == g.out_str * g.n) {
                    prev.n *= g.n;
                    prev.in_str = g.in_str;
                    prev.out_str = g.out_str;
                    continue;
                }
            }
            groups.push_back(g);
        }
    };
    std::vector<HostLoop> outer, inner;
    fold(0, axis, outer);
    fold(axis + 1, src.ndim, inner);
    O = 1; so_in = 0; so_out = 0;
    I = 1; si_in = 0; si_out = 0;
    if (!inner.empty()) {
        I = inner.back().n;
        si_in = inner.back().in_str;
        si_out = inner.back().out_str;
        inner.pop_back();
    }
    if (!outer.empty()) {
        O = outer.back().n;
        so_in = outer.back().in_str;
        so_out = outer.back().out_str;
        outer.pop_back();
    }
    loops.clear();
    for (auto& g : outer) loops.push_back(g);
    for (auto& g : inner) loops.push_back(g);
}

// x / d = (mulhi(x, m) + x) >> s for every x < 2^31, with s = ceil(log2 d)
// and the round-up multiplier m = floor(2^(32+s) / d) - 2^32 + 1 (< 2^32);
// mulhi(x, m) <= x keeps the sum below 2^32.
static void div_magic(rk_u64 d, rk_u32& m, rk_u32& s) {
    rk_u32 sh = 0;
    while ((rk_u64(1) << sh) < d) ++sh;
    m = (rk_u32)(((unsigned __int128)1 << (32 + sh)) / d - ((unsigned __int128)1 << 32) + 1);
    s = sh;
}

void set_div_magic(rk_pass_args& a) {
    a.div_magic = 0;
    if (a.n_inst >= (rk_u64(1) << 31)) return;   // the kernels divide in 64 bits
    div_magic(a.dim_I ? a.dim_I : 1, a.mag_I, a.sh_I);
    div_magic(a.dim_NS ? a.dim_NS : 1, a.mag_NS, a.sh_NS);
    div_magic(a.dim_H ? a.dim_H : 1, a.mag_H, a.sh_H);
    a.div_magic = 1;
}

int describe_pass(const AxisPass& ap, const Layout& ls, const Layout& ld, int axis, TileWidthFn tile_width,
                  PassPlan& pp) {
    memset(&pp.args, 0, sizeof(pp.args));
    pp.es = t_elem_bytes;
    pp.axis = axis;
    pp.L = ap.L;
    pp.kind = ap.blue_single ? 1 : 0;
    if (ap.blue_single) {
        pp.variant = VAR_BLUE;
    } else {
        int pre = (ap.pre_ops & RK_PRE_TWIDDLE) ? 1 : (ap.pre_ops & RK_PRE_CHIRP) ? 2 : 0;
        int post = (ap.post_ops & RK_POST_KMUL_CONJ) ? 1 : (ap.post_ops & RK_POST_CHIRP) ? 2 : 0;
        if ((ap.pre_ops & RK_PRE_TWIDDLE) && (ap.pre_ops & RK_PRE_CHIRP))
            return fail(RK_ERR_UNSUPPORTED, "internal: pass with both chirp and twiddle pre-ops");
        pp.variant = pre | (post << 2);
    }
    // the kernels bound positions one by one only on chirp sides (pad_in /
    // pad_out of fft_device.hpp): every other side covers its whole axis
    if ((!(ap.pre_ops & RK_PRE_CHIRP) && ap.n_valid_in != ap.Nax) ||
        (!(ap.post_ops & RK_POST_CHIRP) && ap.n_valid_out != ap.Nax))
        return fail(RK_ERR_UNSUPPORTED, "internal: a padded pass side without a chirp op (L=%d)", ap.L);
    rk_pass_args& a = pp.args;
    int64_t O, so_in, so_out, I, si_in, si_out;
    fold_batch(ls, ld, axis, O, so_in, so_out, I, si_in, si_out, pp.loops);
    a.dim_I = (rk_u64)I;
    a.dim_NS = (rk_u64)ap.NS;
    a.dim_H = (rk_u64)ap.H;
    a.n_inst = (rk_u64)(O * ap.H * ap.NS * I);
    a.in_off = ls.offset;
    a.out_off = ld.offset;
    a.in_so = so_in;
    a.out_so = so_out;
    a.in_si = si_in;
    a.out_si = si_out;
    a.in_sa = ls.strides[axis];
    a.out_sa = ld.strides[axis];
    a.N = (rk_u64)ap.Nax;
    a.pos_mult = (rk_u64)(ap.Nax / ap.L);
    a.tw_mult = (rk_u64)(ap.Nax / (ap.NS * ap.L));
    a.pre_ops = ap.pre_ops;
    a.post_ops = ap.post_ops;
    a.n_valid_in = (rk_u64)ap.n_valid_in;
    a.n_valid_out = (rk_u64)ap.n_valid_out;
    a.in_sign = 1.0f;
    a.out_sign = 1.0f;
    a.out_scale = 1.0f;
    a.out_scale64 = 1.0;

    // global-memory walk of each side
    const int64_t L = ap.L;
    const int64_t pos_in = (int64_t)a.pos_mult * a.in_sa;
    const int64_t inst_in = I > 1 ? si_in : (ap.NS > 1 || ap.H > 1) ? a.in_sa : so_in;
    const bool flat_in = ap.H == 1 && ap.NS == 1 && I == 1 && a.in_sa == 1 && (O == 1 || so_in == L) &&
                         ap.pre_ops == 0 && ap.n_valid_in == ap.Nax && pp.loops.empty();
    const int64_t pos_out = ap.NS * a.out_sa;
    const int64_t inst_out = I > 1 ? si_out : ap.NS > 1 ? a.out_sa : ap.H > 1 ? L * a.out_sa : so_out;
    const bool flat_out = ap.NS == 1 && I == 1 && a.out_sa == 1 &&
                          (ap.H == 1 ? (O == 1 || so_out == L) : (O == 1 || so_out == ap.Nax)) &&
                          ap.post_ops == 0 && ap.n_valid_out == ap.Nax && pp.loops.empty();
    // A side is column-like when adjacent instances are adjacent in
    // memory and its elements are not; the compute mapping follows
    // the column-like side(s) so they are read/written straight from
    // registers, and a row-like side under that mapping is staged
    // through LDS.
    const bool col_in = pos_in != 1 && inst_in == 1;
    const bool col_out = pos_out != 1 && inst_out == 1;
    const int T = threads_per_instance(ap.L);
    a.cmap = (col_in || col_out) ? 1u : 0u;
    auto side_mode = [&](bool col, bool flat, int64_t pos) -> rk_u32 {
        if (a.cmap) {
            if (col) return RK_MODE_DIRECT;
            if (flat) return RK_MODE_FLAT;
            return pos == 1 ? RK_MODE_ROW : RK_MODE_DIRECT;
        }
        if (flat && T < 32) return RK_MODE_FLAT;   // short rows: stream the tile linearly
        return RK_MODE_DIRECT;
    };
    a.load_mode = side_mode(col_in, flat_in, pos_in);
    a.store_mode = side_mode(col_out, flat_out, pos_out);
    pp.C = tile_width(ap.L, a.cmap != 0);
    a.n_tiles = (a.n_inst + (rk_u64)pp.C - 1) / (rk_u64)pp.C;
    set_div_magic(a);
    pp.inst_per_outer = (uint64_t)(ap.H * ap.NS * I);
    pp.outer_count = O;
    return RK_OK;
}

static int plan_tile_width(int L, bool colmap) { return colmap ? tile_c_col(L) : tile_c_row(L); }

// XCDs of the current device (the dispatcher hands work-group b to XCD
// b % count); 8 on MI355X when there is no device to ask.
static rk_u32 xcd_count() {
    int dev = 0, n = 0;
    if (hipGetDevice(&dev) == hipSuccess && hipDeviceGetAttribute(&n, hipDeviceAttributeNumberOfXccs, dev) == hipSuccess &&
        n > 0)
        return (rk_u32)n;
    (void)hipGetLastError();
    return 8;
}

// Column tiles narrower than 16 instances share 128-byte lines with their
// neighbours: each group of 32 / C consecutive tiles is sent to one XCD, so
// the neighbours meet in that XCD's L2 instead of each fetching whole lines
// from HBM (RK_XCD_GROUP overrides the group size; 1 turns it off).  Measured
// on 64 x 2^20 as two 1024-point passes with C = 4: 953 us without groups,
// 715 us with groups of 4, 669 us with groups of 8.  Tiles of 16 instances
// go in pairs (two tiles: 256 contiguous bytes of each row): 1024 x 2^16 (two
// 256-point passes) 0.326 against 0.310 ms, 64 x 2^20 (two 1024-point
// passes) 0.454 against 0.448 ms.
static void set_xcd_group(PassPlan& pp) {
    static const int forced = [] {
        const char* e = getenv("RK_XCD_GROUP");
        return e && *e ? atoi(e) : 0;
    }();
    rk_u32 g = 1;
    if (forced > 0)
        g = (rk_u32)forced;
    else if ((pp.args.cmap || pp.args.load_mode == RK_MODE_COLUMN || pp.args.store_mode == RK_MODE_COLUMN) && pp.C <= 16)
        g = (rk_u32)(32 / pp.C);
    if (g > 1) {
        pp.args.xcd_group = g;
        pp.args.n_xcd = xcd_count();
    }
}

// Bytes above which a plain or twiddle pass uses the streaming kernel: the
// data it reads and writes (one buffer when in place) then exceeds the
// 256 MB Infinity Cache, so nothing it touches would be read from a cache
// again (RK_STREAM_MIN_BYTES overrides; 0 turns streaming off).
static double stream_min_bytes() {
    static const double v = [] {
        const char* e = getenv("RK_STREAM_MIN_BYTES");
        return e && *e ? atof(e) : 256.0 * 1024 * 1024;
    }();
    return v;
}

static bool stream_pass(const PassPlan& pp) {
    if (pp.variant != VAR_PLAIN && pp.variant != VAR_TWIDDLE) return false;
    if (stream_min_bytes() <= 0) return false;
    const double bytes = (double)pp.es * (double)pp.args.n_inst * (double)pp.L * (pp.src == pp.dst ? 1.0 : 2.0);
    return bytes > stream_min_bytes();
}

// Kernels with the pass's load/store modes compiled in (RK_FIXED_MODES=0
// turns them off).
static bool use_fixed_modes() {
    if (t_tuning && t_tuning->fixed_modes >= 0) return t_tuning->fixed_modes != 0;
    static const bool v = [] {
        const char* e = getenv("RK_FIXED_MODES");
        return !(e && *e && atoi(e) == 0);
    }();
    return v;
}

// Catalogue bit of the twin with this pass's load/store modes compiled in
// (0: the pass's modes have no such twins).
static int fixed_mode_bit(const PassPlan& pp) {
    if (!use_fixed_modes()) return 0;
    const rk_u32 lm = pp.args.load_mode, sm = pp.args.store_mode;
    if (lm == RK_MODE_DIRECT && sm == RK_MODE_DIRECT) return VAR_MODE_DD;
    if (lm == RK_MODE_DIRECT && sm == RK_MODE_FLAT) return VAR_MODE_DF;
    if (lm == RK_MODE_FLAT && sm == RK_MODE_FLAT) return VAR_MODE_FF;
    return 0;
}

// Variant bits of a pass's kernel beyond its Variant: the IO policy and the
// wave-local stages.
static int io_variant_bits(const PassPlan& pp) {
    int v = 0;
    if (pp.nt_load && pp.nt_store)
        v |= VAR_STREAM;
    else if (pp.nt_load)
        v |= VAR_NT_LOAD;
    else if (pp.nt_store)
        v |= VAR_NT_STORE;
    if (pp.wave) v |= VAR_WAVE;
    if (pp.pipe) v |= VAR_PIPE;
    if (pp.half) v |= VAR_HALF;
    if (pp.occ) v |= VAR_OCC;
    if (pp.es == 16) v |= VAR_F64;
    return v;
}

static const KernelEntry* catalog_find(int L, int C, int variant) {
    for (const KernelEntry& e : catalog())
        if (e.L == L && e.C == C && e.variant == variant) return &e;
    return nullptr;
}

// The pass's kernel: the catalogue's, with the pass's load/store modes
// compiled in when such a twin exists, else from the run-time compiler (the
// same entry, any variant number).  Without run-time compilation, a pass
// whose nontemporal twin is missing keeps the plain accesses.
static int find_kernel(PassPlan& pp, uint32_t flags) {
    pp.fixed_modes = false;
    const int mbit = fixed_mode_bit(pp);
    auto take = [&](const KernelEntry* e, bool fixed) {
        pp.aot = e->func;
        pp.threads = e->threads;
        pp.lds_bytes = e->lds_bytes;
        pp.fixed_modes = fixed;
    };
    const int want = pp.variant | io_variant_bits(pp);
    if (mbit)
        if (const KernelEntry* e = catalog_find(pp.L, pp.C, want | mbit)) return take(e, true), RK_OK;
    if (const KernelEntry* e = catalog_find(pp.L, pp.C, want)) return take(e, false), RK_OK;
    if (flags & RK_FFT_NO_JIT) {
        if (pp.nt_load || pp.nt_store) {
            pp.nt_load = pp.nt_store = false;
            return find_kernel(pp, flags);
        }
        return fail(RK_ERR_NO_KERNEL, "no precompiled kernel for L=%d C=%d variant=%d and run-time compilation is off",
                    pp.L, pp.C, want);
    }
    pp.fixed_modes = mbit != 0;
    if (flags & RK_FFT_DRY_RUN) {   // describe the kernel the run-time compiler would build
        pp.threads = pp.C * threads_per_instance(pp.L);
        pp.lds_bytes = tile_lds_bytes_host(pp.L, pp.C, pp.es);
        return RK_OK;
    }
    return jit_pass_kernel(pp.L, pp.C, want | mbit, &pp.jit, &pp.threads, &pp.lds_bytes);
}

// Passes whose instances fill whole waves (T = L / R0 divides 64) run the
// row mapping with wave-local stage exchanges: a column-like side is then
// staged through LDS with lanes across instances (COLUMN), a contiguous
// side read by each instance's own lanes (DIRECT).  RK_WAVE=1 turns it on.
static bool wave_enabled() {
    static const bool v = [] {
        const char* e = getenv("RK_WAVE");
        return e && *e && atoi(e) != 0;
    }();
    return v;
}

// Passes whose loads are DIRECT can run the persistent, software-pipelined
// kernel (each work-group loads its next tile while it computes the current
// one).  RK_PIPE=1 turns it on.
static bool pipe_enabled() {
    static const bool v = [] {
        const char* e = getenv("RK_PIPE");
        return e && *e && atoi(e) != 0;
    }();
    return v;
}

// Work-groups of a persistent pass: one per resident slot of every CU, the
// slots bounded by the tile's LDS and by 32 waves per CU with registers
// bounded for that residency (occ), else one per CU.
static uint64_t pipe_grid(const PassPlan& pp) {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0) {
        (void)hipGetLastError();
        cus = 256;
    }
    const int waves = std::max(1, (pp.threads + 63) / 64);
    const int by_lds = pp.lds_bytes > 0 ? (160 * 1024) / pp.lds_bytes : 32;
    const int slots = pp.occ ? std::max(1, std::min(by_lds, 32 / waves)) : 1;
    return (uint64_t)cus * (uint64_t)slots;
}

// Tiles exchanged through re/im halves even when they fit whole: RK_HALF=1.
static bool half_enabled() {
    static const bool v = [] {
        const char* e = getenv("RK_HALF");
        return e && *e && atoi(e) != 0;
    }();
    return v;
}

// Kernels whose launch bounds ask for as many resident work-groups per CU as
// the tile's LDS allows.  By default: complex64 plain and twiddle passes of
// work-groups of 1024 threads (16 waves) of at most 1024 points with
// DIRECT/DIRECT or DIRECT/FLAT compiled in, the 16 x 1024 column tiles of
// two-pass chains: two work-groups per CU in 64 VGPRs without spills (their
// generic-mode kernels spill 47-55 registers under the same bound).
// Measured on 64 x 2^20: 0.463 ms with the bound, 0.530 without.  RK_OCC=n
// instead bounds every pass of at least n points (0: none).
static int occ_min_length() {
    static const int v = [] {
        const char* e = getenv("RK_OCC");
        return e && *e ? atoi(e) : -1;
    }();
    return v;
}

static bool occ_default(const PassPlan& pp) {
    const int mbit = fixed_mode_bit(pp);
    return pp.es == 8 && (pp.variant == VAR_PLAIN || pp.variant == VAR_TWIDDLE) && pp.L <= 1024 &&
           (int64_t)pp.C * threads_per_instance(pp.L) == 1024 && (mbit == VAR_MODE_DD || mbit == VAR_MODE_DF);
}

static void choose_wave(PassPlan& pp, uint32_t flags) {
    if (!wave_enabled()) return;
    const int T = threads_per_instance(pp.L);
    if (T > 64 || 64 % T != 0) return;
    if (flags & RK_FFT_NO_JIT) {
        bool have = false;
        for (const KernelEntry& e : catalog())
            have = have || (e.L == pp.L && e.C == pp.C && (e.variant & ~(VAR_STREAM | VAR_NT_LOAD | VAR_NT_STORE)) ==
                                                             (pp.variant | VAR_WAVE));
        if (!have) return;
    }
    rk_pass_args& a = pp.args;
    if (a.cmap) {
        auto to_rows = [](rk_u32 m) -> rk_u32 {
            return m == RK_MODE_DIRECT ? RK_MODE_COLUMN : m == RK_MODE_ROW ? RK_MODE_DIRECT : m;
        };
        a.load_mode = to_rows(a.load_mode);
        a.store_mode = to_rows(a.store_mode);
        a.cmap = 0;
    }
    pp.wave = true;
}

// A table of complex values in the pass's element type: FP64 values rounded
// once to complex64, or kept as complex128.
static int64_t add_table(TableBuilder& tb, const PassPlan& pp, const std::string& key, const std::vector<cd>& v) {
    if (pp.es == 16) {
        std::vector<double2> d(v.size());
        for (size_t k = 0; k < v.size(); ++k) d[k] = make_double2(v[k].real(), v[k].imag());
        return tb.add(key + "/c128", d.data(), d.size() * sizeof(double2));
    }
    std::vector<float2> f(v.size());
    for (size_t k = 0; k < v.size(); ++k) f[k] = make_float2((float)v[k].real(), (float)v[k].imag());
    return tb.add(key, f.data(), f.size() * sizeof(float2));
}

int build_tables(TableBuilder& tb, PassPlan& pp, const AxisPass& ap, int64_t nreal) {
    // in-pass stage twiddles W_L^e
    {
        std::vector<cd> w(pp.L);
        for (int e = 0; e < pp.L; ++e) w[e] = unit_root(e, pp.L);
        pp.twl_off = add_table(tb, pp, "twl" + std::to_string(pp.L), w);
    }
    if (ap.pre_ops & RK_PRE_TWIDDLE) {
        int64_t N = ap.Nax;
        int shift = 0;
        while ((int64_t(1) << (2 * shift)) < N) ++shift;   // 2^shift >= sqrt(N)
        const int64_t lo_n = int64_t(1) << shift;
        const int64_t hi_n = (N + lo_n - 1) / lo_n;
        std::vector<double2> hi(hi_n), lo(lo_n);
        for (int64_t k = 0; k < hi_n; ++k) {
            cd z = unit_root(k * lo_n, N);
            hi[k] = make_double2(z.real(), z.imag());
        }
        for (int64_t k = 0; k < lo_n; ++k) {
            cd z = unit_root(k, N);
            lo[k] = make_double2(z.real(), z.imag());
        }
        pp.twhi_off = tb.add("twhi" + std::to_string(N), hi.data(), hi.size() * sizeof(double2));
        pp.twlo_off = tb.add("twlo" + std::to_string(N), lo.data(), lo.size() * sizeof(double2));
        pp.args.tw_shift = (rk_u32)shift;
        pp.args.tw_mask = (rk_u64)(lo_n - 1);
    }
    const bool needs_chirp = (ap.pre_ops & RK_PRE_CHIRP) || (ap.post_ops & RK_POST_CHIRP);
    if (needs_chirp) {
        // exp(-i pi (n^2 mod 2N) / N): the phase index is an exact integer
        std::vector<cd> ch(nreal);
        const int64_t two_n = 2 * nreal;
        for (int64_t k = 0; k < nreal; ++k) {
            const unsigned __int128 sq = (unsigned __int128)k * (unsigned __int128)k;
            const int64_t ph = (int64_t)(sq % (unsigned __int128)two_n);
            ch[k] = unit_root(ph, two_n);   // exp(-2 pi i ph / 2N) = exp(-i pi ph / N)
        }
        pp.chirp_off = add_table(tb, pp, "chirp" + std::to_string(nreal), ch);
    }
    const bool needs_bspec = ap.blue_single || (ap.post_ops & RK_POST_KMUL_CONJ);
    if (needs_bspec) {
        const int64_t m = ap.Nax, two_n = 2 * nreal;
        std::vector<cd> b(m, cd(0, 0));
        for (int64_t k = 0; k < nreal; ++k) {
            const unsigned __int128 sq = (unsigned __int128)k * (unsigned __int128)k;
            const int64_t ph = (int64_t)(sq % (unsigned __int128)two_n);
            cd c = std::conj(unit_root(ph, two_n));   // conj(chirp[k]) = exp(+i pi k^2 / N)
            b[k] = c;
            if (k > 0) b[m - k] = c;
        }
        fft_host(b);
        for (int64_t k = 0; k < m; ++k) b[k] /= (double)m;
        pp.bspec_off = add_table(tb, pp, "bspec" + std::to_string(nreal) + "_" + std::to_string(m), b);
    }
    return RK_OK;
}

int route_passes(const rk_fft_desc& d, const std::vector<int>& work, const Layout& lin, const Layout& lout,
                 const ChainLimits& lim, TileWidthFn tile_width, const PassHook& hook, std::vector<PassPlan>& passes,
                 size_t& ws_elems, bool& two_ws) {
    int cur = BUF_IN;
    int64_t cur_shape[RK_MAX_DIMS];
    for (int k = 0; k < d.ndim; ++k) cur_shape[k] = d.shape[k];
    for (size_t wi = 0; wi < work.size(); ++wi) {
        const int axis = work[wi];
        const int64_t n = d.shape[axis];
        std::vector<AxisPass> aps;
        ChainLimits al = lim;
        if (lim.strided_single_max > 0 && lin.strides[axis] != 1)
            al.single_max = std::min(al.single_max, lim.strided_single_max);
        int rc = axis_passes(n, d.flags, al, aps);
        if (rc != RK_OK) return rc;
        if (aps.size() >= 3) two_ws = true;
        int ws_next = BUF_WS0;
        for (size_t pi = 0; pi < aps.size(); ++pi) {
            const AxisPass& ap = aps[pi];
            const bool axis_last = (pi + 1 == aps.size());
            int dst;
            const bool plan_last = axis_last && (wi + 1 == work.size());
            if (axis_last && (plan_last || !(d.flags & RK_FFT_OUT_LAST_ONLY))) {
                dst = BUF_OUT;
            } else {
                dst = ws_next;
                if (dst == cur) dst = (dst == BUF_WS0) ? BUF_WS1 : BUF_WS0;
                ws_next = (dst == BUF_WS0) ? BUF_WS1 : BUF_WS0;
            }
            // layouts of src and dst for this pass
            int64_t s_shape[RK_MAX_DIMS], d_shape[RK_MAX_DIMS];
            for (int k = 0; k < d.ndim; ++k) s_shape[k] = d_shape[k] = cur_shape[k];
            s_shape[axis] = ap.in_len;
            d_shape[axis] = ap.out_len;
            Layout ls = (cur == BUF_IN) ? lin : (cur == BUF_OUT) ? lout : contiguous_layout(d.ndim, s_shape);
            Layout ld = (dst == BUF_OUT) ? lout : contiguous_layout(d.ndim, d_shape);
            if (dst == BUF_WS0 || dst == BUF_WS1) {
                size_t e = 1;
                for (int k = 0; k < d.ndim; ++k) e *= (size_t)d_shape[k];
                ws_elems = std::max(ws_elems, e);
            }
            PassPlan pp;
            pp.src = cur;
            pp.dst = dst;
            rc = describe_pass(ap, ls, ld, axis, tile_width, pp);
            if (rc != RK_OK) return rc;
            pp.src_stride0 = ls.strides[0];
            pp.dst_stride0 = ld.strides[0];
            rc = hook(pp, ap, n);
            if (rc != RK_OK) return rc;
            passes.push_back(pp);
            cur = dst;
            cur_shape[axis] = ap.out_len;
        }
    }
    return RK_OK;
}

static int plan_create(const rk_fft_desc* d, rk_fft_plan** out) {
    if (!d || !out) return fail(RK_ERR_INVALID, "rk_fft_plan_create: null argument");
    *out = nullptr;
    if (d->dtype != RK_C64 && d->dtype != RK_C128)
        return fail(RK_ERR_UNSUPPORTED, "dtype must be RK_C64 or RK_C128");
    // the element size every sizing decision of this plan uses (tile widths,
    // tables, workspace, chunks), for this thread and this plan only
    struct ElemScope {
        explicit ElemScope(int es) { t_elem_bytes = es; }
        ~ElemScope() { t_elem_bytes = 8; }
    } elem_scope(d->dtype == RK_C128 ? 16 : 8);
    if (d->ndim < 1 || d->ndim > RK_MAX_DIMS) return fail(RK_ERR_INVALID, "ndim must be in [1, %d]", RK_MAX_DIMS);
    for (int k = 0; k < d->ndim; ++k)
        if (d->shape[k] < 1) return fail(RK_ERR_INVALID, "shape[%d] = %lld must be >= 1", k, (long long)d->shape[k]);
    if (d->naxes < 0 || d->naxes > d->ndim) return fail(RK_ERR_INVALID, "naxes out of range");
    std::vector<int> axes;
    if (d->naxes == 0) {
        for (int k = 0; k < d->ndim; ++k) axes.push_back(k);
    } else {
        for (int k = 0; k < d->naxes; ++k) {
            int a = d->axes[k];
            if (a < 0 || a >= d->ndim) return fail(RK_ERR_INVALID, "axis %d out of range for ndim %d", a, d->ndim);
            if (std::find(axes.begin(), axes.end(), a) != axes.end())
                return fail(RK_ERR_INVALID, "axis %d repeated", a);
            axes.push_back(a);
        }
    }

    std::unique_ptr<rk_fft_plan> plan(new rk_fft_plan());
    plan->desc = *d;
    plan->es = t_elem_bytes;
    hipGetDevice(&plan->device);

    Layout lin, lout;
    lin.ndim = lout.ndim = d->ndim;
    for (int k = 0; k < d->ndim; ++k) {
        lin.shape[k] = lout.shape[k] = d->shape[k];
        lin.strides[k] = d->in_strides[k];
        lout.strides[k] = d->out_strides[k];
    }
    lin.offset = d->in_offset;
    lout.offset = d->out_offset;

    // axes with length > 1, innermost first (as the reference orders kernels)
    std::vector<int> work;
    for (auto it = axes.rbegin(); it != axes.rend(); ++it)
        if (d->shape[*it] > 1) work.push_back(*it);

    double norm = 1.0;
    for (int a : work) norm /= (double)d->shape[a];
    plan->norm = norm;

    if (work.empty()) {
        plan->trivial = true;
        CopyArgs& c = plan->copy;
        memset(&c, 0, sizeof(c));
        c.words = plan->es / 8;
        c.ndim = d->ndim;
        c.total = 1;
        for (int k = 0; k < d->ndim; ++k) {
            c.shape[k] = d->shape[k];
            c.in_str[k] = d->in_strides[k];
            c.out_str[k] = d->out_strides[k];
            c.total *= (rk_u64)d->shape[k];
        }
        c.in_off = d->in_offset;
        c.out_off = d->out_offset;
        *out = plan.release();
        return RK_OK;
    }

    TableBuilder tb;
    size_t ws_elems = 0;
    bool two_ws = false;
    PassHook hook = [&](PassPlan& pp, const AxisPass& ap, int64_t n) -> int {
        choose_wave(pp, d->flags);
        return build_tables(tb, pp, ap, n);
    };
    int rc = route_passes(*d, work, lin, lout, default_limits(), plan_tile_width, hook, plan->passes, ws_elems, two_ws);
    if (rc != RK_OK) return rc;
    plan->passes.front().is_first = true;
    plan->passes.back().is_last = true;

    // chunked execution (see rk_fft_plan::chunk_outer): over dim 0 when it is
    // a batch dimension and every pass walks it as the outermost factor of
    // its instance index (no host loops), so that the instances of batch
    // index b are a contiguous range of whole tiles in every pass
    if (plan->passes.size() >= 2 && d->shape[0] > 1 && std::find(work.begin(), work.end(), 0) == work.end()) {
        const int64_t B = d->shape[0];
        bool ok = true;
        int64_t per_row_elems = 1;   // largest src/dst elements per dim-0 index
        for (auto& p : plan->passes) {
            const uint64_t per0 = p.args.n_inst / (uint64_t)B;
            ok = ok && p.loops.empty() && p.args.n_inst % (uint64_t)B == 0 && per0 % (uint64_t)p.C == 0;
            p.inst_per_outer = per0;   // instances per dim-0 index
            per_row_elems = std::max(per_row_elems, std::max(p.src_stride0, p.dst_stride0));
        }
        // Chains whose passes would otherwise stream the whole array through
        // HBM (more than the 256 MB Infinity Cache) run in chunks whose
        // intermediates fill that cache: 256 MiB for a two-pass chain (one
        // workspace), 128 MiB for longer ones (two workspaces in use), with
        // only the plan's first loads and last stores nontemporal.
        // Measured: 64 x 2^20 (three passes) 0.521 ms in 128 MiB chunks
        // against 0.576 unchunked, 0.544 / 0.615 in 96 / 160 MiB chunks;
        // 4096 x 12288 (two passes) 0.284 ms in 256 MiB chunks against 0.300
        // unchunked, 0.309 / 0.324 in 128 / 64 MiB chunks.  RK_CHUNK_BYTES
        // overrides (0 turns chunking off).
        static const int64_t chunk_env = [] {
            const char* e = getenv("RK_CHUNK_BYTES");
            return e && *e ? atoll(e) : (int64_t)-1;
        }();
        const int64_t chunk_bytes = (t_tuning && t_tuning->chunk_bytes >= 0) ? t_tuning->chunk_bytes
                                    : chunk_env >= 0 ? chunk_env
                                    : (two_ws ? (int64_t)128 << 20 : (int64_t)256 << 20);
        const double pass_bytes = 2.0 * t_elem_bytes * (double)plan->passes.front().args.n_inst * plan->passes.front().L;
        if (ok && chunk_bytes > 0 && stream_min_bytes() > 0 && pass_bytes > stream_min_bytes()) {
            const int64_t chunk = std::max<int64_t>(1, chunk_bytes / (per_row_elems * (int64_t)t_elem_bytes));
            if (chunk < B) {
                plan->chunk_outer = chunk;
                plan->n_outer = B;
                // the workspace holds one chunk: its layouts are contiguous with dim 0 outermost
                if (ws_elems) ws_elems = (size_t)((int64_t)ws_elems / B * chunk);
            }
        }
    }
    // the IO policy of every pass, then its kernel.  A chunked chain streams
    // only what comes from and goes to the caller's arrays (its first pass's
    // loads, its last pass's stores): the intermediate stays in the Infinity
    // Cache between the passes of a chunk.
    for (PassPlan& pp : plan->passes) {
        const bool plain = pp.variant == VAR_PLAIN || pp.variant == VAR_TWIDDLE;
        if (plan->chunk_outer) {
            pp.nt_load = plain && pp.is_first && pp.src == BUF_IN && stream_min_bytes() > 0;
            pp.nt_store = plain && pp.is_last && pp.dst == BUF_OUT && stream_min_bytes() > 0;
        } else {
            pp.nt_load = pp.nt_store = stream_pass(pp);
        }
        pp.pipe = pipe_enabled() && pp.args.load_mode == RK_MODE_DIRECT && pp.loops.empty();
        pp.half = half_enabled() && (int64_t)pp.C * lds_pitch_host(pp.L) * pp.es <= 80 * 1024;
        pp.occ = occ_min_length() >= 0 ? occ_min_length() > 0 && pp.L >= occ_min_length() : occ_default(pp);
        rc = find_kernel(pp, d->flags);
        if (rc != RK_OK) return rc;
        set_xcd_group(pp);
        const uint64_t max_grid = (uint64_t(1) << 31) / (uint64_t)pp.threads;
        pp.grid = std::min<uint64_t>(pp.args.n_tiles, pp.pipe ? pipe_grid(pp) : max_grid);
        if (pp.pipe) pp.args.xcd_group = 1;   // the persistent body walks tiles in launch order
    }
    plan->ws_bytes = 0;
    if (ws_elems) {
        const size_t one = ((ws_elems * (size_t)t_elem_bytes) + 255) & ~size_t(255);
        plan->ws1_offset = one;
        plan->ws_bytes = two_ws ? 2 * one : one;
        // a plan whose passes use WS1 without ever needing two at once still
        // gets both slots
        for (auto& p : plan->passes)
            if (p.src == BUF_WS1 || p.dst == BUF_WS1) plan->ws_bytes = 2 * one;
    }
    if (!tb.bytes.empty() && !(d->flags & RK_FFT_DRY_RUN)) {
        RK_HIP(hipMalloc(&plan->tables, tb.bytes.size()));
        RK_HIP(hipMemcpy(plan->tables, tb.bytes.data(), tb.bytes.size(), hipMemcpyHostToDevice));
        // the plan executes on callers' streams, which do not wait for the
        // null stream: the tables must have landed before this returns
        RK_HIP(hipDeviceSynchronize());
    }
    *out = plan.release();
    return RK_OK;
}

// Launches tiles [t0, t1) of a pass (one tile per work-group).
static int launch_pass(const PassPlan& pp, rk_pass_args& a, hipStream_t stream, const rk_user_blob* load_blob,
                       const rk_user_blob* store_blob, rk_u64 t0, rk_u64 t1) {
    // host loops over batch dims the kernel could not fold
    const size_t nl = pp.loops.size();
    int64_t total = 1;
    for (auto& l : pp.loops) total *= l.n;
    const rk_i64 in0 = a.in_off, out0 = a.out_off;
    for (int64_t it = 0; it < total; ++it) {
        int64_t rem = it;
        a.in_off = in0;
        a.out_off = out0;
        for (size_t k = nl; k-- > 0;) {
            const int64_t idx = rem % pp.loops[k].n;
            rem /= pp.loops[k].n;
            a.in_off += idx * pp.loops[k].in_str;
            a.out_off += idx * pp.loops[k].out_str;
        }
        // grids beyond the launch limit are split; the kernel stops at tile t1
        // (a persistent kernel covers [t0, t1) in one launch)
        a.n_tiles = t1;
        for (rk_u64 base = t0; base < t1; base += pp.pipe ? (t1 - t0) : pp.grid) {
            a.tile_base = base;
            const unsigned grid = (unsigned)std::min<rk_u64>(pp.grid, t1 - base);
            // fused kernels take (pass args, load blob, store blob)
            void* params[] = {&a, const_cast<rk_user_blob*>(load_blob), const_cast<rk_user_blob*>(store_blob)};
            if (pp.jit) {
                RK_HIP(hipModuleLaunchKernel(pp.jit, grid, 1, 1, (unsigned)pp.threads, 1, 1, 0, stream, params,
                                             nullptr));
            } else {
                RK_HIP(hipLaunchKernel(pp.aot, dim3(grid), dim3((unsigned)pp.threads), params, 0, stream));
            }
        }
    }
    a.in_off = in0;
    a.out_off = out0;
    return RK_OK;
}

static int plan_execute(const rk_fft_plan* plan, void* out, const void* in, int inverse, void* ws, hipStream_t stream,
                        const void* load_blob, size_t load_bytes, const void* store_blob, size_t store_bytes) {
    if (!plan) return fail(RK_ERR_INVALID, "rk_fft_execute: null plan");
    if (!out || !in) return fail(RK_ERR_INVALID, "rk_fft_execute: null in/out pointer");
    if (plan->trivial) {
        CopyArgs c = plan->copy;
        c.in = (const float2*)in;
        c.out = (float2*)out;
        return launch_copy(c, stream);
    }
    if (plan->desc.flags & RK_FFT_DRY_RUN) return fail(RK_ERR_INVALID, "rk_fft_execute: a dry-run plan cannot execute");
    if (plan->ws_bytes && !ws) return fail(RK_ERR_INVALID, "rk_fft_execute: plan needs %zu bytes of workspace", plan->ws_bytes);
    void* bufs[4] = {const_cast<void*>(in), out, ws, ws ? (char*)ws + plan->ws1_offset : nullptr};
    const char* tables = (const char*)plan->tables;
    auto tab = [&](int64_t off) -> const void* { return off < 0 ? nullptr : tables + off; };
    rk_user_blob lb, sb;
    memset(&lb, 0, sizeof(lb));
    memset(&sb, 0, sizeof(sb));
    if (load_blob && load_bytes) memcpy(lb.b, load_blob, std::min<size_t>(load_bytes, sizeof(lb.b)));
    if (store_blob && store_bytes) memcpy(sb.b, store_blob, std::min<size_t>(store_bytes, sizeof(sb.b)));
    const size_t np = plan->passes.size();
    std::vector<rk_pass_args> args(np);
    for (size_t p = 0; p < np; ++p) {
        const PassPlan& pp = plan->passes[p];
        rk_pass_args& a = args[p];
        a = pp.args;
        a.in = (const float2*)bufs[pp.src];
        a.out = (float2*)bufs[pp.dst];
        a.twl = (const float2*)tab(pp.twl_off);
        a.tw_hi = (const double2*)tab(pp.twhi_off);
        a.tw_lo = (const double2*)tab(pp.twlo_off);
        a.chirp = (const float2*)tab(pp.chirp_off);
        a.bspec = (const float2*)tab(pp.bspec_off);
        const bool norm_forward = (plan->desc.flags & RK_FFT_NORM_FORWARD) != 0;
        if (inverse) {
            if (pp.is_first) a.in_sign = -1.0f;
            if (pp.is_last) {
                a.out_sign = -1.0f;
                if (!norm_forward) {
                    a.out_scale = (float)plan->norm;
                    a.out_scale64 = plan->norm;
                }
            }
        } else if (norm_forward && pp.is_last) {
            a.out_scale = (float)plan->norm;
            a.out_scale64 = plan->norm;
        }
    }
    if (!plan->chunk_outer) {
        for (size_t p = 0; p < np; ++p) {
            if (plan->timing) RK_HIP(hipEventRecord(plan->events[2 * p], stream));
            int rc = launch_pass(plan->passes[p], args[p], stream, &lb, &sb, 0, args[p].n_tiles);
            if (rc != RK_OK) return rc;
            if (plan->timing) RK_HIP(hipEventRecord(plan->events[2 * p + 1], stream));
        }
        return RK_OK;
    }
    // chunked: all passes over `chunk_outer` outer indices, then the next chunk;
    // the workspace side of each pass is rebased to the chunk
    int64_t chunk_idx = 0;
    for (int64_t o0 = 0; o0 < plan->n_outer; o0 += plan->chunk_outer, ++chunk_idx) {
        const int64_t o1 = std::min(plan->n_outer, o0 + plan->chunk_outer);
        for (size_t p = 0; p < np; ++p) {
            const PassPlan& pp = plan->passes[p];
            rk_pass_args a = args[p];
            if (pp.src == BUF_WS0 || pp.src == BUF_WS1) a.in_off -= o0 * pp.src_stride0;
            if (pp.dst == BUF_WS0 || pp.dst == BUF_WS1) a.out_off -= o0 * pp.dst_stride0;
            const rk_u64 t0 = (rk_u64)o0 * pp.inst_per_outer / (rk_u64)pp.C;
            const rk_u64 t1 = (rk_u64)o1 * pp.inst_per_outer / (rk_u64)pp.C;
            const size_t ev = 2 * (chunk_idx * np + p);
            if (plan->timing && ev + 1 < plan->events.size()) RK_HIP(hipEventRecord(plan->events[ev], stream));
            int rc = launch_pass(pp, a, stream, &lb, &sb, t0, t1);
            if (rc != RK_OK) return rc;
            if (plan->timing && ev + 1 < plan->events.size()) RK_HIP(hipEventRecord(plan->events[ev + 1], stream));
        }
    }
    return RK_OK;
}

}  // namespace rk

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------

extern "C" int rk_fft_plan_create(const rk_fft_desc* desc, rk_fft_plan** plan) {
    return rk::plan_create(desc, plan);
}

extern "C" int rk_fft_plan_create_tuned(const rk_fft_desc* desc, const rk_fft_tuning* tuning, rk_fft_plan** plan) {
    if (tuning) {
        if (tuning->factor_max < 0 || tuning->first_factor < 0 || tuning->col_tile < 0 || tuning->row_tile < 0 ||
            tuning->chunk_bytes < -1 || tuning->fixed_modes < -1 || tuning->fixed_modes > 1 || tuning->reserved != 0)
            return fail(RK_ERR_INVALID, "knob out of range");
        if (tuning->factor_max == 1)
            return fail(RK_ERR_INVALID, "factor_max must be 0 or >= 2");
    }
    struct Scope {
        explicit Scope(const rk_fft_tuning* t) { rk::t_tuning = t; }
        ~Scope() { rk::t_tuning = nullptr; }
    } scope(tuning);
    return rk::plan_create(desc, plan);
}

extern "C" int rk_fft_execute_checked(const rk_fft_plan* plan, void* out, size_t out_bytes, const void* in,
                                      size_t in_bytes, int inverse, void* workspace, size_t workspace_bytes,
                                      void* stream) {
    if (!plan) return fail(RK_ERR_INVALID, "null plan");
    const int64_t need_in = rk::layout_extent_bytes(plan->desc, plan->desc.in_strides, plan->desc.in_offset, plan->es);
    const int64_t need_out =
        rk::layout_extent_bytes(plan->desc, plan->desc.out_strides, plan->desc.out_offset, plan->es);
    if (need_in < 0 || need_out < 0)
        return fail(RK_ERR_INVALID, "a layout reaches before its buffer's start");
    if ((int64_t)in_bytes < need_in)
        return fail(RK_ERR_INVALID, "input buffer of %zu bytes; the plan reads %lld", in_bytes, (long long)need_in);
    if ((int64_t)out_bytes < need_out)
        return fail(RK_ERR_INVALID, "output buffer of %zu bytes; the plan writes %lld", out_bytes,
                    (long long)need_out);
    if (workspace_bytes < plan->ws_bytes)
        return fail(RK_ERR_INVALID, "workspace of %zu bytes; the plan needs %zu", workspace_bytes, plan->ws_bytes);
    return rk::plan_execute(plan, out, in, inverse, workspace, (hipStream_t)stream, nullptr, 0, nullptr, 0);
}

extern "C" int rk_fft_plan_workspace_size(const rk_fft_plan* plan, size_t* bytes) {
    if (!plan || !bytes) return fail(RK_ERR_INVALID, "rk_fft_plan_workspace_size: null argument");
    *bytes = plan->ws_bytes;
    return RK_OK;
}

extern "C" int rk_fft_execute(const rk_fft_plan* plan, void* out, const void* in, int inverse, void* workspace,
                              void* stream) {
    return rk::plan_execute(plan, out, in, inverse, workspace, (hipStream_t)stream, nullptr, 0, nullptr, 0);
}

extern "C" int rk_fft_execute_fused(const rk_fft_plan* plan, void* out, const void* in, int inverse, void* workspace,
                                    void* stream, const void* load_blob, size_t load_bytes, const void* store_blob,
                                    size_t store_bytes) {
    if (load_bytes > RK_FUSED_BLOB_BYTES || store_bytes > RK_FUSED_BLOB_BYTES)
        return fail(RK_ERR_INVALID, "fused argument blob larger than %d bytes", RK_FUSED_BLOB_BYTES);
    return rk::plan_execute(plan, out, in, inverse, workspace, (hipStream_t)stream, load_blob, load_bytes, store_blob,
                            store_bytes);
}

// ---------------------------------------------------------------------------
// Execution that allocates nothing: the library's own device scratch array as
// the workspace (no counterpart in the reference).  The array is a file-scope
// device variable of RK_SCRATCH_BYTES, so every GPU the library runs on holds
// one copy, at its own address, until the process ends.  Host code never
// passes the array's name to a kernel (on the host it is only a stand-in for
// the device symbol): the device address comes from hipGetSymbolAddress, once
// per device with that device current, and is kept per device.  The passes
// write every workspace element before they read it, so nothing needs
// clearing between calls; calls whose work overlaps on one GPU (other
// streams, or the same plan on two streams) share the array and race.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char rk_scratch[RK_SCRATCH_BYTES];

static int scratch_address(int device, void** out) {
    static std::mutex mu;
    static std::map<int, void*> addr;   // device -> device address of rk_scratch
    std::lock_guard<std::mutex> lock(mu);
    auto it = addr.find(device);
    if (it == addr.end()) {
        int cur = 0;
        RK_HIP(hipGetDevice(&cur));
        if (cur != device) RK_HIP(hipSetDevice(device));
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(rk_scratch));
        if (cur != device) RK_HIP(hipSetDevice(cur));
        if (e != hipSuccess || !p)
            return fail(RK_ERR_HIP, "rk_fft_execute_scratch: address of the scratch array on device %d: %s", device,
                        hipGetErrorString(e));
        it = addr.emplace(device, p).first;
    }
    *out = it->second;
    return RK_OK;
}

extern "C" int rk_fft_execute_scratch(const rk_fft_plan* plan, void* out, const void* in, int inverse, void* stream) {
    if (!plan) return fail(RK_ERR_INVALID, "rk_fft_execute_scratch: null plan");
    if (plan->ws_bytes > (size_t)RK_SCRATCH_BYTES)
        return fail(RK_ERR_INVALID,
                    "rk_fft_execute_scratch: the plan needs %zu bytes of workspace, more than the %llu bytes of the "
                    "library's scratch array (use rk_fft_execute with a workspace)",
                    plan->ws_bytes, (unsigned long long)RK_SCRATCH_BYTES);
    void* ws = nullptr;
    if (plan->ws_bytes) {
        int rc = scratch_address(plan->device, &ws);
        if (rc != RK_OK) return rc;
    }
    return rk::plan_execute(plan, out, in, inverse, ws, (hipStream_t)stream, nullptr, 0, nullptr, 0);
}

extern "C" int rk_fft_plan_destroy(rk_fft_plan* plan) {
    if (!plan) return RK_OK;
    if (plan->tables) hipFree(plan->tables);
    for (hipEvent_t e : plan->events) hipEventDestroy(e);
    delete plan;
    return RK_OK;
}

extern "C" int rk_fft_plan_num_passes(const rk_fft_plan* plan, int* n) {
    if (!plan || !n) return fail(RK_ERR_INVALID, "rk_fft_plan_num_passes: null argument");
    *n = plan->trivial ? 0 : (int)plan->passes.size();
    return RK_OK;
}

extern "C" int rk_fft_plan_chunking(const rk_fft_plan* plan, int64_t* chunk_outer, int64_t* n_outer) {
    if (!plan || !chunk_outer || !n_outer) return fail(RK_ERR_INVALID, "rk_fft_plan_chunking: null argument");
    *chunk_outer = plan->chunk_outer;
    *n_outer = plan->n_outer;
    return RK_OK;
}

extern "C" int rk_fft_plan_pass_info(const rk_fft_plan* plan, int pass, rk_fft_pass_info* info) {
    if (!plan || !info) return fail(RK_ERR_INVALID, "rk_fft_plan_pass_info: null argument");
    if (pass < 0 || pass >= (int)plan->passes.size()) return fail(RK_ERR_INVALID, "pass index %d out of range", pass);
    const rk::PassPlan& pp = plan->passes[pass];
    memset(info, 0, sizeof(*info));
    info->L = pp.L;
    info->C = pp.C;
    info->threads = pp.threads;
    info->kind = pp.kind;
    info->load_mode = (int)pp.args.load_mode;
    info->store_mode = (int)pp.args.store_mode;
    info->fixed_modes = pp.fixed_modes ? 1 : 0;
    info->src = pp.src;
    info->dst = pp.dst;
    info->axis = pp.axis;
    info->pre_ops = pp.args.pre_ops;
    info->post_ops = pp.args.post_ops;
    info->n_inst = pp.args.n_inst;
    info->grid = pp.grid;
    info->axis_len = pp.args.N;
    info->Ns = pp.args.dim_NS;
    info->lds_bytes = pp.lds_bytes;
    info->jit = pp.jit ? 1 : 0;
    info->variant = pp.variant;
    info->occ = pp.occ ? 1 : 0;
    info->stream = (pp.nt_load && pp.nt_store) ? 1 : 0;
    info->nt_load = pp.nt_load ? 1 : 0;
    info->nt_store = pp.nt_store ? 1 : 0;
    info->wave = pp.wave ? 1 : 0;
    info->pipe = pp.pipe ? 1 : 0;
    info->in_dims[0] = (int64_t)(pp.args.n_inst / (pp.args.dim_I * pp.args.dim_NS * pp.args.dim_H));
    info->in_dims[1] = (int64_t)pp.args.n_valid_in;
    info->in_dims[2] = (int64_t)pp.args.dim_I;
    return RK_OK;
}

extern "C" int rk_fft_plan_set_pass_kernel(rk_fft_plan* plan, int pass, void* function) {
    if (!plan || !function) return fail(RK_ERR_INVALID, "rk_fft_plan_set_pass_kernel: null argument");
    if (pass < 0 || pass >= (int)plan->passes.size()) return fail(RK_ERR_INVALID, "pass index %d out of range", pass);
    plan->passes[pass].jit = (hipFunction_t)function;
    plan->passes[pass].aot = nullptr;
    plan->passes[pass].user_kernel = true;
    plan->passes[pass].fixed_modes = false;
    if (plan->passes[pass].pipe) {   // the replacement is an ordinary (one tile per work-group) kernel
        rk::PassPlan& pp = plan->passes[pass];
        pp.pipe = false;
        pp.grid = std::min<uint64_t>(pp.args.n_tiles, (uint64_t(1) << 31) / (uint64_t)pp.threads);
    }
    return RK_OK;
}

extern "C" int rk_fft_plan_set_timing(rk_fft_plan* plan, int enable) {
    if (!plan) return fail(RK_ERR_INVALID, "rk_fft_plan_set_timing: null plan");
    if (enable && plan->events.empty()) {
        const int64_t chunks =
            plan->chunk_outer ? (plan->n_outer + plan->chunk_outer - 1) / plan->chunk_outer : 1;
        plan->events.resize(2 * plan->passes.size() * (size_t)chunks);
        for (auto& e : plan->events) RK_HIP(hipEventCreate(&e));
    }
    plan->timing = enable != 0;
    return RK_OK;
}

extern "C" int rk_fft_plan_pass_time_ms(const rk_fft_plan* plan, int pass, float* ms) {
    if (!plan || !ms) return fail(RK_ERR_INVALID, "rk_fft_plan_pass_time_ms: null argument");
    *ms = -1.0f;
    const size_t np = plan->passes.size();
    if (!plan->timing || pass < 0 || pass >= (int)np) return RK_OK;
    // sum over the chunks the pass ran in
    float total = 0.0f;
    for (size_t ev = 2 * (size_t)pass; ev + 1 < plan->events.size(); ev += 2 * np) {
        float t = 0.0f;
        RK_HIP(hipEventSynchronize(plan->events[ev + 1]));
        RK_HIP(hipEventElapsedTime(&t, plan->events[ev], plan->events[ev + 1]));
        total += t;
    }
    *ms = total;
    return RK_OK;
}
