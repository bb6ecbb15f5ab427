// rk_jit.cpp — run-time compilation of FFT pass kernels that the
// ahead-of-time catalog does not hold (other smooth lengths, other tile
// widths).  The same fft_pass_body template is compiled by hiprtc from the
// embedded device headers, so there is one kernel code path.  This is the
// counterpart of reikna rendering fft.mako per shape and handing it to the
// driver JIT (reikna/core/computation.py:579-591).
#include <map>
#include <mutex>
#include <string>
#include <tuple>

#include "rk_internal.hpp"
#include "rk_embedded.inc"

namespace rk {

int compile_code_object(const std::string& source, const std::vector<std::string>& header_names,
                        const std::vector<std::string>& header_sources, const std::string& options,
                        std::string* code, std::string* log);

struct JitKernel {
    hipModule_t module;
    hipFunction_t fn;
    int threads, lds;
};

int jit_pass_kernel(int L, int C, int variant, hipFunction_t* fn, int* threads, int* lds_bytes) {
    static std::mutex mu;
    static std::map<std::tuple<int, int, int, int>, JitKernel> cache;
    if (L < 1 || L > 16384 || !is_smooth(L))
        return fail(RK_ERR_NO_KERNEL, "no kernel for DFT length %d (must be 7-smooth and <= 16384)", L);
    const int T = threads_per_instance(L);
    if ((int64_t)C * T > 1024) return fail(RK_ERR_NO_KERNEL, "tile %d x %d threads exceeds 1024", C, T);
    int dev = 0;
    hipGetDevice(&dev);
    std::lock_guard<std::mutex> lock(mu);
    auto key = std::make_tuple(dev, L, C, variant);
    auto it = cache.find(key);
    if (it == cache.end()) {
        // the catalogue's entry, for any variant number (IO policy, modes,
        // wave-local stages included)
        char src[512];
        snprintf(src, sizeof(src),
                 "#include \"fft_device.hpp\"\n"
                 "extern \"C\" __global__ void __launch_bounds__(%d, (rkfft::pass_min_waves<%d, %d, %d>()))\n"
                 "rk_jit_pass(rk_pass_args a) {\n"
                 "    rkfft::fft_pass_entry<%d, %d, %d>(a);\n"
                 "}\n",
                 C * T, L, C, variant, L, C, variant);
        std::string code, log;
        int rc = compile_code_object(src, {"fft_args.h", "fft_device.hpp"}, {k_src_fft_args_h, k_src_fft_device_hpp},
                                     "-Xclang -target-feature -Xclang -packed-fp32-ops", &code, &log);
        if (rc != RK_OK) return rc;
        JitKernel k;
        RK_HIP(hipModuleLoadData(&k.module, code.data()));
        RK_HIP(hipModuleGetFunction(&k.fn, k.module, "rk_jit_pass"));
        k.threads = C * T;
        const int es = (variant & VAR_F64) ? 16 : 8;
        k.lds = (variant & VAR_HALF) ? tile_lds_bytes_host(L, C, es) / 2 : tile_lds_bytes_host(L, C, es);
        it = cache.emplace(key, k).first;
    }
    *fn = it->second.fn;
    *threads = it->second.threads;
    *lds_bytes = it->second.lds;
    return RK_OK;
}

}  // namespace rk
