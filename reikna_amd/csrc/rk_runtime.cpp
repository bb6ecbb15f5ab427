// rk_runtime.cpp — device runtime of the C-ABI: errors, memory, streams,
// events, and run-time compiled modules (hiprtc with an on-disk code cache).
//
// This replaces what reikna takes from grunnur: Array.from_host / empty_like
// (reikna/core/computation.py:401,444,689), Buffer.allocate (:686),
// Queue.synchronize (:597) and StaticKernel's driver JIT (:579-591).
#include <hip/hiprtc.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <mutex>
#include <sstream>

#include "rk_internal.hpp"

namespace rk {

static thread_local std::string g_last_error;

int fail(int code, const char* fmt, ...) {
    char buf[2048];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

void clear_error() { g_last_error.clear(); }

}  // namespace rk

using rk::fail;

extern "C" const char* rk_last_error(void) { return rk::g_last_error.c_str(); }

extern "C" const char* rk_version(void) { return "reikna_amd 0.1 gfx950"; }

// ---------------------------------------------------------------------------
// device and memory
// ---------------------------------------------------------------------------

extern "C" int rk_device_count(int* count) {
    if (!count) return fail(RK_ERR_INVALID, "rk_device_count: null output");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;   // no device: report zero, not an error
    *count = n;
    return RK_OK;
}

extern "C" int rk_set_device(int device) {
    RK_HIP(hipSetDevice(device));
    return RK_OK;
}

extern "C" int rk_get_device_info(int device, rk_device_info* info) {
    if (!info) return fail(RK_ERR_INVALID, "rk_get_device_info: null output");
    hipDeviceProp_t p;
    RK_HIP(hipGetDeviceProperties(&p, device));
    memset(info, 0, sizeof(*info));
    snprintf(info->name, sizeof(info->name), "%s", p.name);
    snprintf(info->arch, sizeof(info->arch), "%s", p.gcnArchName);
    info->compute_units = p.multiProcessorCount;
    info->wavefront_size = p.warpSize;
    info->max_threads_per_block = p.maxThreadsPerBlock;
    info->lds_per_block = p.sharedMemPerBlock;
    info->total_global_mem = p.totalGlobalMem;
    info->l2_cache_bytes = p.l2CacheSize;
    info->clock_khz = p.clockRate;
    info->memory_clock_khz = p.memoryClockRate;
    info->memory_bus_width = p.memoryBusWidth;
    return RK_OK;
}

extern "C" int rk_malloc(void** ptr, size_t bytes) {
    if (!ptr) return fail(RK_ERR_INVALID, "rk_malloc: null output");
    *ptr = nullptr;
    if (bytes == 0) return RK_OK;
    RK_HIP(hipMalloc(ptr, bytes));
    return RK_OK;
}

extern "C" int rk_free(void* ptr) {
    if (ptr) RK_HIP(hipFree(ptr));
    return RK_OK;
}

extern "C" int rk_memcpy_htod(void* dst, const void* src, size_t bytes, void* stream) {
    if (bytes == 0) return RK_OK;
    RK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream));
    RK_HIP(hipStreamSynchronize((hipStream_t)stream));
    return RK_OK;
}

extern "C" int rk_memcpy_dtoh(void* dst, const void* src, size_t bytes, void* stream) {
    if (bytes == 0) return RK_OK;
    RK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream));
    RK_HIP(hipStreamSynchronize((hipStream_t)stream));
    return RK_OK;
}

extern "C" int rk_memcpy_dtod(void* dst, const void* src, size_t bytes, void* stream) {
    if (bytes == 0) return RK_OK;
    RK_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)stream));
    return RK_OK;
}

extern "C" int rk_memset(void* dst, int value, size_t bytes, void* stream) {
    if (bytes == 0) return RK_OK;
    RK_HIP(hipMemsetAsync(dst, value, bytes, (hipStream_t)stream));
    return RK_OK;
}

extern "C" int rk_mem_info(size_t* free_bytes, size_t* total_bytes) {
    size_t f = 0, t = 0;
    RK_HIP(hipMemGetInfo(&f, &t));
    if (free_bytes) *free_bytes = f;
    if (total_bytes) *total_bytes = t;
    return RK_OK;
}

// ---------------------------------------------------------------------------
// streams and events
// ---------------------------------------------------------------------------

extern "C" int rk_stream_create(void** stream) {
    if (!stream) return fail(RK_ERR_INVALID, "rk_stream_create: null output");
    hipStream_t s;
    RK_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = s;
    return RK_OK;
}

extern "C" int rk_stream_destroy(void* stream) {
    if (stream) RK_HIP(hipStreamDestroy((hipStream_t)stream));
    return RK_OK;
}

extern "C" int rk_stream_synchronize(void* stream) {
    RK_HIP(hipStreamSynchronize((hipStream_t)stream));
    return RK_OK;
}

extern "C" int rk_device_synchronize(void) {
    RK_HIP(hipDeviceSynchronize());
    return RK_OK;
}

extern "C" int rk_event_create(void** event) {
    if (!event) return fail(RK_ERR_INVALID, "rk_event_create: null output");
    hipEvent_t e;
    RK_HIP(hipEventCreate(&e));
    *event = e;
    return RK_OK;
}

extern "C" int rk_event_destroy(void* event) {
    if (event) RK_HIP(hipEventDestroy((hipEvent_t)event));
    return RK_OK;
}

extern "C" int rk_event_record(void* event, void* stream) {
    RK_HIP(hipEventRecord((hipEvent_t)event, (hipStream_t)stream));
    return RK_OK;
}

extern "C" int rk_event_synchronize(void* event) {
    RK_HIP(hipEventSynchronize((hipEvent_t)event));
    return RK_OK;
}

extern "C" int rk_event_query(void* event, int* done) {
    hipError_t e = hipEventQuery((hipEvent_t)event);
    if (e == hipSuccess) {
        *done = 1;
        return RK_OK;
    }
    if (e == hipErrorNotReady) {
        *done = 0;
        return RK_OK;
    }
    return fail(RK_ERR_HIP, "hipEventQuery failed: %s", hipGetErrorString(e));
}

extern "C" int rk_event_elapsed_ms(void* start, void* stop, float* ms) {
    RK_HIP(hipEventElapsedTime(ms, (hipEvent_t)start, (hipEvent_t)stop));
    return RK_OK;
}

extern "C" int rk_stream_wait_event(void* stream, void* event) {
    RK_HIP(hipStreamWaitEvent((hipStream_t)stream, (hipEvent_t)event, 0));
    return RK_OK;
}

// ---------------------------------------------------------------------------
// run-time compiled modules
// ---------------------------------------------------------------------------

namespace rk {

static uint64_t fnv1a(const std::string& s, uint64_t h = 1469598103934665603ull) {
    for (unsigned char ch : s) {
        h ^= ch;
        h *= 1099511628211ull;
    }
    return h;
}

static std::string cache_dir() {
    const char* env = getenv("RK_CACHE_DIR");
    std::string d;
    if (env && *env) {
        d = env;
    } else {
        const char* home = getenv("HOME");
        d = std::string(home && *home ? home : "/tmp") + "/.cache/reikna_amd";
    }
    // create the directory chain
    std::string cur;
    std::stringstream ss(d);
    std::string part;
    if (!d.empty() && d[0] == '/') cur = "/";
    while (std::getline(ss, part, '/')) {
        if (part.empty()) continue;
        cur += part + "/";
        mkdir(cur.c_str(), 0755);
    }
    return d;
}

static std::string current_arch() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return "gfx950";
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, dev) != hipSuccess) return "gfx950";
    std::string a = p.gcnArchName;
    size_t colon = a.find(':');
    if (colon != std::string::npos) a = a.substr(0, colon);
    return a;
}

// Compiles `source` (plus named headers) to a code object for the current
// device, or returns the cached one.
int compile_code_object(const std::string& source, const std::vector<std::string>& header_names,
                        const std::vector<std::string>& header_sources, const std::string& options,
                        std::string* code, std::string* log) {
    const std::string arch = current_arch();
    std::string key_text = source + "\x01" + options + "\x01" + arch;
    for (size_t i = 0; i < header_names.size(); ++i) key_text += "\x02" + header_names[i] + "\x03" + header_sources[i];
    char key[32];
    snprintf(key, sizeof(key), "%016llx", (unsigned long long)fnv1a(key_text));
    const std::string path = cache_dir() + "/" + key + "-" + arch + ".co";
    {
        std::ifstream f(path, std::ios::binary);
        if (f) {
            std::stringstream b;
            b << f.rdbuf();
            *code = b.str();
            if (!code->empty()) return RK_OK;
        }
    }
    hiprtcProgram prog;
    std::vector<const char*> hn, hs;
    for (size_t i = 0; i < header_names.size(); ++i) {
        hn.push_back(header_names[i].c_str());
        hs.push_back(header_sources[i].c_str());
    }
    hiprtcResult r = hiprtcCreateProgram(&prog, source.c_str(), "rk_generated.hip", (int)hn.size(),
                                         hs.empty() ? nullptr : hs.data(), hn.empty() ? nullptr : hn.data());
    if (r != HIPRTC_SUCCESS) return fail(RK_ERR_NO_KERNEL, "hiprtcCreateProgram: %s", hiprtcGetErrorString(r));
    std::vector<std::string> opts;
    opts.push_back("--offload-arch=" + arch);
    opts.push_back("-O3");
    opts.push_back("-std=c++17");
    {
        std::stringstream os(options);
        std::string o;
        while (os >> o) opts.push_back(o);
    }
    std::vector<const char*> optp;
    for (auto& o : opts) optp.push_back(o.c_str());
    r = hiprtcCompileProgram(prog, (int)optp.size(), optp.data());
    size_t log_size = 0;
    hiprtcGetProgramLogSize(prog, &log_size);
    std::string lg(log_size + 1, '\0');
    if (log_size) hiprtcGetProgramLog(prog, &lg[0]);
    if (log) *log = lg;
    if (r != HIPRTC_SUCCESS) {
        hiprtcDestroyProgram(&prog);
        return fail(RK_ERR_NO_KERNEL, "hiprtc compile failed: %s\n%s", hiprtcGetErrorString(r), lg.c_str());
    }
    size_t code_size = 0;
    hiprtcGetCodeSize(prog, &code_size);
    code->assign(code_size, '\0');
    hiprtcGetCode(prog, &(*code)[0]);
    hiprtcDestroyProgram(&prog);
    // best-effort cache write (atomic rename)
    const std::string tmp = path + ".tmp" + std::to_string((long long)getpid());
    {
        std::ofstream f(tmp, std::ios::binary);
        if (f) f.write(code->data(), (std::streamsize)code->size());
    }
    rename(tmp.c_str(), path.c_str());
    return RK_OK;
}

}  // namespace rk

extern "C" int rk_module_compile(const char* source, const char* options, void** module, char* log,
                                 size_t log_size) {
    if (!source || !module) return fail(RK_ERR_INVALID, "rk_module_compile: null argument");
    std::string code, lg;
    int rc = rk::compile_code_object(source, {}, {}, options ? options : "", &code, &lg);
    if (log && log_size) {
        snprintf(log, log_size, "%s", lg.c_str());
    }
    if (rc != RK_OK) return rc;
    hipModule_t m;
    RK_HIP(hipModuleLoadData(&m, code.data()));
    *module = m;
    return RK_OK;
}

extern "C" int rk_module_check(const char* source, const char* options, char* log, size_t log_size) {
    if (!source) return fail(RK_ERR_INVALID, "rk_module_check: null source");
    std::string code, lg;
    int rc = rk::compile_code_object(source, {}, {}, options ? options : "", &code, &lg);
    if (log && log_size) snprintf(log, log_size, "%s", lg.c_str());
    return rc;
}

extern "C" int rk_module_get_function(void* module, const char* name, void** function) {
    hipFunction_t f;
    RK_HIP(hipModuleGetFunction(&f, (hipModule_t)module, name));
    *function = f;
    return RK_OK;
}

extern "C" int rk_module_unload(void* module) {
    if (module) RK_HIP(hipModuleUnload((hipModule_t)module));
    return RK_OK;
}

extern "C" int rk_launch(void* function, const uint32_t grid[3], const uint32_t block[3], uint32_t shared_bytes,
                         void* stream, const void* args, size_t args_bytes) {
    if (!function || !grid || !block) return fail(RK_ERR_INVALID, "rk_launch: null argument");
    size_t sz = args_bytes;
    void* config[] = {HIP_LAUNCH_PARAM_BUFFER_POINTER, const_cast<void*>(args), HIP_LAUNCH_PARAM_BUFFER_SIZE, &sz,
                      HIP_LAUNCH_PARAM_END};
    RK_HIP(hipModuleLaunchKernel((hipFunction_t)function, grid[0], grid[1], grid[2], block[0], block[1], block[2],
                                 shared_bytes, (hipStream_t)stream, nullptr, config));
    return RK_OK;
}
