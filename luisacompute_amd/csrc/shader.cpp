// shader.cpp -- run-time compilation of shaders for gfx950 with hiprtc: the HIP counterpart of the
// CUDA backend's NVRTC path (src/backends/cuda/cuda_compiler.cpp, CUDADevice::create_shader
// cuda_device.cpp). The source sees the device library (lc_hip_device.h and the headers it
// includes) as in-memory headers embedded into the library when it is built, and is compiled
// with the flags of the library's own kernels: IEEE single precision without contraction (the
// traversal's watertight test rounds like the CPU oracle), f32 denormals flushed, no packed-FP32
// instructions.
//
// Code objects are cached by a 64-bit hash of everything that decides them (cache format, compile
// options, the device library's text, the source), as the CUDA backend caches PTX by checksum
// (cuda_device.cpp:555-604): in memory for the process, and on disk as kernel_<hash>.hsaco under
// LUISA_HIP_CACHE_DIR (default $HOME/.cache/luisacompute_amd; "0" turns the disk cache off). Each
// file starts with a header holding the hash and the payload size, so a stale or truncated file
// is ignored. LUISA_DUMP_SOURCE=1 also writes kernel_<hash>.hip beside the cache (the reference's
// switch of the same name, cuda_device.cpp:54-62).
#include <hip/hiprtc.h>

#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <sys/stat.h>
#include <unistd.h>
#include <unordered_map>
#include <vector>

#include "lc_hip_internal.h"

namespace lchip {

namespace {

const char *const kRtcOptions[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off",
                                   "-fgpu-flush-denormals-to-zero", "-Xclang", "-target-feature", "-Xclang",
                                   "-packed-fp32-ops", "-Wno-unused-function", "-Wno-parentheses-equality"};
constexpr int kRtcOptionCount = (int)(sizeof(kRtcOptions) / sizeof(kRtcOptions[0]));
constexpr uint64_t kCacheMagic = 0x31434F5350494843ull; // "CHIPSOC1": cache file format 1

uint64_t fnv1a(uint64_t h, const void *data, size_t n) {
    const unsigned char *p = static_cast<const unsigned char *>(data);
    for (size_t i = 0; i < n; i++) {
        h ^= p[i];
        h *= 0x100000001B3ull;
    }
    return h;
}

uint64_t fnv1a_str(uint64_t h, const char *s) { return fnv1a(h, s, std::strlen(s) + 1); }

struct CacheState {
    std::mutex mu;
    std::unordered_map<uint64_t, std::vector<char>> memory;
    uint64_t hits = 0, misses = 0;
};

CacheState &cache_state() {
    static CacheState c;
    return c;
}

// Directory of the disk cache ("" when off); created on first use.
std::string cache_dir() {
    std::string dir;
    if (const char *e = std::getenv("LUISA_HIP_CACHE_DIR")) {
        if (std::strcmp(e, "0") == 0) return "";
        dir = e;
    } else if (const char *h = std::getenv("HOME")) {
        dir = std::string(h) + "/.cache/luisacompute_amd";
    } else {
        return "";
    }
    // mkdir -p
    for (size_t i = 1; i <= dir.size(); i++) {
        if (i == dir.size() || dir[i] == '/') {
            const std::string part = dir.substr(0, i);
            if (mkdir(part.c_str(), 0755) != 0 && errno != EEXIST) return "";
        }
    }
    return dir;
}

std::string hex16(uint64_t v) {
    char b[17];
    std::snprintf(b, sizeof(b), "%016llx", (unsigned long long)v);
    return b;
}

bool read_cached(const std::string &path, uint64_t key, std::vector<char> &code) {
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    uint64_t hdr[3] = {0, 0, 0};
    bool ok = std::fread(hdr, sizeof(hdr), 1, f) == 1 && hdr[0] == kCacheMagic && hdr[1] == key && hdr[2] > 0 &&
              hdr[2] < (1ull << 31);
    if (ok) {
        code.resize(hdr[2]);
        ok = std::fread(code.data(), 1, code.size(), f) == code.size();
    }
    std::fclose(f);
    if (!ok) code.clear();
    return ok;
}

void write_file(const std::string &path, const void *head, size_t head_bytes, const void *data, size_t bytes) {
    const std::string tmp = path + ".tmp" + std::to_string((long long)getpid());
    FILE *f = std::fopen(tmp.c_str(), "wb");
    if (!f) return;
    bool ok = (head_bytes == 0 || std::fwrite(head, head_bytes, 1, f) == 1) && std::fwrite(data, 1, bytes, f) == bytes;
    ok = std::fclose(f) == 0 && ok;
    if (ok) ok = std::rename(tmp.c_str(), path.c_str()) == 0; // readers see a whole file or none
    if (!ok) std::remove(tmp.c_str());
}

bool rtc_compile_uncached(const char *source, const char *name, std::vector<char> &code, std::string &log);

} // namespace

uint64_t rtc_cache_key(const char *source) {
    uint64_t h = 0xCBF29CE484222325ull;
    h = fnv1a(h, &kCacheMagic, sizeof(kCacheMagic));
    for (int i = 0; i < kRtcOptionCount; i++) h = fnv1a_str(h, kRtcOptions[i]);
    for (int i = 0; i < kDeviceHeaderCount; i++) {
        h = fnv1a_str(h, kDeviceHeaderNames[i]);
        h = fnv1a_str(h, kDeviceHeaderSources[i]);
    }
    return fnv1a_str(h, source);
}

void rtc_cache_stats(uint64_t *hits, uint64_t *misses) {
    CacheState &c = cache_state();
    std::lock_guard<std::mutex> lk(c.mu);
    if (hits) *hits = c.hits;
    if (misses) *misses = c.misses;
}

bool rtc_compile(const char *source, const char *name, std::vector<char> &code, std::string &log) {
    code.clear();
    log.clear();
    const uint64_t key = rtc_cache_key(source);
    CacheState &c = cache_state();
    {
        std::lock_guard<std::mutex> lk(c.mu);
        auto it = c.memory.find(key);
        if (it != c.memory.end()) {
            code = it->second;
            c.hits++;
            return true;
        }
    }
    const std::string dir = cache_dir();
    const std::string base = dir.empty() ? std::string() : dir + "/kernel_" + hex16(key);
    const char *dump = std::getenv("LUISA_DUMP_SOURCE");
    if (!base.empty() && dump && std::strcmp(dump, "1") == 0) write_file(base + ".hip", nullptr, 0, source, std::strlen(source));
    bool ok = !base.empty() && read_cached(base + ".hsaco", key, code);
    const bool from_disk = ok;
    if (!ok) {
        ok = rtc_compile_uncached(source, name, code, log);
        if (ok && !base.empty()) {
            const uint64_t hdr[3] = {kCacheMagic, key, (uint64_t)code.size()};
            write_file(base + ".hsaco", hdr, sizeof(hdr), code.data(), code.size());
        }
    }
    if (ok) {
        std::lock_guard<std::mutex> lk(c.mu);
        c.memory.emplace(key, code);
        if (from_disk) c.hits++;
        else c.misses++;
    }
    return ok;
}

namespace {

bool rtc_compile_uncached(const char *source, const char *name, std::vector<char> &code, std::string &log) {
    code.clear();
    log.clear();
    hiprtcProgram prog = nullptr;
    hiprtcResult r = hiprtcCreateProgram(&prog, source, name, kDeviceHeaderCount, kDeviceHeaderSources,
                                         kDeviceHeaderNames);
    if (r != HIPRTC_SUCCESS) {
        log = std::string("hiprtcCreateProgram: ") + hiprtcGetErrorString(r);
        return false;
    }
    r = hiprtcCompileProgram(prog, kRtcOptionCount, kRtcOptions);
    size_t log_size = 0;
    if (hiprtcGetProgramLogSize(prog, &log_size) == HIPRTC_SUCCESS && log_size > 1) {
        log.resize(log_size);
        hiprtcGetProgramLog(prog, &log[0]);
        while (!log.empty() && log.back() == '\0') log.pop_back();
    }
    bool ok = r == HIPRTC_SUCCESS;
    if (ok) {
        size_t size = 0;
        ok = hiprtcGetCodeSize(prog, &size) == HIPRTC_SUCCESS && size > 0;
        if (ok) {
            code.resize(size);
            ok = hiprtcGetCode(prog, code.data()) == HIPRTC_SUCCESS;
        }
    } else if (log.empty()) {
        log = std::string("hiprtcCompileProgram: ") + hiprtcGetErrorString(r);
    }
    hiprtcDestroyProgram(&prog);
    return ok;
}

} // namespace

} // namespace lchip
