// shader.cpp -- run-time compilation of shaders for gfx950 with hiprtc: the HIP counterpart of the
// CUDA backend's NVRTC path (src/backends/cuda/cuda_compiler.cpp, CUDADevice::create_shader
// cuda_device.cpp). The source sees the device library (lc_hip_device.h and the headers it
// includes) as in-memory headers embedded into the library when it is built, and is compiled
// with the flags of the library's own kernels: IEEE single precision without contraction (the
// traversal's watertight test rounds like the CPU oracle), f32 denormals flushed, no packed-FP32
// instructions.
#include <hip/hiprtc.h>

#include <string>
#include <vector>

#include "lc_hip_internal.h"

namespace lchip {

bool rtc_compile(const char *source, const char *name, std::vector<char> &code, std::string &log) {
    code.clear();
    log.clear();
    hiprtcProgram prog = nullptr;
    hiprtcResult r = hiprtcCreateProgram(&prog, source, name, kDeviceHeaderCount, kDeviceHeaderSources,
                                         kDeviceHeaderNames);
    if (r != HIPRTC_SUCCESS) {
        log = std::string("hiprtcCreateProgram: ") + hiprtcGetErrorString(r);
        return false;
    }
    const char *opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off",
                          "-fgpu-flush-denormals-to-zero", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                          "-Wno-unused-function"};
    r = hiprtcCompileProgram(prog, (int)(sizeof(opts) / sizeof(opts[0])), opts);
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

} // namespace lchip
