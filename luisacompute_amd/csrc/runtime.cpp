// runtime.cpp -- host side of libluisa_compute_backend_hip.so: devices, buffers, streams, events, mesh/accel
// builds and trace launches behind the C ABI of include/lc_hip.h.
//
// Structure follows the reference's CUDA backend (CUDADevice / CUDAStream / CUDAMesh /
// CUDAAccel, src/backends/cuda/cuda_device.cpp, cuda_stream.cpp, cuda_mesh.cpp, cuda_accel.cpp)
// with OptiX replaced by the kernels of bvh_build.hip and trace.hip.
#include "lc_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <rocprofiler-sdk-roctx/roctx.h>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>

#include "lc_hip_internal.h"

using namespace lchip;

namespace {

thread_local std::string g_last_error;

int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

int hip_fail(hipError_t e, const char *what) {
    g_last_error = std::string(what) + ": " + hipGetErrorString(e);
    return LC_HIP_ERROR_HIP;
}

#define LC_HIP_TRY(expr)                                      \
    do {                                                      \
        hipError_t _e = (expr);                               \
        if (_e != hipSuccess) return hip_fail(_e, #expr);     \
    } while (0)

// Grows a stream's working buffer (grow(), below); a stream under capture cannot grow.
#define LC_HIP_GROW(args)                                                                                     \
    do {                                                                                                      \
        hipError_t _e = grow args;                                                                            \
        if (_e == hipErrorStreamCaptureUnsupported)                                                           \
            return fail(LC_HIP_ERROR_UNSUPPORTED, "a captured launch needs more working memory than the "    \
                                                  "stream holds: run the work once outside the capture first"); \
        if (_e != hipSuccess) return hip_fail(_e, "stream working memory");                                   \
    } while (0)

template<typename T>
T *obj(uint64_t h) { return (h == 0 || h == LC_HIP_INVALID_HANDLE) ? nullptr : reinterpret_cast<T *>(h); }

template<typename T>
uint64_t handle(T *p) { return reinterpret_cast<uint64_t>(p); }

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(const Device *d) {
        hipGetDevice(&prev);
        if (d && prev != d->index) hipSetDevice(d->index);
    }
    ~DeviceGuard() {
        int cur = -1;
        hipGetDevice(&cur);
        if (prev >= 0 && cur != prev) hipSetDevice(prev);
    }
};

hipStream_t native_stream(uint64_t s) {
    Stream *st = obj<Stream>(s);
    return st ? st->s : nullptr;
}

// Grows a device allocation to at least `bytes` (contents are not preserved).
hipError_t ensure(void *&ptr, size_t &cap, size_t bytes) {
    if (cap >= bytes && ptr) return hipSuccess;
    if (ptr) {
        hipError_t e = hipFree(ptr);
        if (e != hipSuccess) return e;
        ptr = nullptr;
        cap = 0;
    }
    size_t b = std::max<size_t>(bytes, 256);
    hipError_t e = hipMalloc(&ptr, b);
    if (e == hipSuccess) cap = b;
    return e;
}

bool capturing(hipStream_t s) {
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    return hipStreamIsCapturing(s, &st) == hipSuccess && st != hipStreamCaptureStatusNone;
}

// Grows one of a stream's working buffers. While graphs pin the scratch, the old buffer is
// retired (their launches may name it) rather than freed. A stream being captured cannot grow:
// the launches it records must name memory that exists when they replay, so the work is run
// once outside the capture first.
hipError_t grow(StreamScratch *S, hipStream_t s, void *&ptr, size_t &cap, size_t bytes) {
    if (cap >= bytes && ptr) return hipSuccess;
    if (capturing(s)) return hipErrorStreamCaptureUnsupported;
    if (ptr && S->pins) {
        S->retired.push_back(ptr);
        ptr = nullptr;
        cap = 0;
    }
    return ensure(ptr, cap, bytes);
}

// Timed region on a stream: records HIP events around the launches when timing is on (not
// while the stream is being captured: a captured region has no time of its own).
struct TimedRegion {
    Device *dev;
    hipStream_t s;
    const char *name;
    hipEvent_t a = nullptr, b = nullptr;
    TimedRegion(Device *d, hipStream_t st, const char *n) : dev(d), s(st), name(n) {
        if (dev->timing && !capturing(s)) {
            hipEventCreate(&a);
            hipEventCreate(&b);
            hipEventRecord(a, s);
        }
    }
    ~TimedRegion() {
        if (a) {
            hipEventRecord(b, s);
            dev->pending.push_back({name, a, b});
        }
    }
};

// Host-side marker range (roctx, shown by rocprofv3 --marker-trace) around the work of a named
// resource, as the CUDA backend brackets a named accel's build with NVTX ranges
// (src/backends/cuda/cuda_accel.cpp:28,68,101). Nothing happens for unnamed resources.
struct MarkerRange {
    bool on = false;
    MarkerRange(const std::string &name, const char *what) {
        if (!name.empty()) {
            const std::string label = name + "::" + what;
            roctxRangePushA(label.c_str());
            on = true;
        }
    }
    ~MarkerRange() {
        if (on) roctxRangePop();
    }
};

hipError_t collect_timings(Device *dev) {
    hipError_t err = hipSuccess;
    for (auto &p : dev->pending) {
        hipError_t e = hipEventSynchronize(p.b);
        float ms = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, p.a, p.b);
        if (e == hipSuccess) {
            auto &t = dev->timings[p.name];
            t.total_ms += ms;
            t.launches++;
        } else if (err == hipSuccess) {
            err = e;
        }
        hipEventDestroy(p.a);
        hipEventDestroy(p.b);
    }
    dev->pending.clear();
    return err;
}

void stream_worker(Stream *st) {
    for (;;) {
        std::pair<hipEvent_t, std::function<void()>> item;
        {
            std::unique_lock<std::mutex> lk(st->qmu);
            st->qcv.wait(lk, [&] { return st->stop || !st->q.empty(); });
            if (st->q.empty()) return;
            item = std::move(st->q.front());
            st->q.pop_front();
        }
        hipEventSynchronize(item.first);
        hipEventDestroy(item.first);
        item.second();
    }
}

// Reads back the counters of a build (waits for the copy enqueued after the build).
int mesh_stats_sync(Mesh *m) {
    if (!m->built) return fail(LC_HIP_ERROR_NOT_BUILT, "mesh is not built");
    if (!m->stats_valid) {
        if (m->ntris > 0) {
            LC_HIP_TRY(hipEventSynchronize(m->ready));
            m->stats = *m->host;
            if (m->stats.tail_level_count != 0)
                return fail(LC_HIP_ERROR_DEVICE_FAULT, "BVH collapse exceeded the level bound");
            if (m->stats.tri_count != m->ntris)
                return fail(LC_HIP_ERROR_DEVICE_FAULT, "BVH build lost primitives");
            if (m->stats.error & 2u)
                return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "triangle index out of range of the vertex buffer");
        } else {
            std::memset(&m->stats, 0, sizeof(m->stats));
        }
        m->stats_valid = true;
    }
    return LC_HIP_OK;
}

void invert_affine(const float m[12], float out[12]) {
    double a = m[0], b = m[1], c = m[2], d = m[4], e = m[5], f = m[6], g = m[8], h = m[9], i = m[10];
    double A = e * i - f * h, B = -(d * i - f * g), C = d * h - e * g;
    double det = a * A + b * B + c * C;
    double id = det != 0.0 ? 1.0 / det : 0.0;
    double r[9] = {A * id, -(b * i - c * h) * id, (b * f - c * e) * id,
                   B * id, (a * i - c * g) * id, -(a * f - c * d) * id,
                   C * id, -(a * h - b * g) * id, (a * e - b * d) * id};
    double tx = m[3], ty = m[7], tz = m[11];
    for (int row = 0; row < 3; row++) {
        out[row * 4 + 0] = (float)r[row * 3 + 0];
        out[row * 4 + 1] = (float)r[row * 3 + 1];
        out[row * 4 + 2] = (float)r[row * 3 + 2];
        out[row * 4 + 3] = (float)(-(r[row * 3 + 0] * tx + r[row * 3 + 1] * ty + r[row * 3 + 2] * tz));
    }
}

// Primitives per BVH leaf slot by AccelOption hint (resource.h:94-126): FAST_TRACE builds single-
// primitive leaves (more wide nodes, fewer triangle tests: path tracer 2.56 -> 3.28 G rays/s);
// FAST_BUILD allows 3 (fewer nodes, 19% shorter build of 16.7M triangles).
uint32_t leaf_size_for(uint32_t hint) { return hint == LC_HIP_ACCEL_FAST_BUILD ? 3u : 1u; }
// Binned-SAH refinement of the radix tree's clusters for every hint but FAST_BUILD (the reference
// CPU backend builds Embree's high-quality SAH for FAST_TRACE, cpu/accel.rs:58-62).
uint32_t sah_for(const Device *dev, uint32_t hint) { return hint != LC_HIP_ACCEL_FAST_BUILD && dev->sah_refine ? 1u : 0u; }

void copy_counters_async(BuildCounters *dctr, HostCounters *host, uint32_t *error, hipStream_t s) {
    hipMemcpyAsync(host, dctr, offsetof(BuildCounters, level), hipMemcpyDeviceToHost, s);
    hipMemcpyAsync(&host->tail_level_count, level_nodes_word(dctr, kCollapseLaunches), sizeof(uint32_t),
                   hipMemcpyDeviceToHost, s);
    hipMemcpyAsync(&host->error, error, sizeof(uint32_t), hipMemcpyDeviceToHost, s);
}

// Compacting builds (allow_compaction). cuda_mesh.cpp:42-89 builds into a temporary buffer, reads
// the compacted size back and copies into the primitive's own buffer, which only grows. Here the
// tree is built into the device's arena of worst-case wide nodes (n nodes for n primitives) with
// the arena's build scratch; the host then waits for the node count and the used nodes are copied
// into the primitive's own array. A built primitive thus holds node_count nodes instead of n and
// no build scratch; with allow_update it also keeps the refit state (the build's counters and 32 B
// per wide node), copied out of the arena scratch with the nodes. Called with dev->mu held.
struct BuildTarget {
    WideNode *nodes;
    void *scratch;
    size_t scratch_bytes;
};

hipError_t arena_acquire(Device *dev, Mesh *m, uint32_t n, BuildTarget &t) {
    // the previous compaction's copy out of the arena has finished before the arena is reused
    hipError_t e = hipEventSynchronize(dev->arena_done);
    if (e != hipSuccess) return e;
    void *p = dev->arena_nodes;
    size_t cap = dev->arena_node_cap * sizeof(WideNode);
    e = ensure(p, cap, bvh_node_capacity(n) * sizeof(WideNode));
    dev->arena_nodes = static_cast<WideNode *>(p);
    dev->arena_node_cap = cap / sizeof(WideNode);
    if (e != hipSuccess) return e;
    t.nodes = dev->arena_nodes;
    e = ensure(dev->arena_scratch, dev->arena_scratch_bytes, bvh_scratch_bytes(n));
    t.scratch = dev->arena_scratch;
    t.scratch_bytes = dev->arena_scratch_bytes;
    // a primitive that compacted before keeps no scratch of its own
    if (m->scratch) {
        hipFree(m->scratch);
        m->scratch = nullptr;
        m->scratch_bytes = 0;
    }
    return e;
}

// After the build's launches and counter copies on s: waits for the counters, grows the
// primitive's node array to the node count if needed and copies the nodes out of the arena; with
// allow_update (triangle meshes) also the refit state: the build's counters (`ctr`, in the arena
// scratch) into a per-primitive array with room for the nodes' exact boxes.
int arena_compact(Device *dev, Mesh *m, hipStream_t s, const BuildCounters *ctr) {
    MarkerRange mr(m->name, "compact");
    LC_HIP_TRY(hipStreamSynchronize(s));
    const uint32_t count = m->host->node_count;
    if (count == 0 || count > dev->arena_node_cap) return fail(LC_HIP_ERROR_DEVICE_FAULT, "BVH node count out of range");
    if (!m->nodes || m->node_cap < count) {
        if (m->nodes) LC_HIP_TRY(hipFree(m->nodes));
        m->nodes = nullptr;
        m->node_cap = 0;
        LC_HIP_TRY(hipMalloc(&m->nodes, (size_t)count * sizeof(WideNode)));
        m->node_cap = count;
    }
    LC_HIP_TRY(hipMemcpyAsync(m->nodes, dev->arena_nodes, (size_t)count * sizeof(WideNode), hipMemcpyDeviceToDevice, s));
    if (m->allow_update && !m->procedural && ctr) {
        LC_HIP_TRY(ensure(m->refit_state, m->refit_state_bytes, bvh_refit_state_bytes(count)));
        LC_HIP_TRY(hipMemcpyAsync(bvh_refit_state_counters(m->refit_state), ctr, sizeof(BuildCounters),
                                  hipMemcpyDeviceToDevice, s));
        m->refit_nodes = count;
    }
    LC_HIP_TRY(hipEventRecord(dev->arena_done, s));
    return LC_HIP_OK;
}

void free_stream_scratch(StreamScratch *S) {
    if (S->words) hipFree(S->words);
    if (S->spill) hipFree(S->spill);
    if (S->sort) hipFree(S->sort);
    for (void *p : S->retired) hipFree(p);
    delete S;
}

// Drops one graph's (or capture's) pin on a stream's scratch (caller holds dev->mu, and the
// launches that named retired buffers have completed).
void unpin(Device *dev, StreamScratch *S) {
    if (!S || S->pins == 0) return;
    if (--S->pins > 0) return;
    for (void *p : S->retired) hipFree(p);
    S->retired.clear();
    auto it = std::find(dev->orphans.begin(), dev->orphans.end(), S);
    if (it != dev->orphans.end()) {
        dev->orphans.erase(it);
        free_stream_scratch(S);
    }
}

} // namespace

StreamScratch *lchip::stream_scratch(Device *dev, hipStream_t s) {
    auto it = dev->scratch.find(s);
    if (it != dev->scratch.end()) return it->second;
    auto *S = new StreamScratch;
    void *p = nullptr;
    if (hipMalloc(&p, kStreamWordsBytes) != hipSuccess) {
        delete S;
        return nullptr;
    }
    // ordered before any launch on s that uses the words
    if (hipMemsetAsync(p, 0, kStreamWordsBytes, s) != hipSuccess) {
        hipFree(p);
        delete S;
        return nullptr;
    }
    S->words = static_cast<uint32_t *>(p);
    S->ray_counter = S->words;
    S->error = S->words + kRayCounterWords;
    S->stats = reinterpret_cast<unsigned long long *>(static_cast<char *>(p) + 576);
    dev->scratch.emplace(s, S);
    return S;
}

void lchip::stream_scratch_release(Device *dev, hipStream_t s) {
    std::lock_guard<std::mutex> lk(dev->mu);
    auto it = dev->scratch.find(s);
    if (it == dev->scratch.end()) return;
    if (it->second->pins) {
        dev->orphans.push_back(it->second); // graphs still name it: freed with their last pin
    } else {
        free_stream_scratch(it->second);
    }
    dev->scratch.erase(it);
}

bool lchip::defer_while_capturing(Device *dev, std::function<void()> fn) {
    std::lock_guard<std::mutex> lk(dev->mu);
    if (dev->capturing.empty()) return false;
    dev->deferred.push_back(std::move(fn));
    return true;
}

namespace {
// Runs the destroys deferred during captures once none is open (caller does not hold dev->mu).
void run_deferred(Device *dev) {
    std::vector<std::function<void()>> fns;
    {
        std::lock_guard<std::mutex> lk(dev->mu);
        if (!dev->capturing.empty()) return;
        fns.swap(dev->deferred);
    }
    for (auto &f : fns) f();
}
} // namespace

extern "C" {

const char *lc_hip_last_error(void) { return g_last_error.c_str(); }

int lc_hip_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

uint64_t lc_hip_device_create(int32_t index) {
    int n = lc_hip_device_count();
    if (n <= 0) {
        fail(LC_HIP_ERROR_UNSUPPORTED, "no HIP device");
        return LC_HIP_INVALID_HANDLE;
    }
    if (index < 0 || index >= n) index = 0;
    auto dev = std::make_unique<Device>();
    dev->index = index;
    hipDeviceProp_t prop{};
    if (hipGetDeviceProperties(&prop, index) != hipSuccess) {
        fail(LC_HIP_ERROR_HIP, "hipGetDeviceProperties failed");
        return LC_HIP_INVALID_HANDLE;
    }
    dev->cu_count = prop.multiProcessorCount;
    dev->name = prop.gcnArchName;
    DeviceGuard g(dev.get());
    void *p = nullptr;
    if (hipMalloc(&p, 256) != hipSuccess) {
        fail(LC_HIP_ERROR_OUT_OF_MEMORY, "device scratch allocation failed");
        return LC_HIP_INVALID_HANDLE;
    }
    hipMemset(p, 0, 256);
    dev->error = static_cast<uint32_t *>(p);
    // the no-allocation trace's scratch array on this GPU (its entry point fails without it)
    if (noalloc_scratch_lookup(&dev->noalloc_scratch) != hipSuccess) {
        dev->noalloc_scratch = nullptr;
        (void)hipGetLastError();
    }
    if (const char *v = std::getenv("LUISA_HIP_BVH_SAH")) dev->sah_refine = std::atoi(v) != 0 ? 1u : 0u;
    if (const char *v = std::getenv("LUISA_HIP_BVH_WIDTH")) dev->bvh_width = (uint32_t)std::min(8, std::max(2, std::atoi(v)));
    dev->cfg[0] = TraceConfig{256, 16, 32, 0};
    dev->cfg[1] = TraceConfig{256, 16, 32, 0};
    if (hipEventCreateWithFlags(&dev->arena_done, hipEventDisableTiming) != hipSuccess) {
        hipFree(dev->error);
        fail(LC_HIP_ERROR_HIP, "device event creation failed");
        return LC_HIP_INVALID_HANDLE;
    }
    return handle(dev.release());
}

void lc_hip_device_destroy(uint64_t device) {
    Device *dev = obj<Device>(device);
    if (!dev) return;
    DeviceGuard g0(dev);
    {
        // destroys still waiting for a capture that was never ended
        std::vector<std::function<void()>> fns;
        {
            std::lock_guard<std::mutex> lk(dev->mu);
            for (auto &kv : dev->capturing) { // captures left open: discard what they recorded
                hipGraph_t gr = nullptr;
                if (hipStreamEndCapture(kv.first, &gr) == hipSuccess && gr) hipGraphDestroy(gr);
                (void)hipGetLastError();
                unpin(dev, kv.second);
            }
            dev->capturing.clear();
            fns.swap(dev->deferred);
        }
        for (auto &f : fns) f();
    }
    DeviceGuard g(dev);
    hipDeviceSynchronize();
    collect_timings(dev);
    pt_release(dev);
    for (auto &kv : dev->scratch) free_stream_scratch(kv.second);
    dev->scratch.clear();
    for (StreamScratch *S : dev->orphans) free_stream_scratch(S);
    dev->orphans.clear();
    if (dev->error) hipFree(dev->error);
    if (dev->arena_nodes) hipFree(dev->arena_nodes);
    if (dev->arena_scratch) hipFree(dev->arena_scratch);
    if (dev->arena_done) hipEventDestroy(dev->arena_done);
    delete dev;
}

uint32_t lc_hip_device_warp_size(uint64_t device) {
    Device *dev = obj<Device>(device);
    if (!dev) return 0;
    int ws = 64;
    hipDeviceGetAttribute(&ws, hipDeviceAttributeWarpSize, dev->index);
    return (uint32_t)ws;
}

int lc_hip_device_native(uint64_t device) {
    Device *dev = obj<Device>(device);
    return dev ? dev->index : -1;
}

int lc_hip_device_query(uint64_t device, const char *key, char *out, size_t capacity) {
    Device *dev = obj<Device>(device);
    if (!dev || !key || !out || capacity == 0) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "bad query arguments");
    std::string v;
    std::string k(key);
    if (k == "device_name") v = dev->name;
    else if (k == "cu_count") v = std::to_string(dev->cu_count);
    else if (k == "warp_size") v = std::to_string(lc_hip_device_warp_size(device));
    else return fail(LC_HIP_ERROR_UNSUPPORTED, "unknown query key");
    std::snprintf(out, capacity, "%s", v.c_str());
    return LC_HIP_OK;
}

uint64_t lc_hip_buffer_create(uint64_t device, size_t element_stride, size_t element_count, void **native_handle) {
    Device *dev = obj<Device>(device);
    if (!dev || element_stride == 0) {
        fail(LC_HIP_ERROR_INVALID_ARGUMENT, "bad buffer arguments");
        return LC_HIP_INVALID_HANDLE;
    }
    DeviceGuard g(dev);
    auto b = std::make_unique<Buffer>();
    b->dev = dev;
    b->elem_stride = element_stride;
    b->bytes = element_stride * element_count;
    if (hipMalloc(&b->ptr, std::max<size_t>(b->bytes, 16)) != hipSuccess) {
        fail(LC_HIP_ERROR_OUT_OF_MEMORY, "hipMalloc failed for buffer");
        return LC_HIP_INVALID_HANDLE;
    }
    if (native_handle) *native_handle = b->ptr;
    return handle(b.release());
}

uint64_t lc_hip_buffer_create_external(uint64_t device, size_t element_stride, size_t element_count, void *device_ptr) {
    Device *dev = obj<Device>(device);
    if (!dev || element_stride == 0 || !device_ptr) {
        fail(LC_HIP_ERROR_INVALID_ARGUMENT, "bad external buffer arguments");
        return LC_HIP_INVALID_HANDLE;
    }
    DeviceGuard g(dev);
    hipPointerAttribute_t attr{};
    if (hipPointerGetAttributes(&attr, device_ptr) != hipSuccess || attr.type != hipMemoryTypeDevice) {
        (void)hipGetLastError();
        fail(LC_HIP_ERROR_INVALID_ARGUMENT, "external memory is not device memory");
        return LC_HIP_INVALID_HANDLE;
    }
    auto b = std::make_unique<Buffer>();
    b->dev = dev;
    b->elem_stride = element_stride;
    b->bytes = element_stride * element_count;
    b->ptr = device_ptr;
    b->owned = false;
    return handle(b.release());
}

void lc_hip_buffer_destroy(uint64_t device, uint64_t buffer) {
    Device *dev = obj<Device>(device);
    Buffer *b = obj<Buffer>(buffer);
    if (!dev || !b) return;
    if (defer_while_capturing(dev, [device, buffer] { lc_hip_buffer_destroy(device, buffer); })) return;
    DeviceGuard g(dev);
    if (b->owned) hipFree(b->ptr);
    delete b;
}

int lc_hip_buffer_upload(uint64_t device, uint64_t stream, uint64_t buffer, size_t offset, size_t size, const void *data) {
    Device *dev = obj<Device>(device);
    Buffer *b = obj<Buffer>(buffer);
    if (!dev || !b || !data) return fail(LC_HIP_ERROR_INVALID_HANDLE, "bad upload handles");
    if (offset + size > b->bytes) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "upload out of buffer range");
    DeviceGuard g(dev);
    hipStream_t s = native_stream(stream);
    // Pageable source: the copy is staged before hipMemcpyAsync returns; synchronizing the
    // stream here makes "consumed at dispatch" hold for any source memory.
    LC_HIP_TRY(hipMemcpyAsync(static_cast<char *>(b->ptr) + offset, data, size, hipMemcpyHostToDevice, s));
    LC_HIP_TRY(hipStreamSynchronize(s));
    return LC_HIP_OK;
}

int lc_hip_buffer_download(uint64_t device, uint64_t stream, uint64_t buffer, size_t offset, size_t size, void *data) {
    Device *dev = obj<Device>(device);
    Buffer *b = obj<Buffer>(buffer);
    if (!dev || !b || !data) return fail(LC_HIP_ERROR_INVALID_HANDLE, "bad download handles");
    if (offset + size > b->bytes) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "download out of buffer range");
    DeviceGuard g(dev);
    LC_HIP_TRY(hipMemcpyAsync(data, static_cast<char *>(b->ptr) + offset, size, hipMemcpyDeviceToHost,
                              native_stream(stream)));
    return LC_HIP_OK;
}

int lc_hip_buffer_copy(uint64_t device, uint64_t stream, uint64_t src, size_t src_offset, uint64_t dst, size_t dst_offset,
                       size_t size) {
    Device *dev = obj<Device>(device);
    Buffer *a = obj<Buffer>(src), *b = obj<Buffer>(dst);
    if (!dev || !a || !b) return fail(LC_HIP_ERROR_INVALID_HANDLE, "bad copy handles");
    if (src_offset + size > a->bytes || dst_offset + size > b->bytes)
        return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "copy out of buffer range");
    DeviceGuard g(dev);
    LC_HIP_TRY(hipMemcpyAsync(static_cast<char *>(b->ptr) + dst_offset, static_cast<char *>(a->ptr) + src_offset, size,
                              hipMemcpyDeviceToDevice, native_stream(stream)));
    return LC_HIP_OK;
}

uint64_t lc_hip_stream_create(uint64_t device, uint32_t) {
    Device *dev = obj<Device>(device);
    if (!dev) {
        fail(LC_HIP_ERROR_INVALID_HANDLE, "bad device");
        return LC_HIP_INVALID_HANDLE;
    }
    DeviceGuard g(dev);
    auto st = std::make_unique<Stream>();
    st->dev = dev;
    if (hipStreamCreateWithFlags(&st->s, hipStreamNonBlocking) != hipSuccess) {
        fail(LC_HIP_ERROR_HIP, "hipStreamCreate failed");
        return LC_HIP_INVALID_HANDLE;
    }
    Stream *raw = st.get();
    st->worker = std::thread([raw] { stream_worker(raw); });
    return handle(st.release());
}

void lc_hip_stream_destroy(uint64_t device, uint64_t stream) {
    Device *dev = obj<Device>(device);
    Stream *st = obj<Stream>(stream);
    if (!dev || !st) return;
    {
        // another stream's capture is open: destroy this one when it ends
        std::lock_guard<std::mutex> dl(dev->mu);
        bool other = false;
        for (auto &kv : dev->capturing) other = other || kv.first != st->s;
        if (other) {
            dev->deferred.push_back([device, stream] { lc_hip_stream_destroy(device, stream); });
            return;
        }
    }
    DeviceGuard g(dev);
    {
        // a capture left open: discard what it recorded
        std::lock_guard<std::mutex> dl(dev->mu);
        auto it = dev->capturing.find(st->s);
        if (it != dev->capturing.end()) {
            hipGraph_t gr = nullptr;
            if (hipStreamEndCapture(st->s, &gr) == hipSuccess && gr) hipGraphDestroy(gr);
            (void)hipGetLastError();
            unpin(dev, it->second);
            dev->capturing.erase(it);
        }
    }
    hipStreamSynchronize(st->s);
    {
        std::lock_guard<std::mutex> lk(st->qmu);
        st->stop = true;
    }
    st->qcv.notify_all();
    if (st->worker.joinable()) st->worker.join();
    pt_release_stream(dev, st->s);
    stream_scratch_release(dev, st->s);
    hipStreamDestroy(st->s);
    delete st;
    run_deferred(dev);
}

int lc_hip_stream_synchronize(uint64_t device, uint64_t stream) {
    Device *dev = obj<Device>(device);
    if (!dev) return fail(LC_HIP_ERROR_INVALID_HANDLE, "bad device");
    DeviceGuard g(dev);
    LC_HIP_TRY(hipStreamSynchronize(native_stream(stream)));
    return LC_HIP_OK;
}

int lc_hip_stream_add_callback(uint64_t device, uint64_t stream, void (*fn)(void *), void *user) {
    Device *dev = obj<Device>(device);
    Stream *st = obj<Stream>(stream);
    if (!dev || !st || !fn) return fail(LC_HIP_ERROR_INVALID_HANDLE, "bad callback arguments");
    DeviceGuard g(dev);
    hipEvent_t e;
    LC_HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    LC_HIP_TRY(hipEventRecord(e, st->s));
    {
        std::lock_guard<std::mutex> lk(st->qmu);
        st->q.emplace_back(e, [fn, user] { fn(user); });
    }
    st->qcv.notify_one();
    return LC_HIP_OK;
}

void *lc_hip_stream_native(uint64_t, uint64_t stream) { return native_stream(stream); }

// DeviceInterface::set_name (include/luisa/runtime/rhi/device_interface.h:189). Meshes, procedural
// primitives, accels and shaders (Resource::Tag MESH 3, PROCEDURAL_PRIMITIVE 5, ACCEL 7, SHADER 10,
// include/luisa/runtime/rhi/resource.h:179-200) keep the name and bracket their builds / dispatches
// with marker ranges "<name>::build" / "::update" / "::compact" / "::dispatch"; other resources
// accept it and ignore it, as backends without a naming facility do.
int lc_hip_set_name(uint64_t device, uint32_t resource_tag, uint64_t handle, const char *name) {
    Device *dev = obj<Device>(device);
    if (!dev || !obj<char>(handle)) return fail(LC_HIP_ERROR_INVALID_HANDLE, "bad set_name handles");
    const std::string n = name ? name : "";
    switch (resource_tag) {
        case 3u:
        case 5u: {
            Mesh *m = obj<Mesh>(handle);
            std::lock_guard<std::mutex> lk(m->mu);
            m->name = n;
            return LC_HIP_OK;
        }
        case 7u: {
            Accel *a = obj<Accel>(handle);
            std::lock_guard<std::mutex> lk(a->mu);
            a->name = n;
            return LC_HIP_OK;
        }
        case 10u: obj<Shader>(handle)->name = n; return LC_HIP_OK;
        default: return LC_HIP_OK;
    }
}

// ---- stream capture into HIP graphs. The reference's CUDA backend encodes every CommandList it
// is given again (CUDAStream::dispatch, src/backends/cuda/cuda_stream.cpp:194-203); a caller
// that submits the same launches many times (a wavefront path tracer's per-sample sequence)
// records them once here and replays the graph. Between begin and end the stream's launches are
// recorded, not run: shader dispatches, traces (unsorted, and only within the working memory the
// stream already holds: run the sequence once outside the capture first) and buffer copies. Host
// synchronisation, callbacks, events and builds are not capturable. A graph replays on any
// stream; its launches use the capturing stream's working memory, so replays must not overlap
// that stream's own traces. ----
int lc_hip_stream_capture_begin(uint64_t device, uint64_t stream) {
    Device *dev = obj<Device>(device);
    Stream *st = obj<Stream>(stream);
    if (!dev || !st) return fail(LC_HIP_ERROR_INVALID_HANDLE, "bad capture handles");
    DeviceGuard g(dev);
    std::lock_guard<std::mutex> dl(dev->mu);
    if (dev->capturing.count(st->s)) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "the stream is already being captured");
    StreamScratch *S = stream_scratch(dev, st->s); // created (and cleared) before the capture
    if (!S) return fail(LC_HIP_ERROR_OUT_OF_MEMORY, "stream scratch");
    LC_HIP_TRY(hipStreamBeginCapture(st->s, hipStreamCaptureModeRelaxed));
    S->pins++;
    dev->capturing.emplace(st->s, S);
    return LC_HIP_OK;
}

static int capture_end(uint64_t device, uint64_t stream, uint64_t *graph);

int lc_hip_stream_capture_end(uint64_t device, uint64_t stream, uint64_t *graph) {
    const int rc = capture_end(device, stream, graph);
    if (Device *dev = obj<Device>(device)) run_deferred(dev);
    return rc;
}

static int capture_end(uint64_t device, uint64_t stream, uint64_t *graph) {
    Device *dev = obj<Device>(device);
    Stream *st = obj<Stream>(stream);
    if (!dev || !st || !graph) return fail(LC_HIP_ERROR_INVALID_HANDLE, "bad capture handles");
    *graph = LC_HIP_INVALID_HANDLE;
    DeviceGuard g(dev);
    std::lock_guard<std::mutex> dl(dev->mu);
    auto it = dev->capturing.find(st->s);
    if (it == dev->capturing.end()) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "the stream is not being captured");
    StreamScratch *S = it->second;
    dev->capturing.erase(it);
    hipGraph_t gr = nullptr;
    hipError_t e = hipStreamEndCapture(st->s, &gr);
    if (e != hipSuccess || !gr) {
        if (gr) hipGraphDestroy(gr);
        unpin(dev, S);
        (void)hipGetLastError();
        return hip_fail(e != hipSuccess ? e : hipErrorStreamCaptureInvalidated, "stream capture");
    }
    hipGraphExec_t ex = nullptr;
    e = hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        hipGraphDestroy(gr);
        unpin(dev, S);
        return hip_fail(e, "graph instantiate");
    }
    auto *G = new Graph;
    G->dev = dev;
    G->graph = gr;
    G->exec = ex;
    G->pinned = S; // the capture's pin passes to the graph
    *graph = handle(G);
    return LC_HIP_OK;
}

int lc_hip_graph_launch(uint64_t device, uint64_t stream, uint64_t graph) {
    Device *dev = obj<Device>(device);
    Graph *G = obj<Graph>(graph);
    if (!dev || !G || G->dev != dev) return fail(LC_HIP_ERROR_INVALID_HANDLE, "bad graph handles");
    DeviceGuard g(dev);
    LC_HIP_TRY(hipGraphLaunch(G->exec, native_stream(stream)));
    return LC_HIP_OK;
}

void lc_hip_graph_destroy(uint64_t device, uint64_t graph) {
    Device *dev = obj<Device>(device);
    Graph *G = obj<Graph>(graph);
    if (!dev || !G || G->dev != dev) return;
    if (defer_while_capturing(dev, [device, graph] { lc_hip_graph_destroy(device, graph); })) return;
    DeviceGuard g(dev);
    hipDeviceSynchronize(); // replays in flight may still name retired buffers
    hipGraphExecDestroy(G->exec);
    hipGraphDestroy(G->graph);
    {
        std::lock_guard<std::mutex> dl(dev->mu);
        unpin(dev, G->pinned);
    }
    delete G;
}

// ---- events: timeline fences. The reference's event is a timeline semaphore
// (src/backends/cuda/cuda_event.cpp:21-63, src/runtime/event.cpp:49-57): signal(stream, v) sets
// the value to v when the stream reaches that point, wait(stream, v) holds the stream until the
// value is at least v, is_completed / synchronize test / wait for that on the host. Here a signal
// of v records a HIP event on its stream; the timeline has reached v once a signal of a value
// >= v has completed. A wait or synchronize for a value no signal >= v has been issued for yet
// blocks the calling thread until another thread issues one (the GPU never spins on a value). ----
struct Event {
    Device *dev;
    std::mutex mu;
    std::condition_variable cv;
    std::map<uint64_t, hipEvent_t> signals; // issued, not yet known complete, by value
    uint64_t issued = 0;                    // highest value signalled so far
    uint64_t reached = 0;                   // highest value known complete
    bool closing = false;
};

// Moves completed signals into `reached` (caller holds e->mu).
static void event_poll(Event *e) {
    for (auto it = e->signals.begin(); it != e->signals.end();) {
        if (hipEventQuery(it->second) == hipSuccess) {
            e->reached = std::max(e->reached, it->first);
            hipEventDestroy(it->second);
            it = e->signals.erase(it);
        } else {
            ++it;
        }
    }
    // signals at or below the reached value add nothing
    for (auto it = e->signals.begin(); it != e->signals.end() && it->first <= e->reached;) {
        hipEventDestroy(it->second);
        it = e->signals.erase(it);
    }
}

// Blocks until a signal of a value >= v has been issued; returns the HIP event of the smallest
// such signal, or null when the timeline has already reached v (caller holds lk on e->mu).
static hipEvent_t event_await_issue(Event *e, uint64_t v, std::unique_lock<std::mutex> &lk) {
    e->cv.wait(lk, [&] { return e->closing || e->issued >= v || e->reached >= v; });
    event_poll(e);
    if (e->reached >= v) return nullptr;
    auto it = e->signals.lower_bound(v);
    return it == e->signals.end() ? nullptr : it->second;
}

uint64_t lc_hip_event_create(uint64_t device) {
    Device *dev = obj<Device>(device);
    if (!dev) {
        fail(LC_HIP_ERROR_INVALID_HANDLE, "bad device");
        return LC_HIP_INVALID_HANDLE;
    }
    auto *e = new Event;
    e->dev = dev;
    return handle(e);
}

void lc_hip_event_destroy(uint64_t device, uint64_t event) {
    Event *e = obj<Event>(event);
    if (!e || !obj<Device>(device)) return;
    if (defer_while_capturing(e->dev, [device, event] { lc_hip_event_destroy(device, event); })) return;
    DeviceGuard g(e->dev);
    {
        std::lock_guard<std::mutex> lk(e->mu);
        e->closing = true;
        for (auto &kv : e->signals) {
            hipEventSynchronize(kv.second);
            hipEventDestroy(kv.second);
        }
        e->signals.clear();
    }
    e->cv.notify_all();
    delete e;
}

int lc_hip_event_signal(uint64_t device, uint64_t event, uint64_t stream, uint64_t value) {
    Event *e = obj<Event>(event);
    if (!e || !obj<Device>(device)) return fail(LC_HIP_ERROR_INVALID_HANDLE, "bad event");
    DeviceGuard g(e->dev);
    hipEvent_t h;
    LC_HIP_TRY(hipEventCreateWithFlags(&h, hipEventDisableTiming));
    {
        hipError_t r = hipEventRecord(h, native_stream(stream));
        if (r != hipSuccess) {
            hipEventDestroy(h);
            return hip_fail(r, "hipEventRecord");
        }
    }
    {
        std::lock_guard<std::mutex> lk(e->mu);
        auto it = e->signals.find(value);
        if (it != e->signals.end()) {
            hipEventDestroy(it->second);
            it->second = h;
        } else {
            e->signals.emplace(value, h);
        }
        e->issued = std::max(e->issued, value);
        event_poll(e);
    }
    e->cv.notify_all();
    return LC_HIP_OK;
}

int lc_hip_event_wait(uint64_t device, uint64_t event, uint64_t stream, uint64_t value) {
    Event *e = obj<Event>(event);
    if (!e || !obj<Device>(device)) return fail(LC_HIP_ERROR_INVALID_HANDLE, "bad event");
    DeviceGuard g(e->dev);
    std::unique_lock<std::mutex> lk(e->mu);
    hipEvent_t h = event_await_issue(e, value, lk);
    if (h) LC_HIP_TRY(hipStreamWaitEvent(native_stream(stream), h, 0));
    return LC_HIP_OK;
}

int lc_hip_event_synchronize(uint64_t device, uint64_t event, uint64_t value) {
    Event *e = obj<Event>(event);
    if (!e || !obj<Device>(device)) return fail(LC_HIP_ERROR_INVALID_HANDLE, "bad event");
    DeviceGuard g(e->dev);
    std::unique_lock<std::mutex> lk(e->mu);
    hipEvent_t h = event_await_issue(e, value, lk);
    if (!h) return LC_HIP_OK;
    LC_HIP_TRY(hipEventSynchronize(h));
    event_poll(e);
    return LC_HIP_OK;
}

int lc_hip_event_is_completed(uint64_t device, uint64_t event, uint64_t value) {
    Event *e = obj<Event>(event);
    if (!e || !obj<Device>(device)) return fail(LC_HIP_ERROR_INVALID_HANDLE, "bad event");
    DeviceGuard g(e->dev);
    std::lock_guard<std::mutex> lk(e->mu);
    event_poll(e);
    return e->reached >= value ? 1 : 0;
}

// ---- meshes ----
uint64_t lc_hip_mesh_create(uint64_t device, const lc_hip_accel_option *option) {
    Device *dev = obj<Device>(device);
    if (!dev) {
        fail(LC_HIP_ERROR_INVALID_HANDLE, "bad device");
        return LC_HIP_INVALID_HANDLE;
    }
    DeviceGuard g(dev);
    auto m = std::make_unique<Mesh>();
    m->dev = dev;
    if (option) {
        m->hint = option->hint;
        m->allow_compaction = option->allow_compaction != 0;
        m->allow_update = option->allow_update != 0;
    }
    if (hipHostMalloc((void **)&m->host, sizeof(HostCounters), hipHostMallocDefault) != hipSuccess ||
        hipEventCreateWithFlags(&m->ready, hipEventDisableTiming) != hipSuccess) {
        fail(LC_HIP_ERROR_OUT_OF_MEMORY, "mesh host counters");
        return LC_HIP_INVALID_HANDLE;
    }
    return handle(m.release());
}

void lc_hip_mesh_destroy(uint64_t device, uint64_t mesh) {
    Device *dev = obj<Device>(device);
    Mesh *m = obj<Mesh>(mesh);
    if (!dev || !m) return;
    if (defer_while_capturing(dev, [device, mesh] { lc_hip_mesh_destroy(device, mesh); })) return;
    DeviceGuard g(dev);
    hipDeviceSynchronize();
    if (m->nodes) hipFree(m->nodes);
    if (m->tris) hipFree(m->tris);
    if (m->proc) hipFree(m->proc);
    if (m->proc_boxes) hipFree(m->proc_boxes);
    if (m->proc_ids) hipFree(m->proc_ids);
    if (m->scratch) hipFree(m->scratch);
    if (m->refit_state) hipFree(m->refit_state);
    if (m->host) hipHostFree(m->host);
    if (m->ready) hipEventDestroy(m->ready);
    delete m;
}

// ---- procedural primitives: create_procedural_primitive (device_interface.h:175-176) and
// ProceduralPrimitiveBuildCommand (command.h:489-510): a BLAS over AABBs, one per leaf slot ----
uint64_t lc_hip_procedural_create(uint64_t device, const lc_hip_accel_option *option) {
    const uint64_t h = lc_hip_mesh_create(device, option);
    if (h != LC_HIP_INVALID_HANDLE) obj<Mesh>(h)->procedural = true;
    return h;
}

void lc_hip_procedural_destroy(uint64_t device, uint64_t primitive) { lc_hip_mesh_destroy(device, primitive); }

int lc_hip_procedural_build(uint64_t device, uint64_t stream, uint64_t primitive, uint32_t, uint64_t aabb_buffer,
                            size_t aabb_offset, size_t aabb_size) {
    Device *dev = obj<Device>(device);
    Mesh *m = obj<Mesh>(primitive);
    Buffer *ab = obj<Buffer>(aabb_buffer);
    if (!dev || !m || !ab) return fail(LC_HIP_ERROR_INVALID_HANDLE, "bad procedural build handles");
    if (!m->procedural) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "not a procedural primitive");
    if (aabb_offset % 4 != 0 || aabb_size % 24 != 0)
        return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "AABB range must be 4-byte aligned and a multiple of 24 bytes");
    if (aabb_offset + aabb_size > ab->bytes) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "AABB range outside its buffer");
    const uint64_t n64 = aabb_size / 24;
    if (n64 >= 0x7FFFFFFFull) return fail(LC_HIP_ERROR_UNSUPPORTED, "more than 2^31 primitives");
    const uint32_t n = (uint32_t)n64;
    std::lock_guard<std::mutex> lk(m->mu);
    DeviceGuard g(dev);
    hipStream_t s = native_stream(stream);
    if (!m->allow_compaction) {
        void *p = m->nodes;
        size_t cap = m->node_cap * sizeof(WideNode);
        LC_HIP_TRY(ensure(p, cap, bvh_node_capacity(n) * sizeof(WideNode)));
        m->nodes = static_cast<WideNode *>(p); m->node_cap = cap / sizeof(WideNode);
        LC_HIP_TRY(ensure(m->scratch, m->scratch_bytes, bvh_scratch_bytes(n)));
    }
    const size_t pn = std::max<size_t>(n, 1);
    if (pn > m->proc_cap) {
        if (m->proc) hipFree(m->proc);
        if (m->proc_boxes) hipFree(m->proc_boxes);
        if (m->proc_ids) hipFree(m->proc_ids);
        m->proc = nullptr; m->proc_boxes = nullptr; m->proc_ids = nullptr; m->proc_cap = 0;
        LC_HIP_TRY(hipMalloc(&m->proc, pn * sizeof(ProcRecord)));
        LC_HIP_TRY(hipMalloc(&m->proc_boxes, 2 * pn * sizeof(float4)));
        LC_HIP_TRY(hipMalloc(&m->proc_ids, pn * sizeof(uint32_t)));
        m->proc_cap = pn;
    }
    m->ntris = n;
    m->built = true;
    m->stats_valid = false;
    m->build_epoch++;
    if (n == 0) return LC_HIP_OK;
    {
        MarkerRange mr(m->name, "build");
        std::lock_guard<std::mutex> dl(dev->mu);
        BuildTarget t{m->nodes, m->scratch, m->scratch_bytes};
        if (m->allow_compaction) LC_HIP_TRY(arena_acquire(dev, m, n, t));
        StreamScratch *S = stream_scratch(dev, s);
        if (!S) return fail(LC_HIP_ERROR_OUT_OF_MEMORY, "stream scratch");
        LC_HIP_TRY(hipMemsetAsync(S->error, 0, sizeof(uint32_t), s));
        TimedRegion tr(dev, s, "procedural_build");
        const float *aabb = reinterpret_cast<const float *>(static_cast<const char *>(ab->ptr) + aabb_offset);
        LC_HIP_TRY(aabbs_to_boxes(s, aabb, n, m->proc_boxes));
        BuildCounters *ctr = nullptr;
        hipError_t e = bvh_build(s, t.scratch, t.scratch_bytes, nullptr, 0, 0, nullptr, m->proc_boxes, n, t.nodes,
                                 nullptr, m->proc_ids, &ctr, 0, 1u, sah_for(dev, m->hint), dev->bvh_width, dev->cu_count,
                                 S->error, dev->collapse_head);
        if (e != hipSuccess) return hip_fail(e, "procedural bvh_build");
        LC_HIP_TRY(proc_records(s, m->proc_ids, m->proc_boxes, n, m->proc));
        copy_counters_async(ctr, m->host, S->error, s);
        if (m->allow_compaction) {
            const int rc = arena_compact(dev, m, s, nullptr);
            if (rc != LC_HIP_OK) return rc;
        }
    }
    LC_HIP_TRY(hipEventRecord(m->ready, s));
    return LC_HIP_OK;
}

int lc_hip_mesh_build(uint64_t device, uint64_t stream, uint64_t mesh, uint32_t request, uint64_t vertex_buffer,
                      size_t vertex_offset, size_t vertex_size, size_t vertex_stride, uint64_t triangle_buffer,
                      size_t triangle_offset, size_t triangle_size) {
    Device *dev = obj<Device>(device);
    Mesh *m = obj<Mesh>(mesh);
    Buffer *vb = obj<Buffer>(vertex_buffer), *tb = obj<Buffer>(triangle_buffer);
    if (!dev || !m || !vb || !tb) return fail(LC_HIP_ERROR_INVALID_HANDLE, "bad mesh build handles");
    if (m->procedural) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "a procedural primitive is built from AABBs");
    // src/runtime/rtx/mesh.cpp:29-49
    if (vertex_stride == 0 || vertex_stride % 16 != 0) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "vertex stride must be a multiple of 16");
    if (vertex_offset % vertex_stride != 0 || vertex_size % vertex_stride != 0)
        return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "vertex range not aligned to the stride");
    if (triangle_offset % 4 != 0 || triangle_size % 12 != 0)
        return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "triangle range not aligned to uint3");
    if (vertex_offset + vertex_size > vb->bytes || triangle_offset + triangle_size > tb->bytes)
        return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "mesh range outside its buffer");
    uint64_t ntris64 = triangle_size / 12, nverts = vertex_size / vertex_stride;
    if (ntris64 >= 0x7FFFFFFFull) return fail(LC_HIP_ERROR_UNSUPPORTED, "more than 2^31 triangles");
    uint32_t n = (uint32_t)ntris64;
    if (n > 0 && nverts == 0) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "triangles without vertices");
    std::lock_guard<std::mutex> lk(m->mu);
    DeviceGuard g(dev);
    hipStream_t s = native_stream(stream);
    const char *vaddr = static_cast<const char *>(vb->ptr) + vertex_offset;
    const uint32_t *taddr = reinterpret_cast<const uint32_t *>(static_cast<const char *>(tb->ptr) + triangle_offset);
    // update (refit) iff built, allowed, requested and over the same buffer ranges
    // (src/backends/cuda/cuda_mesh.cpp:41-60); anything else rebuilds
    const bool refit = request == LC_HIP_BUILD_PREFER_UPDATE && m->built && m->allow_update && n > 0 &&
                       m->ntris == n && vaddr == m->v_addr && vertex_size == m->v_size &&
                       vertex_stride == m->v_stride && taddr == m->t_addr && triangle_size == m->t_size;
    m->v_addr = vaddr;
    m->v_size = vertex_size;
    m->v_stride = vertex_stride;
    m->t_addr = taddr;
    m->t_size = triangle_size;
    if (refit) {
        MarkerRange mr(m->name, "update");
        const int rc = mesh_stats_sync(m); // the wide depth of the build being refitted
        if (rc != LC_HIP_OK) return rc;
        const uint32_t depth = m->stats.depth;
        m->stats_valid = false;
        m->build_epoch++;
        {
            std::lock_guard<std::mutex> dl(dev->mu);
            StreamScratch *S = stream_scratch(dev, s);
            if (!S) return fail(LC_HIP_ERROR_OUT_OF_MEMORY, "stream scratch");
            LC_HIP_TRY(hipMemsetAsync(S->error, 0, sizeof(uint32_t), s));
            TimedRegion tr(dev, s, "mesh_refit");
            BuildCounters *ctr = nullptr;
            hipError_t e;
            if (m->allow_compaction) { // the refit state kept by the compacting build
                ctr = bvh_refit_state_counters(m->refit_state);
                e = bvh_refit_from_state(s, m->refit_state, m->refit_state_bytes, m->refit_nodes, vaddr, vertex_stride,
                                         nverts, taddr, n, depth, m->nodes, m->tris, dev->cu_count, S->error);
            } else {
                e = bvh_refit(s, m->scratch, m->scratch_bytes, vaddr, vertex_stride, nverts, taddr, n, depth, m->nodes,
                              m->tris, &ctr, dev->cu_count, S->error);
            }
            if (e != hipSuccess) return hip_fail(e, "bvh_refit");
            copy_counters_async(ctr, m->host, S->error, s);
        }
        m->refits++;
        LC_HIP_TRY(hipEventRecord(m->ready, s));
        return LC_HIP_OK;
    }
    void *p;
    size_t cap;
    if (!m->allow_compaction) {
        p = m->nodes; cap = m->node_cap * sizeof(WideNode);
        LC_HIP_TRY(ensure(p, cap, bvh_node_capacity(n) * sizeof(WideNode)));
        m->nodes = static_cast<WideNode *>(p); m->node_cap = cap / sizeof(WideNode);
        LC_HIP_TRY(ensure(m->scratch, m->scratch_bytes, bvh_scratch_bytes(n)));
    }
    p = m->tris; cap = m->tri_cap * sizeof(TriRecord);
    LC_HIP_TRY(ensure(p, cap, std::max<size_t>(n, 1) * sizeof(TriRecord)));
    m->tris = static_cast<TriRecord *>(p); m->tri_cap = cap / sizeof(TriRecord);
    m->ntris = n;
    m->built = true;
    m->stats_valid = false;
    m->build_epoch++;
    if (n == 0) return LC_HIP_OK;
    {
        MarkerRange mr(m->name, "build");
        std::lock_guard<std::mutex> dl(dev->mu);
        BuildTarget t{m->nodes, m->scratch, m->scratch_bytes};
        if (m->allow_compaction) LC_HIP_TRY(arena_acquire(dev, m, n, t));
        StreamScratch *S = stream_scratch(dev, s);
        if (!S) return fail(LC_HIP_ERROR_OUT_OF_MEMORY, "stream scratch");
        LC_HIP_TRY(hipMemsetAsync(S->error, 0, sizeof(uint32_t), s));
        TimedRegion tr(dev, s, "mesh_build");
        BuildCounters *ctr = nullptr;
        hipError_t e = bvh_build(s, t.scratch, t.scratch_bytes, vaddr, vertex_stride, nverts, taddr, nullptr, n,
                                 t.nodes, m->tris, nullptr, &ctr, dev->split_small, leaf_size_for(m->hint),
                                 sah_for(dev, m->hint), dev->bvh_width, dev->cu_count, S->error, dev->collapse_head);
        if (e != hipSuccess) return hip_fail(e, "bvh_build");
        copy_counters_async(ctr, m->host, S->error, s);
        if (m->allow_compaction) {
            const int rc = arena_compact(dev, m, s, ctr);
            if (rc != LC_HIP_OK) return rc;
        }
    }
    LC_HIP_TRY(hipEventRecord(m->ready, s));
    return LC_HIP_OK;
}

int lc_hip_mesh_stats(uint64_t device, uint64_t mesh, lc_hip_bvh_stats *out) {
    Device *dev = obj<Device>(device);
    Mesh *m = obj<Mesh>(mesh);
    if (!dev || !m || !out) return fail(LC_HIP_ERROR_INVALID_HANDLE, "bad mesh stats handles");
    DeviceGuard g(dev);
    int rc = mesh_stats_sync(m);
    if (rc != LC_HIP_OK) return rc;
    out->node_count = m->stats.node_count;
    out->leaf_primitive_count = m->stats.tri_count;
    out->depth = m->stats.depth;
    out->primitive_count = m->ntris;
    std::memcpy(out->bounds, m->stats.bounds, sizeof(out->bounds));
    out->node_bytes = (uint64_t)m->stats.node_count * sizeof(WideNode);
    out->primitive_bytes = (uint64_t)m->stats.tri_count * sizeof(TriRecord);
    return LC_HIP_OK;
}

int lc_hip_mesh_memory(uint64_t device, uint64_t mesh, lc_hip_memory_usage *out) {
    Device *dev = obj<Device>(device);
    Mesh *m = obj<Mesh>(mesh);
    if (!dev || !m || !out) return fail(LC_HIP_ERROR_INVALID_HANDLE, "bad mesh memory handles");
    std::lock_guard<std::mutex> lk(m->mu);
    out->node_bytes = (uint64_t)m->node_cap * sizeof(WideNode);
    out->primitive_bytes = (uint64_t)m->tri_cap * sizeof(TriRecord) +
                           (uint64_t)m->proc_cap * (sizeof(ProcRecord) + 2 * sizeof(float4) + sizeof(uint32_t));
    out->scratch_bytes = (m->scratch ? (uint64_t)m->scratch_bytes : 0) + (m->refit_state ? (uint64_t)m->refit_state_bytes : 0);
    out->total_bytes = out->node_bytes + out->primitive_bytes + out->scratch_bytes;
    return LC_HIP_OK;
}

uint64_t lc_hip_device_build_arena_bytes(uint64_t device) {
    Device *dev = obj<Device>(device);
    if (!dev) return 0;
    std::lock_guard<std::mutex> dl(dev->mu);
    return (uint64_t)dev->arena_node_cap * sizeof(WideNode) + (dev->arena_scratch ? dev->arena_scratch_bytes : 0);
}

int lc_hip_device_release_build_arena(uint64_t device) {
    Device *dev = obj<Device>(device);
    if (!dev) return fail(LC_HIP_ERROR_INVALID_HANDLE, "bad device handle");
    DeviceGuard g(dev);
    std::lock_guard<std::mutex> dl(dev->mu);
    LC_HIP_TRY(hipEventSynchronize(dev->arena_done));
    if (dev->arena_nodes) LC_HIP_TRY(hipFree(dev->arena_nodes));
    dev->arena_nodes = nullptr;
    dev->arena_node_cap = 0;
    if (dev->arena_scratch) LC_HIP_TRY(hipFree(dev->arena_scratch));
    dev->arena_scratch = nullptr;
    dev->arena_scratch_bytes = 0;
    return LC_HIP_OK;
}

int lc_hip_mesh_validate(uint64_t device, uint64_t stream, uint64_t mesh, uint32_t errors[7]) {
    Device *dev = obj<Device>(device);
    Mesh *m = obj<Mesh>(mesh);
    if (!dev || !m || !errors) return fail(LC_HIP_ERROR_INVALID_HANDLE, "bad mesh validate handles");
    DeviceGuard g(dev);
    int rc = mesh_stats_sync(m);
    if (rc != LC_HIP_OK) return rc;
    std::memset(errors, 0, 7 * sizeof(uint32_t));
    if (m->ntris == 0) return LC_HIP_OK;
    hipStream_t s = native_stream(stream);
    void *scratch = nullptr;
    LC_HIP_TRY(hipMalloc(&scratch, sizeof(uint32_t) * (2ull * m->ntris + 16)));
    uint32_t *err = static_cast<uint32_t *>(scratch);
    uint32_t *marks = err + 8;
    hipError_t e = bvh_validate(s, m->nodes, m->stats.node_count, m->tris, m->stats.tri_count, m->ntris, marks, err);
    if (e == hipSuccess) e = hipMemcpyAsync(errors, err, 7 * sizeof(uint32_t), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    hipFree(scratch);
    if (e != hipSuccess) return hip_fail(e, "bvh validate");
    return LC_HIP_OK;
}

// ---- accels ----
uint64_t lc_hip_accel_create(uint64_t device, const lc_hip_accel_option *option) {
    Device *dev = obj<Device>(device);
    if (!dev) {
        fail(LC_HIP_ERROR_INVALID_HANDLE, "bad device");
        return LC_HIP_INVALID_HANDLE;
    }
    DeviceGuard g(dev);
    auto a = std::make_unique<Accel>();
    a->dev = dev;
    if (option) {
        a->hint = option->hint;
        a->allow_compaction = option->allow_compaction != 0;
        a->allow_update = option->allow_update != 0;
    }
    if (hipHostMalloc((void **)&a->tlas_host, sizeof(HostCounters), hipHostMallocDefault) != hipSuccess ||
        hipEventCreateWithFlags(&a->tlas_ready, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&a->staged, hipEventDisableTiming) != hipSuccess) {
        fail(LC_HIP_ERROR_OUT_OF_MEMORY, "accel host counters");
        return LC_HIP_INVALID_HANDLE;
    }
    return handle(a.release());
}

void lc_hip_accel_destroy(uint64_t device, uint64_t accel) {
    Device *dev = obj<Device>(device);
    Accel *a = obj<Accel>(accel);
    if (!dev || !a) return;
    if (defer_while_capturing(dev, [device, accel] { lc_hip_accel_destroy(device, accel); })) return;
    DeviceGuard g(dev);
    hipDeviceSynchronize();
    if (a->records) hipFree(a->records);
    if (a->boxes) hipFree(a->boxes);
    if (a->tlas_nodes) hipFree(a->tlas_nodes);
    if (a->tlas_prims) hipFree(a->tlas_prims);
    if (a->tlas_scratch) hipFree(a->tlas_scratch);
    if (a->tlas_host) hipHostFree(a->tlas_host);
    if (a->tlas_ready) hipEventDestroy(a->tlas_ready);
    if (a->staged) hipEventDestroy(a->staged);
    if (a->stage) hipHostFree(a->stage);
    if (a->state) hipFree(a->state);
    if (a->desc) hipFree(a->desc);
    delete a;
}

// Folds instance writes issued on the device (lc_hip_accel_instance_write) into the host
// instance set. Called with the accel locked and its device current.
static int fold_instance_writes(Accel *a) {
    if (!a->state_written || !a->state) return LC_HIP_OK;
    LC_HIP_TRY(hipDeviceSynchronize());
    std::vector<InstanceState> st((size_t)a->state_count + 1);
    LC_HIP_TRY(hipMemcpy(st.data(), a->state, st.size() * sizeof(InstanceState), hipMemcpyDeviceToHost));
    a->state_written = false;
    uint32_t err = 0;
    std::memcpy(&err, &st[a->state_count], sizeof(err));
    for (uint32_t i = 0; i < a->state_count && i < a->inst.size(); i++) {
        const InstanceState &x = st[i];
        if (x.dirty & (1u << kInstanceTransform)) std::memcpy(a->inst[i].affine, x.affine, sizeof(float) * 12);
        if (x.dirty & (1u << kInstanceUserId)) a->inst[i].user_id = x.user_id;
        if (x.dirty & (1u << kInstanceVisibility)) a->inst[i].vis_mask = x.vis_mask;
    }
    if (err) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "an instance read or write named an instance out of range");
    return LC_HIP_OK;
}

static void xform_point_h(const float m[12], const float p[3], float out[3]) {
    for (int r = 0; r < 3; r++) out[r] = m[4 * r + 0] * p[0] + m[4 * r + 1] * p[1] + m[4 * r + 2] * p[2] + m[4 * r + 3];
}

// Pinned staging of at least `bytes` for a build's uploads, after the previous build's copies
// out of it have run.
static int stage_reserve(Accel *a, size_t bytes) {
    if (a->stage_pending) {
        LC_HIP_TRY(hipEventSynchronize(a->staged));
        a->stage_pending = false;
    }
    if (a->stage_cap >= bytes) return LC_HIP_OK;
    if (a->stage) hipHostFree(a->stage);
    a->stage = nullptr;
    a->stage_cap = 0;
    const size_t cap = std::max<size_t>(bytes, 4096);
    LC_HIP_TRY(hipHostMalloc((void **)&a->stage, cap, hipHostMallocDefault));
    a->stage_cap = cap;
    return LC_HIP_OK;
}

static size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// The request is not consulted: an update rebuilds the TLAS over the instance boxes (a few
// thousand instances at most; the boxes move with the transforms anyway).
int lc_hip_accel_build(uint64_t device, uint64_t stream, uint64_t accel, uint32_t, uint32_t instance_count,
                       const lc_hip_accel_modification *mods, size_t nmods, int update_instance_buffer_only) {
    Device *dev = obj<Device>(device);
    Accel *a = obj<Accel>(accel);
    if (!dev || !a) return fail(LC_HIP_ERROR_INVALID_HANDLE, "bad accel build handles");
    if (nmods > 0 && !mods) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "null modifications");
    std::lock_guard<std::mutex> lk(a->mu);
    MarkerRange mr(a->name, update_instance_buffer_only ? "update" : "build");
    DeviceGuard g(dev);
    hipStream_t s = native_stream(stream);
    // instance writes issued on the device since the last build come first (cpu/accel.rs:335-343
    // applies dirty instances before the modifications)
    {
        const int rc = fold_instance_writes(a);
        if (rc != LC_HIP_OK) return rc;
    }
    // AccelImpl::update (cpu/accel.rs:324-447): resize, then apply each modification in order.
    a->inst.resize(instance_count);
    for (size_t k = 0; k < nmods; k++) {
        const auto &md = mods[k];
        if (md.index >= instance_count) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "modification index out of range");
        InstanceHost &in = a->inst[md.index];
        if (md.flags & LC_HIP_MOD_PRIMITIVE) {
            Mesh *m = obj<Mesh>(md.primitive);
            if (!m) return fail(LC_HIP_ERROR_INVALID_HANDLE, "instance primitive is not a mesh");
            if (!m->built) return fail(LC_HIP_ERROR_NOT_BUILT, "Mesh not built");
            in.mesh = m;
            std::memcpy(in.affine, md.affine, sizeof(in.affine));
            in.vis_mask = 0xFFu;
            in.user_id = md.user_id;
            in.opaque = true;
        }
        if (md.flags & LC_HIP_MOD_OPAQUE_ON) in.opaque = true;
        else if (md.flags & LC_HIP_MOD_OPAQUE_OFF) in.opaque = false;
        if (md.flags & LC_HIP_MOD_TRANSFORM) std::memcpy(in.affine, md.affine, sizeof(in.affine));
        if (md.flags & LC_HIP_MOD_VISIBILITY) in.vis_mask = md.vis_mask;
        if (md.flags & LC_HIP_MOD_USER_ID) in.user_id = md.user_id;
    }
    // instance records (need each BLAS's depth and bounds: wait for its counters)
    a->host_records.resize(instance_count);
    uint32_t max_depth = 0;
    for (uint32_t i = 0; i < instance_count; i++) {
        InstanceHost &in = a->inst[i];
        InstanceRecord &r = a->host_records[i];
        std::memset(&r, 0, sizeof(r));
        std::memcpy(r.fwd, in.affine, sizeof(r.fwd));
        invert_affine(in.affine, r.inv);
        r.vis_mask = in.vis_mask;
        r.user_id = in.user_id;
        r.flags = in.opaque ? kInstOpaque : 0u;
        if (in.mesh) {
            int rc = mesh_stats_sync(in.mesh);
            if (rc != LC_HIP_OK) return rc;
            if (in.mesh->ntris > 0) {
                r.blas_nodes = reinterpret_cast<uint64_t>(in.mesh->nodes);
                r.blas_tris = in.mesh->procedural ? reinterpret_cast<uint64_t>(in.mesh->proc)
                                                  : reinterpret_cast<uint64_t>(in.mesh->tris);
                if (in.mesh->procedural) r.flags |= kInstProcedural;
                r.blas_depth = in.mesh->stats.depth;
                max_depth = std::max(max_depth, r.blas_depth);
            }
        }
    }
    a->max_blas_depth = max_depth;
    // staging layout: records | instance table | TLAS boxes | descriptor
    const size_t rec_bytes = (size_t)instance_count * sizeof(InstanceRecord);
    const size_t state_bytes = ((size_t)instance_count + 1) * sizeof(InstanceState);
    const size_t box_bytes = 2 * sizeof(float4) * (size_t)instance_count;
    const size_t off_state = align_up(rec_bytes, 256);
    const size_t off_box = off_state + align_up(state_bytes, 256);
    const size_t off_desc = off_box + align_up(box_bytes, 256);
    {
        const int rc = stage_reserve(a, off_desc + sizeof(DeviceAccel));
        if (rc != LC_HIP_OK) return rc;
    }
    void *p = a->records;
    size_t cap = a->record_cap * sizeof(InstanceRecord);
    LC_HIP_TRY(ensure(p, cap, std::max<size_t>(instance_count, 1) * sizeof(InstanceRecord)));
    a->records = static_cast<InstanceRecord *>(p);
    a->record_cap = cap / sizeof(InstanceRecord);
    if (instance_count) {
        std::memcpy(a->stage, a->host_records.data(), rec_bytes);
        LC_HIP_TRY(hipMemcpyAsync(a->records, a->stage, rec_bytes, hipMemcpyHostToDevice, s));
    }
    // the device-side instance table the in-kernel instance ops see (state[count] = error word)
    {
        void *sp = a->state;
        size_t scap = a->state_cap * sizeof(InstanceState);
        LC_HIP_TRY(ensure(sp, scap, state_bytes));
        a->state = static_cast<InstanceState *>(sp);
        a->state_cap = scap / sizeof(InstanceState);
        InstanceState *st = reinterpret_cast<InstanceState *>(a->stage + off_state);
        for (uint32_t i = 0; i <= instance_count; i++) st[i] = InstanceState{};
        for (uint32_t i = 0; i < instance_count; i++) {
            std::memcpy(st[i].affine, a->inst[i].affine, sizeof(float) * 12);
            st[i].user_id = a->inst[i].user_id;
            st[i].vis_mask = a->inst[i].vis_mask;
        }
        LC_HIP_TRY(hipMemcpyAsync(a->state, st, state_bytes, hipMemcpyHostToDevice, s));
        a->state_count = instance_count;
        a->state_written = false;
    }
    // update_instance_buffer_only (command.h:531-616) refreshes the instance records over the TLAS
    // of the last build; when the instance count differs from that build's (or nothing was built)
    // that TLAS does not describe the instances, so the accel is built in full instead
    if (update_instance_buffer_only && a->built && instance_count == a->tlas_count) {
        const int rc = accel_write_descriptor(a, s, a->stage + off_desc);
        if (rc != LC_HIP_OK) return rc;
        LC_HIP_TRY(hipEventRecord(a->staged, s));
        a->stage_pending = true;
        return LC_HIP_OK;
    }
    a->built = true;
    a->tlas_count = instance_count;
    a->tlas_valid = false;
    a->tlas_stats_valid = false;
    // world boxes of instances: 8 transformed BLAS-box corners, padded by a relative epsilon
    float *boxes = reinterpret_cast<float *>(a->stage + off_box);
    {
        float wlo[3] = {3e38f, 3e38f, 3e38f}, whi[3] = {-3e38f, -3e38f, -3e38f};
        for (uint32_t i = 0; i < instance_count; i++) {
            float lo[3] = {3e38f, 3e38f, 3e38f}, hi[3] = {-3e38f, -3e38f, -3e38f};
            const InstanceHost &in = a->inst[i];
            if (in.mesh && in.mesh->ntris > 0) {
                const float *b = in.mesh->stats.bounds;
                for (int c = 0; c < 8; c++) {
                    float q[3] = {(c & 1) ? b[3] : b[0], (c & 2) ? b[4] : b[1], (c & 4) ? b[5] : b[2]}, w[3];
                    xform_point_h(in.affine, q, w);
                    for (int k = 0; k < 3; k++) { lo[k] = std::min(lo[k], w[k]); hi[k] = std::max(hi[k], w[k]); }
                }
                for (int k = 0; k < 3; k++) {
                    float pad = 1e-5f * std::max({std::fabs(lo[k]), std::fabs(hi[k]), 1.0f});
                    lo[k] -= pad;
                    hi[k] += pad;
                    wlo[k] = std::min(wlo[k], lo[k]);
                    whi[k] = std::max(whi[k], hi[k]);
                }
            } else {
                // no geometry: a degenerate box far outside any scene (never entered: blas_nodes == 0)
                for (int k = 0; k < 3; k++) { lo[k] = hi[k] = 3e38f; }
            }
            float *o = &boxes[8 * (size_t)i];
            o[0] = lo[0]; o[1] = lo[1]; o[2] = lo[2]; o[3] = 0.0f;
            o[4] = hi[0]; o[5] = hi[1]; o[6] = hi[2]; o[7] = 0.0f;
        }
        for (int k = 0; k < 3; k++) {
            a->world_bounds[k] = wlo[k] <= whi[k] ? wlo[k] : 0.0f;
            a->world_bounds[3 + k] = wlo[k] <= whi[k] ? whi[k] : 0.0f;
        }
    }
    if (instance_count > 1) {
        p = a->boxes; cap = a->box_cap;
        LC_HIP_TRY(ensure(p, cap, box_bytes));
        a->boxes = static_cast<float4 *>(p); a->box_cap = cap;
        LC_HIP_TRY(hipMemcpyAsync(a->boxes, boxes, box_bytes, hipMemcpyHostToDevice, s));
        size_t ncap = bvh_node_capacity(instance_count);
        if (a->tlas_cap < ncap) {
            if (a->tlas_nodes) hipFree(a->tlas_nodes);
            if (a->tlas_prims) hipFree(a->tlas_prims);
            a->tlas_nodes = nullptr;
            a->tlas_prims = nullptr;
            a->tlas_cap = 0;
            LC_HIP_TRY(hipMalloc(&a->tlas_nodes, ncap * sizeof(WideNode)));
            LC_HIP_TRY(hipMalloc(&a->tlas_prims, ncap * sizeof(uint32_t)));
            a->tlas_cap = ncap;
        }
        LC_HIP_TRY(ensure(a->tlas_scratch, a->tlas_scratch_bytes, bvh_scratch_bytes(instance_count)));
        {
            std::lock_guard<std::mutex> dl(dev->mu);
            StreamScratch *S = stream_scratch(dev, s);
            if (!S) return fail(LC_HIP_ERROR_OUT_OF_MEMORY, "stream scratch");
            LC_HIP_TRY(hipMemsetAsync(S->error, 0, sizeof(uint32_t), s));
            TimedRegion tr(dev, s, "accel_build");
            BuildCounters *ctr = nullptr;
            hipError_t e = bvh_build(s, a->tlas_scratch, a->tlas_scratch_bytes, nullptr, 0, 0, nullptr, a->boxes,
                                     instance_count, a->tlas_nodes, nullptr, a->tlas_prims, &ctr, 0,
                                     leaf_size_for(a->hint), sah_for(dev, a->hint), dev->bvh_width,
                                     std::max(1, dev->cu_count),
                                     S->error, dev->collapse_head);
            if (e != hipSuccess) return hip_fail(e, "tlas build");
            copy_counters_async(ctr, a->tlas_host, S->error, s);
        }
        LC_HIP_TRY(hipEventRecord(a->tlas_ready, s));
        a->tlas_valid = true;
    }
    {
        const int rc = accel_write_descriptor(a, s, a->stage + off_desc);
        if (rc != LC_HIP_OK) return rc;
    }
    // the staging is reused by the next build once these copies ran (no host wait here)
    LC_HIP_TRY(hipEventRecord(a->staged, s));
    a->stage_pending = true;
    return LC_HIP_OK;
}

} // extern "C"

// The DeviceAccel descriptor of lc_hip_device.h: the query view of the accel (as
// lc_hip_accel_query builds it) plus the instance table.
int lchip::accel_write_descriptor(Accel *a, hipStream_t s, void *staging) {
    DeviceAccel &d = a->desc_host;
    std::memset(&d, 0, sizeof(d));
    const uint32_t n = (uint32_t)a->inst.size();
    if (n == 0) {
        d.geo.empty = 1;
    } else if (n == 1) {
        d.geo.instances = a->records;
        d.geo.empty = a->host_records[0].blas_nodes == 0 ? 1u : 0u;
    } else {
        d.geo.two = 1;
        d.geo.tlas_nodes = a->tlas_nodes;
        d.geo.tlas_prims = a->tlas_prims;
        d.geo.instances = a->records;
    }
    d.state = a->state;
    d.state_count = a->state_count;
    d.error = a->dev->error;
    if (!a->desc) LC_HIP_TRY(hipMalloc(&a->desc, sizeof(DeviceAccel)));
    std::memcpy(staging, &d, sizeof(d));
    LC_HIP_TRY(hipMemcpyAsync(a->desc, staging, sizeof(d), hipMemcpyHostToDevice, s));
    return LC_HIP_OK;
}

extern "C" {

static int accel_tlas_sync(Accel *a) {
    if (!a->tlas_stats_valid && a->tlas_valid) {
        LC_HIP_TRY(hipEventSynchronize(a->tlas_ready));
        a->tlas_stats = *a->tlas_host;
        if (a->tlas_stats.tail_level_count != 0) return fail(LC_HIP_ERROR_DEVICE_FAULT, "TLAS collapse exceeded the level bound");
        if (a->tlas_stats.tri_count != a->inst.size()) return fail(LC_HIP_ERROR_DEVICE_FAULT, "TLAS build lost instances");
        a->tlas_stats_valid = true;
    }
    return LC_HIP_OK;
}

// What traversal needs from a built accel: one BLAS (single instance) or the TLAS; the stack
// entries a ray can need: per wide level a node group and a postponed leaf group (speculative
// traversal), plus the TLAS leaf entries.
static int accel_geometry(Accel *a, Geometry &G, bool &two, uint32_t &stack_need) {
    std::memset(&G, 0, sizeof(G));
    two = a->inst.size() > 1;
    stack_need = 0;
    if (!two) {
        const InstanceRecord *r = a->host_records.empty() ? nullptr : &a->host_records[0];
        if (!r || r->blas_nodes == 0 || (r->flags & kInstProcedural)) {
            G.empty = 1; // nothing, or a procedural primitive (seen by ray queries only)
        } else {
            G.blas_nodes = reinterpret_cast<const WideNode *>(r->blas_nodes);
            G.blas_tris = reinterpret_cast<const TriRecord *>(r->blas_tris);
            std::memcpy(G.inst_inv, r->inv, sizeof(G.inst_inv));
            G.inst_mask = r->vis_mask;
            stack_need = 2 * (r->blas_depth + 1);
        }
    } else {
        int rc = accel_tlas_sync(a);
        if (rc != LC_HIP_OK) return rc;
        G.tlas_nodes = a->tlas_nodes;
        G.tlas_prims = a->tlas_prims;
        G.instances = a->records;
        stack_need = 2 * (a->tlas_stats.depth + 2 + a->max_blas_depth + 1);
    }
    return LC_HIP_OK;
}

int lc_hip_accel_stats(uint64_t device, uint64_t accel, lc_hip_bvh_stats *out) {
    Device *dev = obj<Device>(device);
    Accel *a = obj<Accel>(accel);
    if (!dev || !a || !out) return fail(LC_HIP_ERROR_INVALID_HANDLE, "bad accel stats handles");
    DeviceGuard g(dev);
    if (!a->built) return fail(LC_HIP_ERROR_NOT_BUILT, "accel is not built");
    std::memset(out, 0, sizeof(*out));
    out->primitive_count = (uint32_t)a->inst.size();
    if (a->tlas_valid) {
        int rc = accel_tlas_sync(a);
        if (rc != LC_HIP_OK) return rc;
        out->node_count = a->tlas_stats.node_count;
        out->leaf_primitive_count = a->tlas_stats.tri_count;
        out->depth = a->tlas_stats.depth;
        std::memcpy(out->bounds, a->tlas_stats.bounds, sizeof(out->bounds));
        out->node_bytes = (uint64_t)out->node_count * sizeof(WideNode);
        out->primitive_bytes = (uint64_t)a->inst.size() * sizeof(InstanceRecord);
    }
    return LC_HIP_OK;
}

int lc_hip_accel_instance_info(uint64_t device, uint64_t accel, uint32_t index, float affine[12], uint32_t *vis_mask,
                               uint32_t *user_id) {
    Accel *a = obj<Accel>(accel);
    if (!obj<Device>(device) || !a) return fail(LC_HIP_ERROR_INVALID_HANDLE, "bad accel");
    std::lock_guard<std::mutex> lk(a->mu);
    DeviceGuard g(a->dev);
    {
        const int rc = fold_instance_writes(a);
        if (rc != LC_HIP_OK) return rc;
    }
    if (index >= a->inst.size()) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "instance index out of range");
    if (affine) std::memcpy(affine, a->inst[index].affine, sizeof(float) * 12);
    if (vis_mask) *vis_mask = a->inst[index].vis_mask;
    if (user_id) *user_id = a->inst[index].user_id;
    return LC_HIP_OK;
}

// In-kernel instance ops as batches over index buffers (see lc_hip.h).
static int instance_batch(uint64_t device, uint64_t stream, uint64_t accel, uint32_t field, bool write,
                          uint64_t index_buffer, size_t index_offset, uint64_t data_buffer, size_t data_offset,
                          uint32_t count) {
    Device *dev = obj<Device>(device);
    Accel *a = obj<Accel>(accel);
    Buffer *ib = obj<Buffer>(index_buffer), *db = obj<Buffer>(data_buffer);
    if (!dev || !a || !ib || !db) return fail(LC_HIP_ERROR_INVALID_HANDLE, "bad instance op handles");
    if (field > kInstanceVisibility) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "instance field: 0 transform, 1 user id, 2 visibility");
    const size_t item = field == kInstanceTransform ? 12 * sizeof(float) : sizeof(uint32_t);
    if (index_offset % 4 != 0 || data_offset % 4 != 0) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "offsets must be 4-byte aligned");
    if (index_offset + (size_t)count * 4 > ib->bytes || data_offset + (size_t)count * item > db->bytes)
        return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "instance op range outside its buffer");
    std::lock_guard<std::mutex> lk(a->mu);
    if (!a->state) return fail(LC_HIP_ERROR_NOT_BUILT, "accel not built");
    DeviceGuard g(dev);
    hipStream_t s = native_stream(stream);
    const uint32_t *idx = reinterpret_cast<const uint32_t *>(static_cast<const char *>(ib->ptr) + index_offset);
    void *data = static_cast<char *>(db->ptr) + data_offset;
    uint32_t *err = reinterpret_cast<uint32_t *>(a->state + a->state_count);
    LC_HIP_TRY(instance_ops(s, a->state, a->state_count, field, write, idx, count, data, err));
    if (write || count) a->state_written = true; // also reports out-of-range reads at the next fold
    return LC_HIP_OK;
}

int lc_hip_accel_instance_read(uint64_t device, uint64_t stream, uint64_t accel, uint32_t field, uint64_t index_buffer,
                               size_t index_offset, uint64_t out_buffer, size_t out_offset, uint32_t count) {
    return instance_batch(device, stream, accel, field, false, index_buffer, index_offset, out_buffer, out_offset, count);
}

int lc_hip_accel_instance_write(uint64_t device, uint64_t stream, uint64_t accel, uint32_t field,
                                uint64_t index_buffer, size_t index_offset, uint64_t in_buffer, size_t in_offset,
                                uint32_t count) {
    return instance_batch(device, stream, accel, field, true, index_buffer, index_offset, in_buffer, in_offset, count);
}

// One launch over up to kMaxBatches ray batches (queue = the batches back to back). With
// `count_ptr` (one batch) the batch's ray count is its capacity and the launch traces the first
// min(*count_ptr, capacity) rays, the count being read on the device when the launch runs.
static int trace_launch_batches(Device *dev, Accel *a, hipStream_t s, const lc_hip_ray_batch *batches, size_t nb,
                                bool any, bool counting, uint64_t counts[5], const uint32_t *count_ptr = nullptr) {
    TraceParams P;
    std::memset(&P, 0, sizeof(P));
    uint64_t total = 0;
    for (size_t k = 0; k < nb; k++) {
        const lc_hip_ray_batch &bt = batches[k];
        Buffer *rb = obj<Buffer>(bt.ray_buffer);
        Buffer *hb = counting ? nullptr : obj<Buffer>(bt.hit_buffer);
        if (!rb || (!counting && !hb)) return fail(LC_HIP_ERROR_INVALID_HANDLE, "bad trace buffer handles");
        if (bt.ray_offset % 16 != 0) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "ray offset must be 16-byte aligned");
        if (bt.ray_offset + (size_t)bt.ray_count * 32 > rb->bytes)
            return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "rays outside the ray buffer");
        if (!counting) {
            const size_t stride = any ? 1 : 24;
            if (!any && bt.hit_offset % 8 != 0) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "hit offset must be 8-byte aligned");
            if (bt.hit_offset + (size_t)bt.ray_count * stride > hb->bytes)
                return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "hits outside the hit buffer");
        }
        RayBatch &rbt = P.batches.b[k];
        rbt.rays = reinterpret_cast<const float4 *>(static_cast<char *>(rb->ptr) + bt.ray_offset);
        rbt.out = counting ? nullptr : static_cast<char *>(hb->ptr) + bt.hit_offset;
        rbt.start = (uint32_t)total;
        rbt.mask = bt.vis_mask;
        total += bt.ray_count;
    }
    // the launch's slot counter runs past the ray count by at most one chunk per wave
    if (total > 0xF0000000ull) return fail(LC_HIP_ERROR_UNSUPPORTED, "more than 0xF0000000 rays in one launch");
    P.batches.n = (uint32_t)nb;
    P.count = (uint32_t)total;
    P.count_ptr = count_ptr;
    if (counting && counts) std::memset(counts, 0, 5 * sizeof(uint64_t));
    if (total == 0) return LC_HIP_OK;
    bool two = false;
    uint32_t stack_need = 0;
    int grc = accel_geometry(a, P.geo, two, stack_need);
    if (grc != LC_HIP_OK) return grc;
    std::lock_guard<std::mutex> dl(dev->mu);
    const TraceConfig &cfg = dev->cfg[any ? 1 : 0];
    // the counting form runs the timed form's grid (same waves, same queue hand-out)
    int bpc = cfg.blocks_per_cu > 0 ? cfg.blocks_per_cu : trace_blocks_per_cu(any, two, false, cfg.block, cfg.slds);
    uint64_t want = (total + cfg.block - 1) / cfg.block;
    int grid = (int)std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)bpc * dev->cu_count, want));
    P.spill_depth = stack_need > (uint32_t)cfg.slds ? stack_need - (uint32_t)cfg.slds : 0u;
    size_t spill_bytes = (size_t)grid * cfg.block * P.spill_depth * sizeof(uint2);
    StreamScratch *S = stream_scratch(dev, s);
    if (!S) return fail(LC_HIP_ERROR_OUT_OF_MEMORY, "stream scratch");
    LC_HIP_GROW((S, s, S->spill, S->spill_bytes, spill_bytes));
    P.spill = static_cast<uint2 *>(S->spill);
    P.ray_counter = S->ray_counter;
    P.error = S->error;
    P.stats = S->stats;
    P.steps = cfg.steps;
    P.fetch_chunk = cfg.fetch_chunk;
    P.fetch_segments = cfg.fetch_segments;
    P.tri_quarters = cfg.tri_quarters;
    P.refill_below = cfg.refill_below;
    if (counting) LC_HIP_TRY(hipMemsetAsync(S->stats, 0, 5 * sizeof(unsigned long long), s));
    // per-ray results do not depend on the order; the counting form traverses the same sorted
    // queue as the timed form, so its node / triangle totals describe that launch. Queues counted
    // on the device sort their first *count_ptr rays (the sort reads the count when it runs). A
    // captured launch sorts when the stream already holds the sort's working memory (captures
    // cannot allocate) and otherwise traces the queue in its own order.
    bool sort = cfg.sort_rays && total > 1 && !P.geo.empty;
    if (sort && capturing(s) && S->sort_bytes < ray_sort_scratch_bytes(P.count)) sort = false;
    if (sort) LC_HIP_GROW((S, s, S->sort, S->sort_bytes, ray_sort_scratch_bytes(P.count)));
    {
        // "trace_*": the whole entry point (sort + traversal); "trace_*_kernel": traversal only
        TimedRegion tr(dev, s, counting ? "trace_count" : (any ? "trace_any" : "trace_closest"));
        if (sort) {
            hipError_t e = ray_sort(s, P.batches, P.count, count_ptr, a->world_bounds, S->sort, S->sort_bytes, &P.order);
            if (e != hipSuccess) return hip_fail(e, "ray sort");
        }
        TimedRegion tk(dev, s, counting ? "trace_count_kernel" : (any ? "trace_any_kernel" : "trace_closest_kernel"));
        hipError_t e = trace_launch(P, s, any, two, counting, grid, cfg.block, cfg.slds);
        if (e != hipSuccess) return hip_fail(e, "trace launch");
    }
    if (counting) {
        unsigned long long st[5] = {0, 0, 0, 0, 0};
        LC_HIP_TRY(hipMemcpyAsync(st, S->stats, sizeof(st), hipMemcpyDeviceToHost, s));
        LC_HIP_TRY(hipStreamSynchronize(s));
        if (counts)
            for (int k = 0; k < 5; k++) counts[k] = st[k];
    }
    return LC_HIP_OK;
}

static int trace_batches(uint64_t device, uint64_t stream, uint64_t accel, const lc_hip_ray_batch *batches, size_t nb,
                         bool any, bool counting, uint64_t counts[5]) {
    Device *dev = obj<Device>(device);
    Accel *a = obj<Accel>(accel);
    if (!dev || !a) return fail(LC_HIP_ERROR_INVALID_HANDLE, "bad trace handles");
    if (nb > 0 && !batches) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "null batch table");
    std::lock_guard<std::mutex> lk(a->mu);
    if (!a->built) return fail(LC_HIP_ERROR_NOT_BUILT, "accel is not built");
    DeviceGuard g(dev);
    hipStream_t s = native_stream(stream);
    uint64_t sum[5] = {0, 0, 0, 0, 0};
    for (size_t k = 0; k < nb; k += kMaxBatches) {
        uint64_t c[5] = {0, 0, 0, 0, 0};
        int rc = trace_launch_batches(dev, a, s, batches + k, std::min<size_t>(kMaxBatches, nb - k), any, counting, c);
        if (rc != LC_HIP_OK) return rc;
        for (int j = 0; j < 5; j++) sum[j] += c[j];
    }
    if (counts)
        for (int j = 0; j < 5; j++) counts[j] = sum[j];
    return LC_HIP_OK;
}

int lc_hip_accel_trace_closest(uint64_t device, uint64_t stream, uint64_t accel, uint64_t ray_buffer, size_t ray_offset,
                               uint64_t hit_buffer, size_t hit_offset, uint32_t ray_count, uint32_t vis_mask) {
    const lc_hip_ray_batch b{ray_buffer, ray_offset, hit_buffer, hit_offset, ray_count, vis_mask};
    return trace_batches(device, stream, accel, &b, 1, false, false, nullptr);
}

int lc_hip_accel_trace_any(uint64_t device, uint64_t stream, uint64_t accel, uint64_t ray_buffer, size_t ray_offset,
                           uint64_t hit_buffer, size_t hit_offset, uint32_t ray_count, uint32_t vis_mask) {
    const lc_hip_ray_batch b{ray_buffer, ray_offset, hit_buffer, hit_offset, ray_count, vis_mask};
    return trace_batches(device, stream, accel, &b, 1, true, false, nullptr);
}

int lc_hip_accel_trace_closest_batched(uint64_t device, uint64_t stream, uint64_t accel, const lc_hip_ray_batch *batches,
                                       size_t batch_count) {
    return trace_batches(device, stream, accel, batches, batch_count, false, false, nullptr);
}

int lc_hip_accel_trace_any_batched(uint64_t device, uint64_t stream, uint64_t accel, const lc_hip_ray_batch *batches,
                                   size_t batch_count) {
    return trace_batches(device, stream, accel, batches, batch_count, true, false, nullptr);
}

int lc_hip_accel_trace_count(uint64_t device, uint64_t stream, uint64_t accel, uint64_t ray_buffer, size_t ray_offset,
                             uint32_t ray_count, uint32_t vis_mask, int any, uint64_t *nodes_visited,
                             uint64_t *triangles_tested) {
    const lc_hip_ray_batch b{ray_buffer, ray_offset, 0, 0, ray_count, vis_mask};
    uint64_t c[5] = {0, 0, 0, 0, 0};
    const int rc = trace_batches(device, stream, accel, &b, 1, any != 0, true, c);
    if (nodes_visited) *nodes_visited = c[0];
    if (triangles_tested) *triangles_tested = c[1];
    return rc;
}

// Lane accounting of the counting form (no reference counterpart: measurement of the persistent
// traversal): node visits and triangle tests next to the loop iterations lanes spent live and the
// wave iterations issued, so useful lane-steps / (64 x wave iterations) is the kernel's lane
// utilisation in steps.
int lc_hip_accel_trace_steps(uint64_t device, uint64_t stream, uint64_t accel, uint64_t ray_buffer, size_t ray_offset,
                             uint32_t ray_count, uint32_t vis_mask, int any, lc_hip_trace_steps *out) {
    if (!out) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "null trace steps output");
    const lc_hip_ray_batch b{ray_buffer, ray_offset, 0, 0, ray_count, vis_mask};
    uint64_t c[5] = {0, 0, 0, 0, 0};
    const int rc = trace_batches(device, stream, accel, &b, 1, any != 0, true, c);
    out->nodes_visited = c[0];
    out->triangles_tested = c[1];
    out->live_lane_steps = c[2];
    out->wave_steps = c[3];
    out->refill_rounds = c[4];
    return rc;
}

// Traces over a queue whose length is produced on the GPU (a shader's appends): `capacity` rays
// lie in the ray buffer and hit slots in the hit buffer; the launch traces the first
// min(count, capacity) of them, count being the uint32 at count_offset of count_buffer when the
// launch runs. Stands in for a kernel dispatched over an IndirectDispatchBuffer whose body calls
// accel.trace_closest / trace_any (IndirectDispatchBuffer, include/luisa/runtime/dispatch_buffer.h;
// lc_accel_trace_closest, src/backends/cuda/cuda_builtin/cuda_device_resource.h:1783-1829). The
// queue is traced in its own order (no ray sort).
static int trace_indirect(uint64_t device, uint64_t stream, uint64_t accel, uint64_t ray_buffer, size_t ray_offset,
                          uint64_t hit_buffer, size_t hit_offset, uint32_t capacity, uint64_t count_buffer,
                          size_t count_offset, uint32_t vis_mask, bool any) {
    Device *dev = obj<Device>(device);
    Accel *a = obj<Accel>(accel);
    Buffer *cb = obj<Buffer>(count_buffer);
    if (!dev || !a || !cb) return fail(LC_HIP_ERROR_INVALID_HANDLE, "bad indirect trace handles");
    if (count_offset % 4 != 0 || count_offset + 4 > cb->bytes)
        return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "count outside the count buffer (uint32, 4-byte aligned)");
    std::lock_guard<std::mutex> lk(a->mu);
    if (!a->built) return fail(LC_HIP_ERROR_NOT_BUILT, "accel is not built");
    DeviceGuard g(dev);
    hipStream_t s = native_stream(stream);
    const lc_hip_ray_batch b{ray_buffer, ray_offset, hit_buffer, hit_offset, capacity, vis_mask};
    const uint32_t *count_ptr = reinterpret_cast<const uint32_t *>(static_cast<const char *>(cb->ptr) + count_offset);
    return trace_launch_batches(dev, a, s, &b, 1, any, false, nullptr, count_ptr);
}

int lc_hip_accel_trace_closest_indirect(uint64_t device, uint64_t stream, uint64_t accel, uint64_t ray_buffer,
                                        size_t ray_offset, uint64_t hit_buffer, size_t hit_offset, uint32_t capacity,
                                        uint64_t count_buffer, size_t count_offset, uint32_t vis_mask) {
    return trace_indirect(device, stream, accel, ray_buffer, ray_offset, hit_buffer, hit_offset, capacity, count_buffer,
                          count_offset, vis_mask, false);
}

int lc_hip_accel_trace_any_indirect(uint64_t device, uint64_t stream, uint64_t accel, uint64_t ray_buffer,
                                    size_t ray_offset, uint64_t hit_buffer, size_t hit_offset, uint32_t capacity,
                                    uint64_t count_buffer, size_t count_offset, uint32_t vis_mask) {
    return trace_indirect(device, stream, accel, ray_buffer, ray_offset, hit_buffer, hit_offset, capacity, count_buffer,
                          count_offset, vis_mask, true);
}

// DeviceRadixSort::SortPairs of lcub (include/luisa/backends/ext/cuda/lcub/device_radix_sort.h:15-27)
// for uint32 keys and values: the temp_storage_size query and the sort command, on the library's
// own radix kernels (radix.hip). The command allocates nothing and records into a graph.
int lc_hip_radix_sort_pairs_temp_bytes(uint32_t num_items, size_t *temp_bytes) {
    if (!temp_bytes) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "null temp_bytes");
    if (num_items > 0x7FFFFFFFu) return fail(LC_HIP_ERROR_UNSUPPORTED, "more than 2^31 - 1 items");
    *temp_bytes = radix_sort_temp_bytes(num_items);
    return LC_HIP_OK;
}

int lc_hip_radix_sort_pairs(uint64_t device, uint64_t stream, uint64_t temp_buffer, uint64_t keys_in,
                            uint64_t keys_out, uint64_t values_in, uint64_t values_out, uint32_t num_items,
                            int begin_bit, int end_bit, uint64_t count_buffer, size_t count_offset) {
    Device *dev = obj<Device>(device);
    Buffer *tb = obj<Buffer>(temp_buffer), *ki = obj<Buffer>(keys_in), *ko = obj<Buffer>(keys_out);
    Buffer *vi = obj<Buffer>(values_in), *vo = obj<Buffer>(values_out);
    if (!dev || !tb || !ki || !ko || !vo || (values_in && !vi)) return fail(LC_HIP_ERROR_INVALID_HANDLE, "bad sort handles");
    if (begin_bit < 0 || end_bit > 32 || begin_bit >= end_bit)
        return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "bits must satisfy 0 <= begin_bit < end_bit <= 32");
    if (num_items > 0x7FFFFFFFu) return fail(LC_HIP_ERROR_UNSUPPORTED, "more than 2^31 - 1 items");
    const size_t bytes = (size_t)num_items * 4;
    if (ki->bytes < bytes || ko->bytes < bytes || vo->bytes < bytes || (vi && vi->bytes < bytes))
        return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "a key or value buffer holds fewer than num_items words");
    if (tb->bytes < radix_sort_temp_bytes(num_items))
        return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "temp buffer smaller than lc_hip_radix_sort_pairs_temp_bytes");
    if (ki == ko || (vi && vi == vo) || tb == ki || tb == ko || tb == vo || (vi && tb == vi) || ko == vo)
        return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "inputs, outputs and temp storage must be distinct buffers");
    const uint32_t *count_ptr = nullptr;
    if (count_buffer) {
        Buffer *cb = obj<Buffer>(count_buffer);
        if (!cb) return fail(LC_HIP_ERROR_INVALID_HANDLE, "bad count buffer");
        if (count_offset % 4 != 0 || count_offset + 4 > cb->bytes)
            return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "count outside the count buffer (uint32, 4-byte aligned)");
        count_ptr = reinterpret_cast<const uint32_t *>(static_cast<const char *>(cb->ptr) + count_offset);
    }
    if (num_items == 0) return LC_HIP_OK;
    DeviceGuard g(dev);
    hipStream_t s = native_stream(stream);
    TimedRegion tr(dev, s, "radix_sort_pairs");
    hipError_t e = radix_sort_pairs(s, tb->ptr, tb->bytes, static_cast<const uint32_t *>(ki->ptr),
                                    static_cast<uint32_t *>(ko->ptr), vi ? static_cast<const uint32_t *>(vi->ptr) : nullptr,
                                    static_cast<uint32_t *>(vo->ptr), num_items, count_ptr, (uint32_t)begin_bit,
                                    (uint32_t)end_bit);
    if (e != hipSuccess) return hip_fail(e, "radix sort");
    return LC_HIP_OK;
}

// Closest hits without allocating: the working memory that lc_hip_accel_trace_closest takes from
// the stream's buffers (slot counters, error word, stack spill) comes from this GPU's static
// scratch array, and the rays are traced in queue order (the ray sort's scratch grows with the
// ray count). Per-ray results equal lc_hip_accel_trace_closest's. The grid shrinks when the
// spill area cannot hold a full grid's stacks. Calls that overlap on one GPU share the array
// and race.
int lc_hip_accel_trace_closest_noalloc(uint64_t device, uint64_t stream, uint64_t accel, uint64_t ray_buffer,
                                       size_t ray_offset, uint64_t hit_buffer, size_t hit_offset, uint32_t ray_count,
                                       uint32_t vis_mask) {
    Device *dev = obj<Device>(device);
    Accel *a = obj<Accel>(accel);
    Buffer *rb = obj<Buffer>(ray_buffer), *hb = obj<Buffer>(hit_buffer);
    if (!dev || !a || !rb || !hb) return fail(LC_HIP_ERROR_INVALID_HANDLE, "bad trace handles");
    if (!dev->noalloc_scratch) return fail(LC_HIP_ERROR_UNSUPPORTED, "no-allocation scratch array unavailable on this device");
    if (ray_offset % 16 != 0) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "ray offset must be 16-byte aligned");
    if (hit_offset % 8 != 0) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "hit offset must be 8-byte aligned");
    if (ray_offset + (size_t)ray_count * 32 > rb->bytes) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "rays outside the ray buffer");
    if (hit_offset + (size_t)ray_count * 24 > hb->bytes) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "hits outside the hit buffer");
    if (ray_count > 0xF0000000u) return fail(LC_HIP_ERROR_UNSUPPORTED, "more than 0xF0000000 rays in one launch");
    std::lock_guard<std::mutex> lk(a->mu);
    if (!a->built) return fail(LC_HIP_ERROR_NOT_BUILT, "accel is not built");
    DeviceGuard g(dev);
    hipStream_t s = native_stream(stream);
    if (ray_count == 0) return LC_HIP_OK;
    TraceParams P;
    std::memset(&P, 0, sizeof(P));
    P.batches.n = 1;
    P.batches.b[0].rays = reinterpret_cast<const float4 *>(static_cast<char *>(rb->ptr) + ray_offset);
    P.batches.b[0].out = static_cast<char *>(hb->ptr) + hit_offset;
    P.batches.b[0].start = 0;
    P.batches.b[0].mask = vis_mask;
    P.count = ray_count;
    bool two = false;
    uint32_t stack_need = 0;
    const int grc = accel_geometry(a, P.geo, two, stack_need);
    if (grc != LC_HIP_OK) return grc;
    std::lock_guard<std::mutex> dl(dev->mu);
    const TraceConfig &cfg = dev->cfg[0];
    const int bpc = cfg.blocks_per_cu > 0 ? cfg.blocks_per_cu : trace_blocks_per_cu(false, two, false, cfg.block, cfg.slds);
    const uint64_t want = ((uint64_t)ray_count + cfg.block - 1) / cfg.block;
    uint64_t grid = std::max<uint64_t>(1, std::min<uint64_t>((uint64_t)bpc * dev->cu_count, want));
    P.spill_depth = stack_need > (uint32_t)cfg.slds ? stack_need - (uint32_t)cfg.slds : 0u;
    const size_t spill_per_block = (size_t)cfg.block * P.spill_depth * sizeof(uint2);
    const size_t spill_cap = kNoAllocScratchBytes - kNoAllocWordsBytes;
    if (spill_per_block > spill_cap)
        return fail(LC_HIP_ERROR_UNSUPPORTED, "traversal stacks of one workgroup need " + std::to_string(spill_per_block) +
                                                  " spill bytes, more than the no-allocation scratch holds (" +
                                                  std::to_string(spill_cap) + ")");
    if (spill_per_block) grid = std::min<uint64_t>(grid, spill_cap / spill_per_block);
    static_assert(576 + 2 * sizeof(unsigned long long) <= kNoAllocWordsBytes, "no-allocation words");
    char *base = static_cast<char *>(dev->noalloc_scratch);
    P.ray_counter = reinterpret_cast<uint32_t *>(base);
    P.error = reinterpret_cast<uint32_t *>(base) + kRayCounterWords;
    P.stats = reinterpret_cast<unsigned long long *>(base + 576);
    P.spill = reinterpret_cast<uint2 *>(base + kNoAllocWordsBytes);
    P.steps = cfg.steps;
    P.fetch_chunk = cfg.fetch_chunk;
    P.fetch_segments = cfg.fetch_segments;
    P.tri_quarters = cfg.tri_quarters;
    P.refill_below = cfg.refill_below;
    const hipError_t e = trace_launch(P, s, false, two, false, (int)grid, cfg.block, cfg.slds);
    if (e != hipSuccess) return hip_fail(e, "trace launch");
    return LC_HIP_OK;
}

// Ray queries (accel.traverse / traverse_any, include/luisa/dsl/rtx/ray_query.h:11-243) over a
// batch with the built-in candidate callbacks of query.hip. Unlike plain traces they see the
// procedural instances too; stack entries: per wide level one node group and one held group of
// leaf records (wave-uniform steps), plus the TLAS leaf entries.
int lc_hip_accel_query(uint64_t device, uint64_t stream, uint64_t accel, uint64_t ray_buffer, size_t ray_offset,
                       uint64_t hit_buffer, size_t hit_offset, uint32_t ray_count, uint32_t vis_mask,
                       const lc_hip_query_policy *policy) {
    Device *dev = obj<Device>(device);
    Accel *a = obj<Accel>(accel);
    Buffer *rb = obj<Buffer>(ray_buffer), *hb = obj<Buffer>(hit_buffer);
    if (!dev || !a || !rb || !hb) return fail(LC_HIP_ERROR_INVALID_HANDLE, "bad query handles");
    if (!policy) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "null query policy");
    if (policy->surface_mode > LC_HIP_QUERY_SURFACE_CUTOUT || policy->procedural_mode > LC_HIP_QUERY_PROCEDURAL_SPHERE)
        return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "unknown query callback");
    if (ray_offset % 16 != 0) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "ray offset must be 16-byte aligned");
    if (hit_offset % 8 != 0) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "hit offset must be 8-byte aligned");
    if (ray_offset + (size_t)ray_count * 32 > rb->bytes) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "rays outside the ray buffer");
    if (hit_offset + (size_t)ray_count * 24 > hb->bytes) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "hits outside the hit buffer");
    std::lock_guard<std::mutex> lk(a->mu);
    if (!a->built) return fail(LC_HIP_ERROR_NOT_BUILT, "accel is not built");
    DeviceGuard g(dev);
    hipStream_t s = native_stream(stream);
    if (ray_count == 0) return LC_HIP_OK;
    QueryParams P;
    std::memset(&P, 0, sizeof(P));
    uint32_t stack_need = 0;
    const uint32_t n_inst = (uint32_t)a->inst.size();
    if (n_inst == 0) {
        P.geo.empty = 1;
    } else if (n_inst == 1) {
        const InstanceRecord &r = a->host_records[0];
        P.geo.instances = a->records;
        P.geo.empty = r.blas_nodes == 0 ? 1u : 0u;
        stack_need = 2 * (r.blas_depth + 2); // a node group and a held leaf group per level
    } else {
        int rc = accel_tlas_sync(a);
        if (rc != LC_HIP_OK) return rc;
        P.geo.two = 1;
        P.geo.tlas_nodes = a->tlas_nodes;
        P.geo.tlas_prims = a->tlas_prims;
        P.geo.instances = a->records;
        stack_need = 2 * (a->tlas_stats.depth + 2 + a->max_blas_depth + 2);
    }
    if (ray_count > 0xF0000000u) return fail(LC_HIP_ERROR_UNSUPPORTED, "more than 0xF0000000 rays in one query");
    P.rays = reinterpret_cast<const float4 *>(static_cast<char *>(rb->ptr) + ray_offset);
    P.hits = static_cast<char *>(hb->ptr) + hit_offset;
    P.count = ray_count;
    P.mask = vis_mask;
    P.first = policy->terminate_on_first ? 1u : 0u;
    P.surface_mode = policy->surface_mode;
    P.cutout = policy->cutout;
    P.procedural_mode = policy->procedural_mode;
    std::lock_guard<std::mutex> dl(dev->mu);
    const int block = query_block();
    const uint64_t want = ((uint64_t)ray_count + block - 1) / block;
    const int grid = (int)std::max<uint64_t>(
        1, std::min<uint64_t>((uint64_t)query_blocks_per_cu() * std::max(1, dev->cu_count), want));
    const uint32_t slds = (uint32_t)query_lds_entries();
    P.spill_depth = stack_need > slds ? stack_need - slds : 0u;
    StreamScratch *S = stream_scratch(dev, s);
    if (!S) return fail(LC_HIP_ERROR_OUT_OF_MEMORY, "stream scratch");
    LC_HIP_GROW((S, s, S->spill, S->spill_bytes, (size_t)grid * block * P.spill_depth * sizeof(uint2)));
    P.spill = static_cast<uint2 *>(S->spill);
    P.error = S->error;
    P.counter = S->ray_counter;
    // the query queue is sorted like trace queues when ray sorting is on (per-ray answers of the
    // nearest-commit query do not depend on the order); under capture as for traces
    bool sort = dev->cfg[P.first ? 1 : 0].sort_rays && ray_count > 1 && !P.geo.empty;
    if (sort && capturing(s) && S->sort_bytes < ray_sort_scratch_bytes(ray_count)) sort = false;
    if (sort) LC_HIP_GROW((S, s, S->sort, S->sort_bytes, ray_sort_scratch_bytes(ray_count)));
    TimedRegion tr(dev, s, P.first ? "query_any" : "query");
    if (sort) {
        BatchTable bt;
        std::memset(&bt, 0, sizeof(bt));
        bt.n = 1;
        bt.b[0].rays = P.rays;
        bt.b[0].mask = vis_mask;
        hipError_t e = ray_sort(s, bt, ray_count, nullptr, a->world_bounds, S->sort, S->sort_bytes, &P.order);
        if (e != hipSuccess) return hip_fail(e, "query ray sort");
    }
    hipError_t e = query_launch(P, s, grid);
    if (e != hipSuccess) return hip_fail(e, "query launch");
    return LC_HIP_OK;
}

// Light and camera of a lc_hip_pt_scene; the light's normal and area are normalize(cross(u, v))
// and |cross(u, v)| in single precision. False for a degenerate light quad.
static bool pt_scene_from(const lc_hip_pt_scene &scene, PtScene &S) {
    std::memset(&S, 0, sizeof(S));
    for (int k = 0; k < 3; k++) {
        S.light_pos[k] = scene.light_position[k];
        S.light_u[k] = scene.light_u[k];
        S.light_v[k] = scene.light_v[k];
        S.light_e[k] = scene.light_emission[k];
        S.cam_o[k] = scene.camera_origin[k];
        S.cam_r[k] = scene.camera_right[k];
        S.cam_u[k] = scene.camera_up[k];
        S.cam_f[k] = scene.camera_forward[k];
    }
    const float *u = S.light_u, *v = S.light_v;
    const float c[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
    const float len = std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]);
    if (!(len > 0.0f)) return false;
    for (int k = 0; k < 3; k++) S.light_n[k] = c[k] / len;
    S.light_area = len;
    return true;
}

// Shading table of an accel's instances (caller holds a->mu, instance_count == a->inst.size()).
static std::vector<PtInstance> pt_table_from(Accel *a, const lc_hip_pt_instance *pt_instances, uint32_t instance_count) {
    std::vector<PtInstance> table(instance_count);
    for (uint32_t i = 0; i < instance_count; i++) {
        const lc_hip_pt_instance &pi = pt_instances[i];
        const Mesh *m = a->inst[i].mesh;
        PtInstance &t = table[i];
        std::memset(&t, 0, sizeof(t));
        t.tris = (m && m->built) ? m->tris : nullptr; // hits only name records of built meshes
        std::memcpy(t.fwd, a->inst[i].affine, sizeof(t.fwd));
        std::memcpy(t.albedo, pi.albedo, sizeof(t.albedo));
        t.emissive = pi.emissive ? 1u : 0u;
    }
    return table;
}

static bool pt_fetch_from(uint32_t chunk, uint32_t segments, PtOptions &po) {
    if ((segments > 1 && segments != kFetchSegments) || chunk > 4096 || chunk % 64 != 0) return false;
    // default: 128-slot claims from one queue segment per XCD (each XCD's L2 keeps its own
    // neighbouring paths; measured best on the box room: 4.24 vs 3.51 G rays/s for exact claims)
    po.fetch_chunk = segments == 0 ? 128u : chunk;
    po.fetch_segments = segments == 0 ? kFetchSegments : segments;
    return true;
}

int lc_hip_pathtrace(uint64_t device, uint64_t stream, uint64_t accel, const lc_hip_pt_instance *pt_instances,
                     uint32_t instance_count, const lc_hip_pt_scene *scene, const lc_hip_pt_options *options,
                     uint64_t image_buffer, size_t image_offset, uint64_t counters_buffer) {
    Device *dev = obj<Device>(device);
    Accel *a = obj<Accel>(accel);
    Buffer *img = obj<Buffer>(image_buffer);
    if (!dev || !a || !img) return fail(LC_HIP_ERROR_INVALID_HANDLE, "bad pathtrace handles");
    if (!scene || !options || (instance_count > 0 && !pt_instances))
        return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "null pathtrace argument");
    const lc_hip_pt_options &o = *options;
    if (o.width == 0 || o.height == 0 || o.spp == 0) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "empty image or no samples");
    if ((uint64_t)o.width * o.height >= (1ull << 31)) return fail(LC_HIP_ERROR_UNSUPPORTED, "image too large");
    if (o.max_depth < 1 || o.max_depth > 12) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "max_depth must be in 1..12");
    if (o.mode > 1) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "unknown pathtrace mode");
    if ((uint64_t)o.sample_offset + o.spp > 0xFFFFFFFFull) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "sample index overflow");
    const uint64_t npix = (uint64_t)o.width * o.height;
    if (image_offset % 16 != 0 || image_offset + npix * 16 > img->bytes)
        return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "image outside the image buffer (float4 per pixel, 16-B aligned)");
    unsigned long long *counts = nullptr;
    if (counters_buffer) {
        Buffer *cb = obj<Buffer>(counters_buffer);
        if (!cb || cb->bytes < 16) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "counters buffer must hold two uint64");
        counts = static_cast<unsigned long long *>(cb->ptr);
    }
    std::lock_guard<std::mutex> lk(a->mu);
    if (!a->built) return fail(LC_HIP_ERROR_NOT_BUILT, "accel is not built");
    if (instance_count != a->inst.size())
        return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "one lc_hip_pt_instance per accel instance");
    std::vector<PtInstance> table = pt_table_from(a, pt_instances, instance_count);
    PtScene S;
    if (!pt_scene_from(*scene, S)) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "degenerate light quad");
    S.width = o.width;
    S.height = o.height;
    S.spp = o.spp;
    S.sample_offset = o.sample_offset;
    S.max_depth = o.max_depth;
    S.rr_depth = o.rr_depth;
    S.seed = o.seed;
    S.sequence = o.sequence;
    PtOptions po{o.mode == LC_HIP_PT_WAVEFRONT ? 1u : 0u, o.chunk_spp, o.sort ? 1u : 0u, o.graph ? 1u : 0u, 0u, 0u};
    if (!pt_fetch_from(o.fetch_chunk, o.fetch_segments, po))
        return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "fetch_chunk: a multiple of 64 up to 4096; fetch_segments: 0, 1 or 8");
    DeviceGuard g(dev);
    hipStream_t s = native_stream(stream);
    Geometry G;
    bool two = false;
    uint32_t stack_need = 0;
    int rc = accel_geometry(a, G, two, stack_need);
    if (rc != LC_HIP_OK) return rc;
    float4 *image = reinterpret_cast<float4 *>(static_cast<char *>(img->ptr) + image_offset);
    std::string err;
    int r;
    {
        TimedRegion tr(dev, s, po.wavefront ? "pathtrace_wavefront" : "pathtrace_megakernel");
        r = pt_render(dev, s, G, two, stack_need, a->world_bounds, S, table, po, image, counts, err);
    }
    if (r == -2) return fail(LC_HIP_ERROR_UNSUPPORTED, err);
    if (r != 0) return fail(LC_HIP_ERROR_HIP, err);
    return LC_HIP_OK;
}

// ---- render serving ----
uint64_t lc_hip_server_create(uint64_t device, const lc_hip_server_options *options) {
    Device *dev = obj<Device>(device);
    if (!dev || !options) {
        fail(LC_HIP_ERROR_INVALID_HANDLE, "bad server arguments");
        return LC_HIP_INVALID_HANDLE;
    }
    const lc_hip_server_options &o = *options;
    ServeOptions so;
    so.hbm_budget = o.hbm_budget_bytes;
    so.wave_paths = o.wave_paths;
    so.max_active = o.max_active;
    so.pt = PtOptions{1u, 0u, o.sort ? 1u : 0u, o.graph ? 1u : 0u, 0u, 0u};
    if (so.wave_paths == 0 || so.wave_paths >= (1ull << 31) || so.max_active == 0 || so.hbm_budget == 0 ||
        !pt_fetch_from(o.fetch_chunk, o.fetch_segments, so.pt)) {
        fail(LC_HIP_ERROR_INVALID_ARGUMENT, "server options: wave_paths in 1..2^31-1, max_active and budget > 0, "
                                            "fetch as for lc_hip_pathtrace");
        return LC_HIP_INVALID_HANDLE;
    }
    DeviceGuard g(dev);
    std::string err;
    Server *s = serve_create(dev, so, err);
    if (!s) {
        fail(LC_HIP_ERROR_OUT_OF_MEMORY, err);
        return LC_HIP_INVALID_HANDLE;
    }
    return handle(s);
}

void lc_hip_server_destroy(uint64_t device, uint64_t server) {
    Device *dev = obj<Device>(device);
    Server *s = obj<Server>(server);
    if (!dev || !s) return;
    if (defer_while_capturing(dev, [device, server] { lc_hip_server_destroy(device, server); })) return;
    DeviceGuard g(dev);
    serve_destroy(s);
}

int lc_hip_server_add_scene(uint64_t device, uint64_t server, uint64_t accel, const lc_hip_pt_instance *pt_instances,
                            uint32_t instance_count, const lc_hip_pt_scene *scene, uint32_t max_depth,
                            uint32_t rr_depth, uint32_t *scene_id) {
    Device *dev = obj<Device>(device);
    Server *s = obj<Server>(server);
    Accel *a = obj<Accel>(accel);
    if (!dev || !s || !a) return fail(LC_HIP_ERROR_INVALID_HANDLE, "bad server scene handles");
    if (!scene || !scene_id || (instance_count > 0 && !pt_instances))
        return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "null server scene argument");
    if (max_depth < 1 || max_depth > 12) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "max_depth must be in 1..12");
    PtBinding B;
    if (!pt_scene_from(*scene, B.S)) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "degenerate light quad");
    B.S.max_depth = max_depth;
    B.S.rr_depth = rr_depth;
    DeviceGuard g(dev);
    {
        std::lock_guard<std::mutex> lk(a->mu);
        if (!a->built) return fail(LC_HIP_ERROR_NOT_BUILT, "accel is not built");
        if (instance_count != a->inst.size())
            return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "one lc_hip_pt_instance per accel instance");
        B.inst = pt_table_from(a, pt_instances, instance_count);
        int rc = accel_geometry(a, B.G, B.two, B.stack_need);
        if (rc != LC_HIP_OK) return rc;
        std::memcpy(B.bounds, a->world_bounds, sizeof(B.bounds));
    }
    std::string err;
    const int id = serve_add_scene(s, B, err);
    if (id < 0) return fail(LC_HIP_ERROR_HIP, err);
    *scene_id = (uint32_t)id;
    return LC_HIP_OK;
}

int lc_hip_server_submit(uint64_t device, uint64_t server, const lc_hip_render_request *request, uint64_t *request_id) {
    Device *dev = obj<Device>(device);
    Server *s = obj<Server>(server);
    if (!dev || !s) return fail(LC_HIP_ERROR_INVALID_HANDLE, "bad server handles");
    if (!request || !request_id) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "null request");
    const lc_hip_render_request &q = *request;
    Buffer *img = obj<Buffer>(q.image_buffer);
    if (!img) return fail(LC_HIP_ERROR_INVALID_HANDLE, "bad request image buffer");
    const uint64_t npix = (uint64_t)q.width * q.height;
    if (q.image_offset % 16 != 0 || q.image_offset + npix * 16 > img->bytes)
        return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "image outside the image buffer (float4 per pixel, 16-B aligned)");
    if ((uint64_t)q.spp > 0xFFFFFFFFull / 64) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "too many samples");
    std::string err;
    float4 *out = reinterpret_cast<float4 *>(static_cast<char *>(img->ptr) + q.image_offset);
    if (serve_submit(s, q.scene, q.width, q.height, q.spp, q.seed, q.sequence, out, request_id, err) != 0)
        return fail(LC_HIP_ERROR_INVALID_ARGUMENT, err);
    return LC_HIP_OK;
}

int lc_hip_server_query(uint64_t device, uint64_t server, uint64_t request_id, lc_hip_request_status *status) {
    Server *s = obj<Server>(server);
    if (!obj<Device>(device) || !s || !status) return fail(LC_HIP_ERROR_INVALID_HANDLE, "bad server handles");
    ServeStatus st;
    if (serve_query(s, request_id, &st) != 0) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "unknown request");
    status->state = (uint32_t)st.state;
    status->samples_done = st.samples_done;
    status->t_submit = st.t_submit;
    status->t_admit = st.t_admit;
    status->t_done = st.t_done;
    return LC_HIP_OK;
}

int lc_hip_server_wait(uint64_t device, uint64_t server, uint64_t request_id, double timeout_seconds) {
    Server *s = obj<Server>(server);
    if (!obj<Device>(device) || !s) return fail(LC_HIP_ERROR_INVALID_HANDLE, "bad server handles");
    const int r = serve_wait(s, request_id, timeout_seconds);
    if (r == 1) return fail(LC_HIP_ERROR_TIMEOUT, "request not finished within the timeout");
    if (r == -1) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "unknown request");
    if (r == -2) return fail(LC_HIP_ERROR_HIP, "request failed");
    return LC_HIP_OK;
}

int lc_hip_server_drain(uint64_t device, uint64_t server, double timeout_seconds) {
    Server *s = obj<Server>(server);
    if (!obj<Device>(device) || !s) return fail(LC_HIP_ERROR_INVALID_HANDLE, "bad server handles");
    if (serve_drain(s, timeout_seconds) != 0) return fail(LC_HIP_ERROR_TIMEOUT, "requests still running");
    return LC_HIP_OK;
}

int lc_hip_server_stats(uint64_t device, uint64_t server, lc_hip_serving_stats *stats) {
    Device *dev = obj<Device>(device);
    Server *s = obj<Server>(server);
    if (!dev || !s || !stats) return fail(LC_HIP_ERROR_INVALID_HANDLE, "bad server handles");
    DeviceGuard g(dev);
    ServeStats st;
    serve_stats(s, &st);
    std::memset(stats, 0, sizeof(*stats));
    stats->waves = st.waves;
    stats->rays_closest = st.rays_closest;
    stats->rays_shadow = st.rays_shadow;
    stats->requests_done = st.requests_done;
    stats->peak_active = st.peak_active;
    stats->active = st.active;
    stats->queued = st.queued;
    stats->peak_bytes = st.peak_bytes;
    return LC_HIP_OK;
}

int lc_hip_trace_set_config(uint64_t device, int kind, int block, int lds_stack, int steps, int blocks_per_cu) {
    Device *dev = obj<Device>(device);
    if (!dev || kind < 0 || kind > 1) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "bad trace config target");
    if (!trace_config_supported(block, lds_stack) || steps < 1 || blocks_per_cu < 0)
        return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "unsupported trace config");
    std::lock_guard<std::mutex> lk(dev->mu);
    TraceConfig &c = dev->cfg[kind];
    c.block = block;
    c.slds = lds_stack;
    c.steps = steps;
    c.blocks_per_cu = blocks_per_cu;
    return LC_HIP_OK;
}

int lc_hip_trace_set_fetch(uint64_t device, int kind, uint32_t chunk, uint32_t segments) {
    Device *dev = obj<Device>(device);
    if (!dev || kind < 0 || kind > 1) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "bad trace config target");
    if (chunk > 4096 || (chunk != 0 && chunk % 64 != 0) || (segments != 1 && segments != kFetchSegments))
        return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "fetch chunk must be 0 or a multiple of 64 up to 4096; segments 1 or 8");
    std::lock_guard<std::mutex> lk(dev->mu);
    dev->cfg[kind].fetch_chunk = chunk;
    dev->cfg[kind].fetch_segments = segments;
    return LC_HIP_OK;
}

// Traversal step policy of a trace kind (autotuned): the triangle-step threshold of closest-hit
// traversal (4 x lanes holding triangles >= tri_quarters x live lanes, 1..4) and the early refill
// (a wave refills once fewer than refill_below of its 64 lanes hold a ray; 0 off).
int lc_hip_trace_set_steps(uint64_t device, int kind, uint32_t tri_quarters, uint32_t refill_below) {
    Device *dev = obj<Device>(device);
    if (!dev || kind < 0 || kind > 1) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "bad trace kind");
    if (tri_quarters < 1 || tri_quarters > 4 || refill_below > 64)
        return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "tri_quarters in 1..4, refill_below in 0..64");
    std::lock_guard<std::mutex> dl(dev->mu);
    dev->cfg[kind].tri_quarters = tri_quarters;
    dev->cfg[kind].refill_below = refill_below;
    return LC_HIP_OK;
}

int lc_hip_trace_get_steps(uint64_t device, int kind, uint32_t *tri_quarters, uint32_t *refill_below) {
    Device *dev = obj<Device>(device);
    if (!dev || kind < 0 || kind > 1 || !tri_quarters || !refill_below)
        return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "bad trace kind");
    std::lock_guard<std::mutex> dl(dev->mu);
    *tri_quarters = dev->cfg[kind].tri_quarters;
    *refill_below = dev->cfg[kind].refill_below;
    return LC_HIP_OK;
}

int lc_hip_trace_get_fetch(uint64_t device, int kind, uint32_t *chunk, uint32_t *segments) {
    Device *dev = obj<Device>(device);
    if (!dev || kind < 0 || kind > 1) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "bad trace config target");
    std::lock_guard<std::mutex> lk(dev->mu);
    if (chunk) *chunk = dev->cfg[kind].fetch_chunk;
    if (segments) *segments = dev->cfg[kind].fetch_segments;
    return LC_HIP_OK;
}

int lc_hip_trace_get_config(uint64_t device, int kind, int *block, int *lds_stack, int *steps, int *blocks_per_cu) {
    Device *dev = obj<Device>(device);
    if (!dev || kind < 0 || kind > 1) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "bad trace config target");
    std::lock_guard<std::mutex> lk(dev->mu);
    const TraceConfig &c = dev->cfg[kind];
    if (block) *block = c.block;
    if (lds_stack) *lds_stack = c.slds;
    if (steps) *steps = c.steps;
    if (blocks_per_cu) *blocks_per_cu = c.blocks_per_cu;
    return LC_HIP_OK;
}

int lc_hip_trace_set_ray_sort(uint64_t device, int kind, int enable) {
    Device *dev = obj<Device>(device);
    if (!dev || kind < 0 || kind > 1) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "bad trace config target");
    std::lock_guard<std::mutex> lk(dev->mu);
    dev->cfg[kind].sort_rays = enable != 0;
    return LC_HIP_OK;
}

int lc_hip_trace_get_ray_sort(uint64_t device, int kind, int *enable) {
    Device *dev = obj<Device>(device);
    if (!dev || kind < 0 || kind > 1 || !enable) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "bad trace config target");
    std::lock_guard<std::mutex> lk(dev->mu);
    *enable = dev->cfg[kind].sort_rays ? 1 : 0;
    return LC_HIP_OK;
}

int lc_hip_build_set_width(uint64_t device, int width) {
    Device *dev = obj<Device>(device);
    if (!dev) return fail(LC_HIP_ERROR_INVALID_HANDLE, "bad device");
    if (width < 2 || width > 8) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "BVH width must be 2..8");
    std::lock_guard<std::mutex> lk(dev->mu);
    dev->bvh_width = (uint32_t)width;
    return LC_HIP_OK;
}

int lc_hip_build_set_collapse_head(uint64_t device, int levels) {
    Device *dev = obj<Device>(device);
    if (!dev) return fail(LC_HIP_ERROR_INVALID_HANDLE, "bad device");
    if (levels < 0 || levels > (int)kCollapseLaunches)
        return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "collapse head levels must be 0 (automatic) .. 98");
    std::lock_guard<std::mutex> lk(dev->mu);
    dev->collapse_head = (uint32_t)levels;
    return LC_HIP_OK;
}

int lc_hip_build_get_width(uint64_t device) {
    Device *dev = obj<Device>(device);
    if (!dev) return fail(LC_HIP_ERROR_INVALID_HANDLE, "bad device");
    return (int)dev->bvh_width;
}

int lc_hip_build_set_sah(uint64_t device, int enable) {
    Device *dev = obj<Device>(device);
    if (!dev) return fail(LC_HIP_ERROR_INVALID_HANDLE, "bad device");
    std::lock_guard<std::mutex> lk(dev->mu);
    dev->sah_refine = enable ? 1u : 0u;
    return LC_HIP_OK;
}

int lc_hip_build_set_split_small(uint64_t device, int enable) {
    Device *dev = obj<Device>(device);
    if (!dev) return fail(LC_HIP_ERROR_INVALID_HANDLE, "bad device");
    dev->split_small = enable ? 1u : 0u;
    return LC_HIP_OK;
}

int lc_hip_timing_enable(uint64_t device, int enable) {
    Device *dev = obj<Device>(device);
    if (!dev) return fail(LC_HIP_ERROR_INVALID_HANDLE, "bad device");
    std::lock_guard<std::mutex> lk(dev->mu);
    dev->timing = enable != 0;
    return LC_HIP_OK;
}

int lc_hip_timing_get(uint64_t device, const char *name, double *total_ms, uint64_t *launches) {
    Device *dev = obj<Device>(device);
    if (!dev || !name) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "bad timing query");
    DeviceGuard g(dev);
    std::lock_guard<std::mutex> lk(dev->mu);
    LC_HIP_TRY(collect_timings(dev));
    auto it = dev->timings.find(name);
    if (total_ms) *total_ms = it == dev->timings.end() ? 0.0 : it->second.total_ms;
    if (launches) *launches = it == dev->timings.end() ? 0 : it->second.launches;
    return LC_HIP_OK;
}

int lc_hip_timing_reset(uint64_t device) {
    Device *dev = obj<Device>(device);
    if (!dev) return fail(LC_HIP_ERROR_INVALID_HANDLE, "bad device");
    DeviceGuard g(dev);
    std::lock_guard<std::mutex> lk(dev->mu);
    hipError_t e = collect_timings(dev);
    dev->timings.clear();
    if (e != hipSuccess) return hip_fail(e, "kernel timing");
    return LC_HIP_OK;
}

// ---- images (2-D, one level, linear rows) ----
} // extern "C"

uint32_t lchip::pixel_storage_bytes(uint32_t storage) {
    if (storage > LC_HIP_PIXEL_FLOAT4) return 0;
    static const uint32_t channel_bytes[5] = {1, 2, 4, 2, 4}; // byte, short, int, half, float
    static const uint32_t channels[3] = {1, 2, 4};
    return channel_bytes[storage / 3] * channels[storage % 3];
}

extern "C" {

static uint64_t texture_create(uint64_t device, uint32_t storage, uint32_t width, uint32_t height, uint32_t depth) {
    Device *dev = obj<Device>(device);
    const uint32_t pb = pixel_storage_bytes(storage);
    if (!dev || pb == 0 || width == 0 || height == 0 || depth == 0 || width > 65536 || height > 65536 ||
        depth > 65536 || (size_t)width * height * depth > (1ull << 34)) {
        fail(LC_HIP_ERROR_UNSUPPORTED, "image storage or size not supported");
        return LC_HIP_INVALID_HANDLE;
    }
    DeviceGuard g(dev);
    auto t = std::make_unique<Texture>();
    t->dev = dev;
    t->storage = storage; t->width = width; t->height = height; t->depth = depth; t->pixel_bytes = pb;
    const size_t bytes = (size_t)width * height * depth * pb;
    hipError_t e = hipMalloc(&t->ptr, bytes);
    if (e == hipSuccess) e = hipMemset(t->ptr, 0, bytes);
    if (e != hipSuccess) {
        if (t->ptr) (void)hipFree(t->ptr);
        hip_fail(e, "image allocation");
        return LC_HIP_INVALID_HANDLE;
    }
    return reinterpret_cast<uint64_t>(t.release());
}

uint64_t lc_hip_texture_create(uint64_t device, uint32_t storage, uint32_t width, uint32_t height) {
    return texture_create(device, storage, width, height, 1);
}

uint64_t lc_hip_volume_create(uint64_t device, uint32_t storage, uint32_t width, uint32_t height, uint32_t depth) {
    return texture_create(device, storage, width, height, depth);
}

void lc_hip_texture_destroy(uint64_t device, uint64_t texture) {
    Device *dev = obj<Device>(device);
    Texture *t = obj<Texture>(texture);
    if (!dev || !t) return;
    if (defer_while_capturing(dev, [device, texture] { lc_hip_texture_destroy(device, texture); })) return;
    DeviceGuard g(dev);
    (void)hipDeviceSynchronize();
    if (t->ptr) (void)hipFree(t->ptr);
    delete t;
}

int lc_hip_texture_upload(uint64_t device, uint64_t stream, uint64_t texture, const void *data) {
    Device *dev = obj<Device>(device);
    Texture *t = obj<Texture>(texture);
    if (!dev || !t || !data) return fail(LC_HIP_ERROR_INVALID_HANDLE, "bad image upload arguments");
    DeviceGuard g(dev);
    hipStream_t s = native_stream(stream);
    // the source is consumed before the call returns, as for buffer uploads
    LC_HIP_TRY(hipMemcpyAsync(t->ptr, data, (size_t)t->width * t->height * t->depth * t->pixel_bytes, hipMemcpyHostToDevice, s));
    LC_HIP_TRY(hipStreamSynchronize(s));
    return LC_HIP_OK;
}

int lc_hip_texture_download(uint64_t device, uint64_t stream, uint64_t texture, void *data) {
    Device *dev = obj<Device>(device);
    Texture *t = obj<Texture>(texture);
    if (!dev || !t || !data) return fail(LC_HIP_ERROR_INVALID_HANDLE, "bad image download arguments");
    DeviceGuard g(dev);
    hipStream_t s = native_stream(stream);
    LC_HIP_TRY(hipMemcpyAsync(data, t->ptr, (size_t)t->width * t->height * t->depth * t->pixel_bytes, hipMemcpyDeviceToHost, s));
    return LC_HIP_OK;
}

int lc_hip_texture_buffer_copy(uint64_t device, uint64_t stream, uint64_t texture, uint64_t buffer, size_t buffer_offset,
                               int to_texture) {
    Device *dev = obj<Device>(device);
    Texture *t = obj<Texture>(texture);
    Buffer *b = obj<Buffer>(buffer);
    if (!dev || !t || !b) return fail(LC_HIP_ERROR_INVALID_HANDLE, "bad image / buffer copy handles");
    const size_t bytes = (size_t)t->width * t->height * t->depth * t->pixel_bytes;
    if (buffer_offset > b->bytes || bytes > b->bytes - buffer_offset)
        return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "image copy outside the buffer");
    DeviceGuard g(dev);
    char *bp = static_cast<char *>(b->ptr) + buffer_offset;
    LC_HIP_TRY(hipMemcpyAsync(to_texture ? t->ptr : (void *)bp, to_texture ? (const void *)bp : t->ptr, bytes,
                              hipMemcpyDeviceToDevice, native_stream(stream)));
    return LC_HIP_OK;
}

int lc_hip_texture_copy(uint64_t device, uint64_t stream, uint64_t src, uint64_t dst) {
    Device *dev = obj<Device>(device);
    Texture *a = obj<Texture>(src), *b = obj<Texture>(dst);
    if (!dev || !a || !b) return fail(LC_HIP_ERROR_INVALID_HANDLE, "bad image copy handles");
    if (a->width != b->width || a->height != b->height || a->depth != b->depth || a->pixel_bytes != b->pixel_bytes)
        return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "image copy between images of different sizes");
    DeviceGuard g(dev);
    LC_HIP_TRY(hipMemcpyAsync(b->ptr, a->ptr, (size_t)a->width * a->height * a->depth * a->pixel_bytes,
                              hipMemcpyDeviceToDevice, native_stream(stream)));
    return LC_HIP_OK;
}

int lc_hip_texture_info(uint64_t device, uint64_t texture, uint32_t *storage, uint32_t size[3]) {
    Texture *t = obj<Texture>(texture);
    if (!obj<Device>(device) || !t) return fail(LC_HIP_ERROR_INVALID_HANDLE, "bad image handle");
    if (storage) *storage = t->storage;
    if (size) {
        size[0] = t->width;
        size[1] = t->height;
        size[2] = t->depth;
    }
    return LC_HIP_OK;
}

// ---- bindless arrays (buffer slots): create_bindless_array + BindlessArrayUpdateCommand ----
uint64_t lc_hip_bindless_create(uint64_t device, size_t slots) {
    Device *dev = obj<Device>(device);
    if (!dev || slots == 0 || slots > (1u << 24)) {
        fail(LC_HIP_ERROR_INVALID_ARGUMENT, "bindless array needs 1..2^24 slots");
        return LC_HIP_INVALID_HANDLE;
    }
    DeviceGuard g(dev);
    auto b = std::make_unique<Bindless>();
    b->dev = dev;
    b->host.assign(slots, BindlessSlot{nullptr, 0});
    b->host_tex.assign(slots, BindlessTex2D{nullptr, 0, 0, 0, 0, {0, 0}});
    hipError_t e = hipMalloc(&b->slots, slots * sizeof(BindlessSlot));
    if (e == hipSuccess) e = hipMemset(b->slots, 0, slots * sizeof(BindlessSlot));
    if (e == hipSuccess) e = hipMalloc(&b->tex, slots * sizeof(BindlessTex2D));
    if (e == hipSuccess) e = hipMemset(b->tex, 0, slots * sizeof(BindlessTex2D));
    if (e != hipSuccess) {
        if (b->slots) (void)hipFree(b->slots);
        if (b->tex) (void)hipFree(b->tex);
        hip_fail(e, "bindless array");
        return LC_HIP_INVALID_HANDLE;
    }
    return reinterpret_cast<uint64_t>(b.release());
}

void lc_hip_bindless_destroy(uint64_t device, uint64_t array) {
    Device *dev = obj<Device>(device);
    Bindless *b = obj<Bindless>(array);
    if (!dev || !b) return;
    if (defer_while_capturing(dev, [device, array] { lc_hip_bindless_destroy(device, array); })) return;
    DeviceGuard g(dev);
    (void)hipDeviceSynchronize();
    if (b->slots) (void)hipFree(b->slots);
    if (b->tex) (void)hipFree(b->tex);
    delete b;
}

int lc_hip_bindless_update(uint64_t device, uint64_t stream, uint64_t array, const lc_hip_bindless_modification *mods,
                           size_t count) {
    Device *dev = obj<Device>(device);
    Bindless *b = obj<Bindless>(array);
    if (!dev || !b) return fail(LC_HIP_ERROR_INVALID_HANDLE, "bad bindless update handles");
    if (count > 0 && !mods) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "null bindless modifications");
    std::vector<BindlessSlot> next = b->host;
    std::vector<BindlessTex2D> next_tex = b->host_tex;
    for (size_t i = 0; i < count; i++) {
        const lc_hip_bindless_modification &m = mods[i];
        if (m.slot >= next.size()) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "bindless slot out of range");
        if (m.texture_op == LC_HIP_BINDLESS_EMPLACE) {
            Texture *t = obj<Texture>(m.texture);
            if (!t) return fail(LC_HIP_ERROR_INVALID_HANDLE, "bad bindless texture");
            if (t->depth != 1) return fail(LC_HIP_ERROR_UNSUPPORTED, "bindless texture slots hold 2-D images");
            if (m.sampler > 15u) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "unknown sampler code");
            next_tex[m.slot] = BindlessTex2D{t->ptr, t->width, t->height, t->storage, m.sampler, {0, 0}};
        } else if (m.texture_op == LC_HIP_BINDLESS_REMOVE) {
            next_tex[m.slot] = BindlessTex2D{nullptr, 0, 0, 0, 0, {0, 0}};
        } else if (m.texture_op != LC_HIP_BINDLESS_NONE) {
            return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "unknown bindless texture operation");
        }
        if (m.buffer_op == LC_HIP_BINDLESS_EMPLACE) {
            Buffer *buf = obj<Buffer>(m.buffer);
            if (!buf) return fail(LC_HIP_ERROR_INVALID_HANDLE, "bad bindless buffer");
            if (m.offset_bytes > buf->bytes) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "bindless offset beyond the buffer");
            next[m.slot] = BindlessSlot{static_cast<char *>(buf->ptr) + m.offset_bytes, buf->bytes - m.offset_bytes};
        } else if (m.buffer_op == LC_HIP_BINDLESS_REMOVE) {
            next[m.slot] = BindlessSlot{nullptr, 0};
        } else if (m.buffer_op != LC_HIP_BINDLESS_NONE) {
            return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "unknown bindless operation");
        }
    }
    DeviceGuard g(dev);
    hipStream_t s = native_stream(stream);
    LC_HIP_TRY(hipMemcpyAsync(b->slots, next.data(), next.size() * sizeof(BindlessSlot), hipMemcpyHostToDevice, s));
    LC_HIP_TRY(hipMemcpyAsync(b->tex, next_tex.data(), next_tex.size() * sizeof(BindlessTex2D), hipMemcpyHostToDevice, s));
    LC_HIP_TRY(hipStreamSynchronize(s)); // the staging is this call's
    b->host.swap(next);
    b->host_tex.swap(next_tex);
    return LC_HIP_OK;
}

// ---- run-time shaders (create_shader / ShaderDispatchCommand over HIP source, shader.cpp) ----
static void copy_log(const std::string &log, char *out, size_t cap) {
    if (!out || cap == 0) return;
    const size_t n = std::min(cap - 1, log.size());
    std::memcpy(out, log.data(), n);
    out[n] = '\0';
}

int lc_hip_shader_cache_stats(uint64_t *hits, uint64_t *misses) {
    rtc_cache_stats(hits, misses);
    return LC_HIP_OK;
}

uint64_t lc_hip_shader_cache_key(const char *source) { return source ? rtc_cache_key(source) : 0; }

int lc_hip_shader_compile(const char *source, char *log, size_t log_capacity, size_t *code_size) {
    if (!source) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "null shader source");
    std::vector<char> code;
    std::string msg;
    const bool ok = rtc_compile(source, "shader.hip", code, msg);
    copy_log(msg, log, log_capacity);
    if (code_size) *code_size = code.size();
    if (!ok) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "shader compile failed: " + msg.substr(0, 2000));
    return LC_HIP_OK;
}

uint64_t lc_hip_shader_create(uint64_t device, const char *source, const char *entry, uint32_t block_x,
                              uint32_t block_y, uint32_t block_z, char *log, size_t log_capacity) {
    Device *dev = obj<Device>(device);
    if (!dev || !source || !entry) {
        fail(LC_HIP_ERROR_INVALID_ARGUMENT, "bad shader create arguments");
        return LC_HIP_INVALID_HANDLE;
    }
    const uint64_t threads = (uint64_t)block_x * block_y * block_z;
    if (threads == 0 || threads > 1024) {
        fail(LC_HIP_ERROR_INVALID_ARGUMENT, "shader block size must hold 1..1024 threads");
        return LC_HIP_INVALID_HANDLE;
    }
    std::vector<char> code;
    std::string msg;
    const bool ok = rtc_compile(source, "shader.hip", code, msg);
    copy_log(msg, log, log_capacity);
    if (!ok) {
        fail(LC_HIP_ERROR_INVALID_ARGUMENT, "shader compile failed: " + msg.substr(0, 2000));
        return LC_HIP_INVALID_HANDLE;
    }
    DeviceGuard g(dev);
    auto sh = std::make_unique<Shader>();
    sh->dev = dev;
    sh->block[0] = block_x; sh->block[1] = block_y; sh->block[2] = block_z;
    // a failed load or lookup is reported here and taken off HIP's last-error slot, which later
    // launches read with hipGetLastError
    hipError_t e = hipModuleLoadData(&sh->module, code.data());
    if (e != hipSuccess) {
        (void)hipGetLastError();
        hip_fail(e, "shader module load");
        return LC_HIP_INVALID_HANDLE;
    }
    e = hipModuleGetFunction(&sh->fn, sh->module, entry);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        (void)hipModuleUnload(sh->module);
        hip_fail(e, "shader entry lookup");
        return LC_HIP_INVALID_HANDLE;
    }
    return reinterpret_cast<uint64_t>(sh.release());
}

int lc_hip_shader_set_argument_usage(uint64_t device, uint64_t shader, const uint32_t *usage, size_t count) {
    Shader *sh = obj<Shader>(shader);
    if (!obj<Device>(device) || !sh || (count && !usage)) return fail(LC_HIP_ERROR_INVALID_HANDLE, "bad shader usage arguments");
    for (size_t i = 0; i < count; i++)
        if (usage[i] > 3u) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "usage is NONE, READ, WRITE or READ_WRITE");
    sh->usage.assign(usage, usage + count);
    return LC_HIP_OK;
}

int lc_hip_shader_argument_usage(uint64_t device, uint64_t shader, size_t index) {
    Shader *sh = obj<Shader>(shader);
    if (!obj<Device>(device) || !sh) return fail(LC_HIP_ERROR_INVALID_HANDLE, "bad shader");
    if (index >= sh->usage.size()) return 3; // unknown: read and write, as a conservative answer
    return (int)sh->usage[index];
}

void lc_hip_shader_destroy(uint64_t device, uint64_t shader) {
    Device *dev = obj<Device>(device);
    Shader *sh = obj<Shader>(shader);
    if (!dev || !sh) return;
    if (defer_while_capturing(dev, [device, shader] { lc_hip_shader_destroy(device, shader); })) return;
    DeviceGuard g(dev);
    hipDeviceSynchronize();
    if (sh->module) hipModuleUnload(sh->module);
    delete sh;
}

// Packs the arguments (16-B aligned, in order), then launches once per dispatch size with
// ls_kid = {size, kernel id} appended; sizes: 3 per dispatch (ids 0..count-1) or, for indirect
// dispatches, read from the device.
static int shader_launch(uint64_t device, uint64_t stream, uint64_t shader, const lc_hip_shader_argument *args,
                         size_t argc, const void *uniforms, size_t uniform_size, const uint32_t *sizes,
                         size_t count, const uint32_t *kernel_ids) {
    Device *dev = obj<Device>(device);
    Shader *sh = obj<Shader>(shader);
    if (!dev || !sh) return fail(LC_HIP_ERROR_INVALID_HANDLE, "bad shader dispatch handles");
    if (argc > 0 && !args) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "null shader arguments");
    constexpr size_t kMaxArgBytes = 4096; // kernel argument segment
    alignas(16) unsigned char buf[kMaxArgBytes];
    size_t off = 0;
    auto put = [&](const void *p, size_t n) -> bool {
        off = (off + 15) & ~size_t(15);
        if (off + n > kMaxArgBytes) return false;
        std::memcpy(buf + off, p, n);
        off += n;
        return true;
    };
    for (size_t i = 0; i < argc; i++) {
        const lc_hip_shader_argument &x = args[i];
        if (x.tag == LC_HIP_ARGUMENT_BUFFER) {
            Buffer *b = obj<Buffer>(x.handle);
            if (!b) return fail(LC_HIP_ERROR_INVALID_HANDLE, "bad buffer argument");
            if (x.offset > b->bytes) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "buffer argument offset outside its buffer");
            const uint64_t size = x.size ? x.size : b->bytes - x.offset;
            if (x.offset + size > b->bytes) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "buffer argument outside its buffer");
            const uint64_t binding[2] = {reinterpret_cast<uint64_t>(static_cast<char *>(b->ptr) + x.offset), size};
            if (!put(binding, sizeof(binding))) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "shader arguments exceed 4 KiB");
        } else if (x.tag == LC_HIP_ARGUMENT_ACCEL) {
            Accel *a = obj<Accel>(x.handle);
            if (!a) return fail(LC_HIP_ERROR_INVALID_HANDLE, "bad accel argument");
            std::lock_guard<std::mutex> lk(a->mu);
            if (!a->built || !a->desc) return fail(LC_HIP_ERROR_NOT_BUILT, "accel argument is not built");
            uint32_t need = 0;
            if (a->inst.size() == 1) {
                need = 2 * (a->host_records[0].blas_depth + 2);
            } else if (a->inst.size() > 1) {
                const int rc = accel_tlas_sync(a);
                if (rc != LC_HIP_OK) return rc;
                need = 2 * (a->tlas_stats.depth + 2 + a->max_blas_depth + 2);
            }
            if (need > 64) return fail(LC_HIP_ERROR_UNSUPPORTED, "accel too deep for a shader's traversal stack (64)");
            a->state_written = true; // the shader may write instances: fold them at the next build
            const uint64_t binding[2] = {reinterpret_cast<uint64_t>(a->desc), 0};
            if (!put(binding, sizeof(binding))) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "shader arguments exceed 4 KiB");
        } else if (x.tag == LC_HIP_ARGUMENT_TEXTURE) {
            Texture *t = obj<Texture>(x.handle);
            if (!t) return fail(LC_HIP_ERROR_INVALID_HANDLE, "bad image argument");
            const uint64_t binding[4] = {reinterpret_cast<uint64_t>(t->ptr), (uint64_t)t->width | ((uint64_t)t->height << 32),
                                         (uint64_t)t->storage | ((uint64_t)t->depth << 32), 0};
            if (!put(binding, sizeof(binding))) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "shader arguments exceed 4 KiB");
        } else if (x.tag == LC_HIP_ARGUMENT_BINDLESS_ARRAY) {
            Bindless *b = obj<Bindless>(x.handle);
            if (!b) return fail(LC_HIP_ERROR_INVALID_HANDLE, "bad bindless array argument");
            const uint64_t binding[4] = {reinterpret_cast<uint64_t>(b->slots), (uint64_t)b->host.size(),
                                         reinterpret_cast<uint64_t>(b->tex), 0};
            if (!put(binding, sizeof(binding))) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "shader arguments exceed 4 KiB");
        } else if (x.tag == LC_HIP_ARGUMENT_UNIFORM) {
            if (!uniforms || x.offset + x.size > uniform_size)
                return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "uniform argument outside the uniform bytes");
            if (!put(static_cast<const char *>(uniforms) + x.offset, x.size))
                return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "shader arguments exceed 4 KiB");
        } else {
            return fail(LC_HIP_ERROR_UNSUPPORTED, "unknown shader argument kind");
        }
    }
    const size_t args_end = off;
    MarkerRange mr(sh->name, "dispatch");
    DeviceGuard g(dev);
    hipStream_t s = native_stream(stream);
    for (size_t k = 0; k < count; k++) {
        const uint32_t dx = sizes[3 * k], dy = sizes[3 * k + 1], dz = sizes[3 * k + 2];
        off = args_end;
        const uint32_t ls_kid[4] = {dx, dy, dz, kernel_ids ? kernel_ids[k] : (uint32_t)k};
        if (!put(ls_kid, sizeof(ls_kid))) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "shader arguments exceed 4 KiB");
        if ((uint64_t)dx * dy * dz == 0) continue;
        const uint32_t gx = (dx + sh->block[0] - 1) / sh->block[0];
        const uint32_t gy = (dy + sh->block[1] - 1) / sh->block[1];
        const uint32_t gz = (dz + sh->block[2] - 1) / sh->block[2];
        // the kernel's one argument is the packed struct (its size, `off`, comes from the code
        // object's argument metadata); passed as a kernel parameter, which stream capture records
        void *params[] = {buf};
        hipError_t e = hipModuleLaunchKernel(sh->fn, gx, gy, gz, sh->block[0], sh->block[1], sh->block[2], 0, s,
                                             params, nullptr);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return hip_fail(e, "shader launch");
        }
    }
    return LC_HIP_OK;
}

int lc_hip_shader_dispatch(uint64_t device, uint64_t stream, uint64_t shader, const lc_hip_shader_argument *args,
                           size_t argc, const void *uniforms, size_t uniform_size, uint32_t dx, uint32_t dy,
                           uint32_t dz) {
    const uint32_t size[3] = {dx, dy, dz};
    return shader_launch(device, stream, shader, args, argc, uniforms, uniform_size, size, 1, nullptr);
}

int lc_hip_shader_dispatch_multiple(uint64_t device, uint64_t stream, uint64_t shader, const lc_hip_shader_argument *args,
                                    size_t argc, const void *uniforms, size_t uniform_size, const uint32_t *sizes,
                                    size_t count) {
    if (count > 0 && !sizes) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "null dispatch sizes");
    return shader_launch(device, stream, shader, args, argc, uniforms, uniform_size, sizes, count, nullptr);
}

int lc_hip_shader_dispatch_indirect(uint64_t device, uint64_t stream, uint64_t shader, const lc_hip_shader_argument *args,
                                    size_t argc, const void *uniforms, size_t uniform_size, uint64_t indirect_buffer,
                                    uint32_t offset, uint32_t max_dispatch) {
    Device *dev = obj<Device>(device);
    Shader *sh = obj<Shader>(shader);
    Buffer *ib = obj<Buffer>(indirect_buffer);
    if (!dev || !sh || !ib) return fail(LC_HIP_ERROR_INVALID_HANDLE, "bad indirect dispatch handles");
    if (ib->bytes < 16) return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "indirect buffer without a header");
    const size_t capacity = (ib->bytes - 16) / 32;
    // the sizes are produced on the device: wait for them (the stream's earlier work included)
    DeviceGuard g(dev);
    hipStream_t s = native_stream(stream);
    std::vector<uint32_t> host((size_t)4 + 8 * capacity);
    LC_HIP_TRY(hipMemcpyAsync(host.data(), ib->ptr, 16 + 32 * capacity, hipMemcpyDeviceToHost, s));
    LC_HIP_TRY(hipStreamSynchronize(s));
    const size_t n = std::min<size_t>({(size_t)host[0], (size_t)max_dispatch, capacity > offset ? capacity - offset : 0});
    std::vector<uint32_t> sizes(3 * n), ids(n);
    for (size_t k = 0; k < n; k++) {
        const uint32_t *e = &host[4 + 8 * (offset + k)]; // {block x, y, z, pad, size x, y, z, kernel id}
        if ((e[0] || e[1] || e[2]) && (e[0] != sh->block[0] || e[1] != sh->block[1] || e[2] != sh->block[2]))
            return fail(LC_HIP_ERROR_INVALID_ARGUMENT, "indirect dispatch block size differs from the shader's");
        sizes[3 * k] = e[4]; sizes[3 * k + 1] = e[5]; sizes[3 * k + 2] = e[6];
        ids[k] = e[7];
    }
    return shader_launch(device, stream, shader, args, argc, uniforms, uniform_size, sizes.data(), n, ids.data());
}

} // extern "C"
