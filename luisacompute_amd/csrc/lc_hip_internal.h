// lc_hip_internal.h -- host-side objects of the HIP backend and the kernel entry points they use.
#pragma once

#include <hip/hip_runtime.h>

#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <functional>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "lc_hip_bvh.h"

namespace lchip {

// ---- kernels (bvh_build.hip) ----
size_t bvh_scratch_bytes(uint32_t n);
size_t bvh_node_capacity(uint32_t n);
hipError_t bvh_build(hipStream_t stream, void *scratch, size_t scratch_bytes, const void *vbase, uint64_t vstride,
                     uint64_t nverts, const uint32_t *tri, const float4 *boxes, uint32_t n, WideNode *nodes,
                     TriRecord *tris, uint32_t *prim_out, BuildCounters **counters_out, uint32_t split_small,
                     uint32_t leaf_max, uint32_t sah_refine, uint32_t width, int cu_count, uint32_t *error_flag,
                     uint32_t collapse_head);
hipError_t bvh_refit(hipStream_t stream, void *scratch, size_t scratch_bytes, const void *vbase, uint64_t vstride,
                     uint64_t nverts, const uint32_t *tri, uint32_t n, uint32_t depth, WideNode *nodes,
                     TriRecord *tris, BuildCounters **counters_out, int cu_count, uint32_t *error_flag);
// Refit state of a compacting build with allow_update (the build's scratch is not kept): the
// counters followed by 2 float4 per wide node; copy the build's counters into it (same stream).
size_t bvh_refit_state_bytes(uint32_t node_count);
BuildCounters *bvh_refit_state_counters(void *state);
hipError_t bvh_refit_from_state(hipStream_t stream, void *state, size_t state_bytes, uint32_t node_count,
                                const void *vbase, uint64_t vstride, uint64_t nverts, const uint32_t *tri, uint32_t n,
                                uint32_t depth, WideNode *nodes, TriRecord *tris, int cu_count, uint32_t *error_flag);
// Procedural primitives: packed 24-B AABBs -> 2 float4 per box; leaf records from the build's
// leaf-order primitive ids.
hipError_t aabbs_to_boxes(hipStream_t stream, const float *aabb, uint32_t n, float4 *boxes);
hipError_t proc_records(hipStream_t stream, const uint32_t *ids, const float4 *boxes, uint32_t n, ProcRecord *out);
hipError_t bvh_validate(hipStream_t stream, const WideNode *nodes, uint32_t node_count, const TriRecord *tris,
                        uint32_t tri_count, uint32_t nprims, uint32_t *marks, uint32_t *err7);

// ---- kernels (instances.hip) ----
// Item i reads (write = false: into data) or writes (from data) field `field` of instance idx[i];
// 12 floats per item for transforms, one uint32 otherwise. Indices >= count set *err.
hipError_t instance_ops(hipStream_t stream, InstanceState *st, uint32_t count, uint32_t field, bool write,
                        const uint32_t *idx, uint32_t n, void *data, uint32_t *err);

// ---- kernels (trace.hip) ----
// A batch of rays traced by one launch: rays [start, start + count) of the launch's queue are
// this batch's rays 0..count-1. Up to kMaxBatches batches ride in the kernel arguments.
constexpr int kMaxBatches = 64;
struct RayBatch {
    const float4 *rays; // 2 float4 per ray (Ray layout)
    void *out;          // 24-B SurfaceHit per ray (closest) or 1 byte per ray (any); null when counting
    uint32_t start;     // first queue index of the batch (ascending across batches)
    uint32_t mask;      // ray visibility mask of the batch
};
struct BatchTable {
    RayBatch b[kMaxBatches];
    uint32_t n;
};

// A trace launch splits its queue into kFetchSegments contiguous segments, one per XCD, each
// with its own slot counter on its own 64-byte line (ray_counter[s * kCounterStride]).
constexpr uint32_t kFetchSegments = 8;
constexpr uint32_t kCounterStride = 16;
constexpr uint32_t kRayCounterWords = kFetchSegments * kCounterStride;

struct TraceParams {
    BatchTable batches;
    uint32_t count;        // rays over all batches (the capacity when count_ptr is set)
    const uint32_t *count_ptr; // optional: ray count in device memory (queues built on the GPU)
    uint32_t *ray_counter;
    const uint32_t *order; // optional: queue indices in traversal order (sorted queue), else identity
    Geometry geo;
    uint2 *spill;
    uint32_t spill_depth; // spill entries per lane
    int steps;            // traversal steps between ray refills
    unsigned long long *stats; // [nodes visited, triangles tested] when counting
    uint32_t *error;
    uint32_t hit_record;  // closest hits carry the leaf record index in their padding word
    uint32_t fetch_chunk;    // queue slots a wave claims per atomic (0: exactly what its idle lanes need)
    uint32_t fetch_segments; // 1: one queue counter; kFetchSegments: one segment + counter per XCD
    uint32_t tri_quarters;   // closest hit: a step tests triangles when 4 x (lanes holding some) >=
                             // tri_quarters x (live lanes)
    uint32_t refill_below;   // a wave stops its steps early to refill once fewer lanes than this
                             // hold a ray (0: refill only after `steps` steps)
};

// ---- kernels (radix.hip) ----
// Stable sort of (key, value) pairs by key bits [begin_bit, end_bit); `vals_in` null sorts the
// identity (item indices). Items are the first min(*count_ptr, cap) when count_ptr is given.
// keys_in is left unchanged and must not alias keys_out; cap < 2^31.
size_t radix_sort_temp_bytes(uint32_t cap);
hipError_t radix_sort_pairs(hipStream_t stream, void *temp, size_t temp_bytes, const uint32_t *keys_in,
                            uint32_t *keys_out, const uint32_t *vals_in, uint32_t *vals_out, uint32_t cap,
                            const uint32_t *count_ptr, uint32_t begin_bit, uint32_t end_bit);
// The same with the ping-pong buffers given (cap words each, distinct from the inputs and outputs);
// count_temp holds radix_sort_count_bytes(cap) bytes of digit counts.
size_t radix_sort_count_bytes(uint32_t cap);
hipError_t radix_sort_pairs_with(hipStream_t stream, void *count_temp, size_t count_bytes, uint32_t *alt_keys,
                                 uint32_t *alt_vals, const uint32_t *keys_in, uint32_t *keys_out,
                                 const uint32_t *vals_in, uint32_t *vals_out, uint32_t cap, const uint32_t *count_ptr,
                                 uint32_t begin_bit, uint32_t end_bit);
// out[i] = in[0] + ... + in[i-1] (one workgroup); in and out may alias.
hipError_t exclusive_scan_u32(hipStream_t stream, const uint32_t *in, uint32_t *out, uint32_t n);
// The same over ceil(*count_ptr / div) words, the count read on the device.
hipError_t exclusive_scan_u32_counted(hipStream_t stream, const uint32_t *in, uint32_t *out, const uint32_t *count_ptr,
                                      uint32_t div);

// ---- kernels (raysort.hip) ----
// count_ptr (device word, may be null): only the first min(*count_ptr, n) rays are sorted.
size_t ray_sort_scratch_bytes(uint32_t n);
hipError_t ray_sort(hipStream_t stream, const BatchTable &batches, uint32_t n, const uint32_t *count_ptr,
                    const float bounds[6], void *scratch, size_t scratch_bytes, const uint32_t **order_out);

// Queue index -> batch (batches ordered by start).
__host__ __device__ inline uint32_t find_batch(const BatchTable &t, uint32_t q) {
    uint32_t b = 0;
    for (uint32_t step = kMaxBatches / 2; step > 0; step >>= 1)
        if (b + step < t.n && t.b[b + step].start <= q) b += step;
    return b;
}

// ---- kernels (pathtracer.hip) ----
// Shading data of one accel instance: its BLAS leaf records (the triangles' vertices as the
// build copied them), object->world affine, material.
struct PtInstance {
    const TriRecord *tris;
    float fwd[12];              // object -> world, row-major 3x4
    float albedo[3];
    uint32_t emissive;          // 1: a hit returns the light's emission
};

// One render's constants (kernel argument).
struct PtScene {
    const PtInstance *inst;
    float light_pos[3], light_u[3], light_v[3], light_e[3];
    float light_n[3], light_area; // normalize(cross(u, v)), |cross(u, v)|
    float cam_o[3], cam_r[3], cam_u[3], cam_f[3];
    uint32_t width, height, spp, sample_offset, max_depth, rr_depth;
    uint64_t seed, sequence;      // PCG32 initstate; pixel i draws from stream initseq = sequence + i
};

struct PtOptions {
    uint32_t wavefront;   // 0: megakernel, 1: wavefront
    uint32_t chunk_spp;   // wavefront: samples per wave
    uint32_t sort;        // wavefront: bucket-sort continuation rays by origin cell + direction octant
    uint32_t graph;       // wavefront: capture the launches in a hipGraph and replay it
    uint32_t fetch_chunk;    // wavefront traces: queue slots a wave claims per atomic (see TraceParams)
    uint32_t fetch_segments; // wavefront traces: 1 or kFetchSegments queue segments
};

// One job's share of a wavefront wave: samples [sample0, sample0 + nsamples) of every pixel of
// one image. A wave holds up to kMaxWaveJobs jobs; its paths are laid out job after job,
// sample-major then pixel within a job.
constexpr uint32_t kMaxWaveJobs = 256;
struct PtJob {
    float4 *acc;          // running sum over the job's samples (float4 per pixel)
    float4 *out;          // final image, written by the job's last wave (may equal acc)
    uint32_t width, height;
    uint32_t sample0;     // absolute index of the job's first sample in this wave
    uint32_t nsamples;    // samples of the job in this wave
    uint32_t path_base;   // first path of the job in the wave
    uint32_t pix_base;    // first accumulate thread of the job in the wave
    uint32_t first;       // 1: the accumulator starts from zero
    uint32_t finish_spp;  // != 0: the wave ends the job: out = acc / finish_spp with alpha 1
    uint64_t seed, sequence; // PCG32 initstate; pixel i draws from stream initseq = sequence + i
};
static_assert(sizeof(PtJob) == 64, "PtJob is 64 bytes");
struct PtWaveTable {
    uint32_t njobs, paths, pixels, _pad;
    PtJob jobs[kMaxWaveJobs];
};

// A scene as the wavefront kernels see it.
struct PtBinding {
    Geometry G;
    bool two = false;
    uint32_t stack_need = 0;
    float bounds[6] = {0, 0, 0, 0, 0, 0};
    PtScene S{};          // light, camera, max_depth, rr_depth (per-render fields unused)
    std::vector<PtInstance> inst;
};

struct Device;
// Enqueues one render on `stream`; image = float4 per pixel (mean radiance over the samples, 1).
int pt_render(Device *dev, hipStream_t stream, const Geometry &G, bool two, uint32_t stack_need, const float bounds[6],
              const PtScene &scene, const std::vector<PtInstance> &inst, const PtOptions &opt, float4 *image,
              unsigned long long *counts, std::string &err);
// Enqueues one wavefront wave of `njobs` jobs (at most `capacity` paths) on `stream`, with the
// stream's workspace; counts[0..1] += closest-hit / shadow rays traced.
int pt_wave(Device *dev, hipStream_t stream, const PtBinding &B, const PtJob *jobs, uint32_t njobs, size_t capacity,
            const PtOptions &opt, unsigned long long *counts, std::string &err);
void pt_release(Device *dev);
void pt_release_stream(Device *dev, hipStream_t stream);

// ---- serving loop (serve.cpp) ----
struct ServeOptions {
    size_t hbm_budget = 1ull << 30; // accumulator arena (admission budget), bytes
    size_t wave_paths = 32u << 20;  // path capacity of a scene's wave
    uint32_t max_active = 64;       // admitted requests at once
    PtOptions pt{};                 // wavefront options of the waves
};
struct ServeStatus {
    int state;
    uint32_t samples_done;
    double t_submit, t_admit, t_done; // seconds since the server started (-1: not yet)
};
struct ServeStats {
    uint64_t waves, rays_closest, rays_shadow, requests_done;
    uint32_t peak_active, active, queued;
    uint64_t peak_bytes;
};
struct Server;
Server *serve_create(Device *dev, const ServeOptions &opt, std::string &err);
void serve_destroy(Server *s);
int serve_add_scene(Server *s, const PtBinding &B, std::string &err); // scene id, or -1
int serve_submit(Server *s, uint32_t scene, uint32_t width, uint32_t height, uint32_t spp, uint64_t seed,
                 uint64_t sequence, float4 *out, uint64_t *id, std::string &err);
int serve_query(Server *s, uint64_t id, ServeStatus *st);   // -1: unknown id
int serve_wait(Server *s, uint64_t id, double timeout_s);   // 0 done, 1 timeout, -1 unknown, -2 failed
int serve_drain(Server *s, double timeout_s);               // 0: every request finished, 1 timeout
void serve_stats(Server *s, ServeStats *st);

// ---- ray queries (query.hip, ray_query.h) ----
// built-in candidate callbacks of the batched query (values of lc_hip_query_policy)
enum { kQuerySurfaceCommit = 0, kQuerySurfaceIgnore = 1, kQuerySurfaceCutout = 2 };
enum { kQueryProceduralIgnore = 0, kQueryProceduralSphere = 1 };
struct QueryParams {
    const float4 *rays; // 2 float4 per ray
    void *hits;         // 24-B CommittedHit per ray
    const uint32_t *order; // optional: queue slot -> ray (sorted queue), else identity
    uint32_t *counter;  // queue slot counter (zeroed by the launch)
    uint32_t count;
    uint32_t mask;
    uint32_t first;     // terminate on the first commit (traverse_any)
    QueryGeometry geo;
    uint2 *spill;
    uint32_t spill_depth;
    uint32_t *error;
    uint32_t surface_mode;
    float cutout;
    uint32_t procedural_mode;
};
int query_block();
int query_lds_entries();
int query_blocks_per_cu();
hipError_t query_launch(const QueryParams &P, hipStream_t stream, int grid);

bool trace_config_supported(int block, int slds);
int trace_blocks_per_cu(bool any, bool two, bool count, int block, int slds);
// Working memory of the no-allocation trace (trace.hip): a static device array of
// kNoAllocScratchBytes per GPU; the lookup runs once per device, with that device current.
constexpr size_t kNoAllocScratchBytes = 32ull << 20;
constexpr size_t kNoAllocWordsBytes = 4096; // slot counters, error word, stats; the spill area follows
hipError_t noalloc_scratch_lookup(void **device_ptr);
hipError_t trace_launch(const TraceParams &P, hipStream_t stream, bool any, bool two, bool count, int grid, int block,
                        int slds);

// ---- host objects ----
struct TraceConfig {
    int block = 256;
    int slds = 16;
    int steps = 32;
    int blocks_per_cu = 0; // 0: occupancy limit
    bool sort_rays = false; // traverse a queue sorted by origin Morton code + direction octant
    uint32_t fetch_chunk = 0;    // see TraceParams
    uint32_t fetch_segments = 1;
    uint32_t tri_quarters = 2;   // see TraceParams
    uint32_t refill_below = 0;
};

struct TimingStat {
    double total_ms = 0.0;
    uint64_t launches = 0;
};

// Working memory of the trace, query and build launches of one stream: launches on different
// streams never share it (streams run concurrently); launches on one stream are ordered by the
// stream. words: [0, 512) the per-segment slot counters, 512 the error word, 576 the two 64-bit
// counter-mode totals.
struct StreamScratch {
    uint32_t *words = nullptr;
    uint32_t *ray_counter = nullptr;
    uint32_t *error = nullptr;
    unsigned long long *stats = nullptr;
    void *spill = nullptr;
    size_t spill_bytes = 0;
    void *sort = nullptr;
    size_t sort_bytes = 0;
    // graphs captured on the stream replay launches that name these buffers: while any such
    // graph (or a capture in progress) holds a pin, a buffer that has to grow is retired instead
    // of freed, and the scratch outlives its stream; retired buffers go when the last pin does
    uint32_t pins = 0;
    std::vector<void *> retired;
};
constexpr size_t kStreamWordsBytes = 1024;

struct Device {
    int index = 0;
    int cu_count = 0;
    std::string name;
    std::mutex mu;
    TraceConfig cfg[2]; // [closest, any]
    uint32_t split_small = 1;
    uint32_t sah_refine = 1; // SAH refinement of FAST_TRACE builds (LUISA_HIP_BVH_SAH=0: off)
    uint32_t bvh_width = 8;  // children per wide node, 2..8 (LUISA_HIP_BVH_WIDTH)
    uint32_t collapse_head = 0; // wide levels launched over the GPU before the one-workgroup tail (0: automatic)
    // error word of run-time shaders' traversal stacks (DeviceAccel::error)
    uint32_t *error = nullptr;
    // per-stream working memory of trace / query / build launches (guarded by mu)
    std::map<hipStream_t, StreamScratch *> scratch;
    // build arena of compacting BLAS builds (guarded by mu): worst-case wide nodes, and the build
    // scratch of primitives without allow_update; free again once arena_done has completed
    WideNode *arena_nodes = nullptr;
    size_t arena_node_cap = 0;
    void *arena_scratch = nullptr;
    size_t arena_scratch_bytes = 0;
    hipEvent_t arena_done = nullptr;
    // kernel timing with HIP events on the launch stream
    bool timing = false;
    struct Pending {
        std::string name;
        hipEvent_t a, b;
    };
    std::vector<Pending> pending;
    std::map<std::string, TimingStat> timings;
    void *pt = nullptr; // path-tracer workspace (pathtracer.hip), released with the device
    // this GPU's copy of the no-allocation trace's static scratch array (null if the lookup failed)
    void *noalloc_scratch = nullptr;
    // stream captures in progress (guarded by mu): the stream's scratch, pinned for the capture
    std::map<hipStream_t, StreamScratch *> capturing;
    // scratch of destroyed streams still pinned by graphs (guarded by mu)
    std::vector<StreamScratch *> orphans;
    // destroys requested while a capture was open (their frees and device-wide waits would
    // invalidate it); run when the last capture on the device ends (guarded by mu)
    std::vector<std::function<void()>> deferred;
};

// Defers a destroy while a capture is open on the device: true when `fn` was queued (run when the
// last capture ends), false when the caller should destroy now.
bool defer_while_capturing(Device *dev, std::function<void()> fn);

// A captured sequence of launches (lc_hip_stream_capture_end), replayed by lc_hip_graph_launch.
struct Graph {
    Device *dev = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    StreamScratch *pinned = nullptr; // the capturing stream's scratch, which the launches name
};

// The stream's working memory, created on first use (caller holds dev->mu); null when the
// allocation fails. Released with the stream (after it drained) or with the device.
StreamScratch *stream_scratch(Device *dev, hipStream_t s);
void stream_scratch_release(Device *dev, hipStream_t s);

struct Buffer {
    Device *dev = nullptr;
    void *ptr = nullptr;
    size_t bytes = 0;
    size_t elem_stride = 0;
    bool owned = true; // false: a view of external device memory (not freed by destroy)
};

struct Stream {
    Device *dev = nullptr;
    hipStream_t s = nullptr;
    std::mutex mu;
    // host callbacks run on a worker thread once the GPU has passed their point in the stream
    std::thread worker;
    std::mutex qmu;
    std::condition_variable qcv;
    std::deque<std::pair<hipEvent_t, std::function<void()>>> q;
    bool stop = false;
};

struct HostCounters {
    uint32_t node_count, tri_count, depth, root_ref;
    int32_t cbounds[6];
    float bounds[6];
    uint32_t tail_level_count; // nodes at wide level kCollapseLaunches: must be 0
    uint32_t error;
};

struct Mesh {
    Device *dev = nullptr;
    std::string name; // DeviceInterface::set_name: build ranges are marked when set
    uint32_t hint = 0;
    bool allow_compaction = true, allow_update = false;
    std::mutex mu;
    uint32_t ntris = 0;
    size_t node_cap = 0, tri_cap = 0;
    WideNode *nodes = nullptr;
    TriRecord *tris = nullptr;
    void *scratch = nullptr;
    size_t scratch_bytes = 0;
    // compacting builds with allow_update: what a refit needs (bvh_refit_state_bytes), sized for
    // refit_nodes wide nodes
    void *refit_state = nullptr;
    size_t refit_state_bytes = 0;
    uint32_t refit_nodes = 0;
    HostCounters *host = nullptr; // pinned
    hipEvent_t ready = nullptr;   // counters copied to host
    bool built = false;
    bool stats_valid = false;
    HostCounters stats{};
    uint64_t build_epoch = 0;     // bumps on every build or refit (TLAS records refer to it)
    uint64_t refits = 0;          // builds served by a refit
    // procedural primitive (an AABB BLAS; created by lc_hip_procedural_create): leaf records in
    // `proc`, the build's input boxes and leaf-order ids in proc_boxes / proc_ids
    bool procedural = false;
    ProcRecord *proc = nullptr;
    float4 *proc_boxes = nullptr;
    uint32_t *proc_ids = nullptr;
    size_t proc_cap = 0;
    // buffer ranges of the last build: an update over the same ranges refits
    const char *v_addr = nullptr;
    const uint32_t *t_addr = nullptr;
    size_t v_size = 0, v_stride = 0, t_size = 0;
};

struct InstanceHost {
    Mesh *mesh = nullptr;
    float affine[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
    uint32_t vis_mask = 0xFFu;
    uint32_t user_id = 0;
    bool opaque = true;
};

struct Accel {
    Device *dev = nullptr;
    std::string name; // DeviceInterface::set_name: build ranges are marked when set
    uint32_t hint = 0;
    bool allow_compaction = true, allow_update = false;
    std::mutex mu;
    std::vector<InstanceHost> inst;
    InstanceRecord *records = nullptr;
    size_t record_cap = 0;
    std::vector<InstanceRecord> host_records;
    // TLAS
    float4 *boxes = nullptr;
    size_t box_cap = 0;
    WideNode *tlas_nodes = nullptr;
    uint32_t *tlas_prims = nullptr;
    size_t tlas_cap = 0;
    void *tlas_scratch = nullptr;
    size_t tlas_scratch_bytes = 0;
    HostCounters *tlas_host = nullptr;
    hipEvent_t tlas_ready = nullptr;
    bool tlas_valid = false;      // TLAS matches the instance set
    bool tlas_stats_valid = false;
    HostCounters tlas_stats{};
    bool built = false;
    uint32_t max_blas_depth = 0;
    float world_bounds[6] = {0, 0, 0, 0, 0, 0}; // union of instance world boxes (ray-sort keys)
    // device-side instance table for the in-kernel instance ops (state[count] is the
    // out-of-range-index flag word); written at every build, folded back into `inst` before the
    // next build when instance writes were issued
    InstanceState *state = nullptr;
    size_t state_cap = 0;
    uint32_t state_count = 0;
    bool state_written = false;
    // descriptor a run-time shader's accel argument points at (lc_hip_device.h DeviceAccel),
    // rewritten by every build
    void *desc = nullptr;
    DeviceAccel desc_host{};
    // pinned staging of a build's uploads (instance records, instance table, TLAS input boxes,
    // descriptor): the copies read it after the build call returned; `staged` marks their end and
    // is waited for before the next build writes it again
    char *stage = nullptr;
    size_t stage_cap = 0;
    hipEvent_t staged = nullptr;
    bool stage_pending = false;
    uint32_t tlas_count = 0; // instance count of the last full build (its TLAS, or its single instance)
};

// ---- images: 2-D, linear rows, one level ----
struct Texture {
    Device *dev = nullptr;
    void *ptr = nullptr;
    uint32_t storage = 0, width = 0, height = 0, pixel_bytes = 0;
    uint32_t depth = 1; // > 1 for volumes (3-D images)
};
// bytes per pixel of a PixelStorage value (include/luisa/runtime/rhi/pixel.h:14-35), 0 if unsupported
uint32_t pixel_storage_bytes(uint32_t storage);

// ---- bindless arrays (buffer slots) ----
struct Bindless {
    Device *dev = nullptr;
    BindlessSlot *slots = nullptr;  // device table of buffer slots
    BindlessTex2D *tex = nullptr;   // device table of 2-D texture slots
    std::vector<BindlessSlot> host; // the tables as last uploaded
    std::vector<BindlessTex2D> host_tex;
};

// ---- run-time shaders (hiprtc; shader.cpp) ----
struct Shader {
    Device *dev = nullptr;
    std::string name; // DeviceInterface::set_name: dispatch ranges are marked when set
    hipModule_t module = nullptr;
    hipFunction_t fn = nullptr;
    uint32_t block[3] = {128, 1, 1};
    std::vector<uint32_t> usage; // per argument: Usage (1 read, 2 write), from the frontend's AST
};
// The device library headers (lc_hip_device.h and what it includes), embedded at build time.
extern const char *const kDeviceHeaderNames[];
extern const char *const kDeviceHeaderSources[];
extern const int kDeviceHeaderCount;
// Compiles `source` for gfx950 with the device library; code object into `code`, diagnostics
// into `log`. False on a compile error.
bool rtc_compile(const char *source, const char *name, std::vector<char> &code, std::string &log);
uint64_t rtc_cache_key(const char *source);
void rtc_cache_stats(uint64_t *hits, uint64_t *misses);
// Writes the accel's DeviceAccel descriptor (allocating it on first use) on `stream`, copying
// from `staging` (pinned, sizeof(DeviceAccel) bytes, untouched until the copy ran).
int accel_write_descriptor(Accel *a, hipStream_t stream, void *staging);

} // namespace lchip
