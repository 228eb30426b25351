/*
 * lc_hip.h -- C ABI of the MI355X (gfx950) ray-tracing backend, library libluisa_compute_backend_hip.so.
 * The reference's own pure-C backend entry point (luisa_compute_lib_interface) is in lc_api.h.
 *
 * Plain C: handles are uint64_t (the backend object pointer, as in the reference where a
 * resource handle is the backend object's address), sizes are size_t, buffers are device
 * memory owned by the library. No C++ or torch types cross this boundary.
 *
 * Each entry point names the reference interface it stands in for. Paths are relative to the
 * LuisaCompute checkout:
 *   DeviceInterface           include/luisa/runtime/rhi/device_interface.h:95-212
 *   Command classes           include/luisa/runtime/rhi/command.h
 *   pure-C device fn-table    include/luisa/rust/api_types.h:597-654 (LCDeviceInterface)
 *
 * Error behaviour: the reference aborts on fatal errors and returns the invalid handle
 * (~0ull) for unsupported requests (include/luisa/runtime/rhi/resource.h:18-33). Here every
 * call that can fail returns LC_HIP_OK (0) or a negative status, creators return
 * LC_HIP_INVALID_HANDLE, and lc_hip_last_error() describes the last failure of the calling
 * thread; the C++ DeviceInterface layer turns these into the reference's abort.
 */
#ifndef LC_HIP_H
#define LC_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LC_HIP_INVALID_HANDLE UINT64_MAX

enum {
    LC_HIP_OK = 0,
    LC_HIP_ERROR_INVALID_ARGUMENT = -1,
    LC_HIP_ERROR_INVALID_HANDLE = -2,
    LC_HIP_ERROR_NOT_BUILT = -3,
    LC_HIP_ERROR_HIP = -4,
    LC_HIP_ERROR_OUT_OF_MEMORY = -5,
    LC_HIP_ERROR_UNSUPPORTED = -6,
    LC_HIP_ERROR_DEVICE_FAULT = -7,
    LC_HIP_ERROR_TIMEOUT = -8
};

/* Ray: include/luisa/runtime/rtx/ray.h:10-16 (32 B, align 16). */
typedef struct lc_hip_ray {
    float origin[3];
    float t_min;
    float direction[3];
    float t_max;
} lc_hip_ray;

/* SurfaceHit: include/luisa/runtime/rtx/hit.h:30-37 (24 B, align 8). Miss: inst = prim = ~0u,
   bary = (0, 0), t = ray t_max (src/rust/luisa_compute_backend_impl/src/cpu/accel.rs:500-507). */
typedef struct lc_hip_surface_hit {
    uint32_t inst;
    uint32_t prim;
    float bary[2];
    float committed_ray_t;
    uint32_t _pad;
} lc_hip_surface_hit;

/* CommittedHit: include/luisa/runtime/rtx/hit.h:18-26 (24 B, align 8), the result of a ray
   query. hit_type: LC_HIP_HIT_MISS / _SURFACE / _PROCEDURAL (HitType, hit.h:7-14). Miss: inst =
   prim = ~0u, bary = (0, 0), committed_ray_t = ray t_max; procedural hits report bary = (0, 0). */
#define LC_HIP_HIT_MISS 0u
#define LC_HIP_HIT_SURFACE 1u
#define LC_HIP_HIT_PROCEDURAL 2u
typedef struct lc_hip_committed_hit {
    uint32_t inst;
    uint32_t prim;
    float bary[2];
    uint32_t hit_type;
    float committed_ray_t;
} lc_hip_committed_hit;

/* AccelOption: include/luisa/runtime/rhi/resource.h:94-126. FAST_TRACE builds one primitive per
   BVH leaf slot, FAST_BUILD up to three (fewer nodes, shorter build, more triangle tests). */
#define LC_HIP_ACCEL_FAST_TRACE 0u
#define LC_HIP_ACCEL_FAST_BUILD 1u
typedef struct lc_hip_accel_option {
    uint32_t hint;
    uint8_t allow_compaction;
    uint8_t allow_update;
    uint8_t _pad[2];
} lc_hip_accel_option;

/* AccelBuildCommand::Modification: include/luisa/runtime/rhi/command.h:534-551 (80 B, align 16).
   flags: 1 primitive, 2 transform, 4 opaque on, 8 opaque off, 16 visibility, 32 user id. */
typedef struct lc_hip_accel_modification {
    uint32_t index;
    uint32_t user_id;
    uint32_t flags;
    uint32_t vis_mask;
    float affine[12]; /* row-major 3x4, object -> world (command.h:557-571) */
    uint64_t primitive; /* mesh handle */
    uint64_t _pad;
} lc_hip_accel_modification;

#define LC_HIP_MOD_PRIMITIVE 1u
#define LC_HIP_MOD_TRANSFORM 2u
#define LC_HIP_MOD_OPAQUE_ON 4u
#define LC_HIP_MOD_OPAQUE_OFF 8u
#define LC_HIP_MOD_VISIBILITY 16u
#define LC_HIP_MOD_USER_ID 32u

/* AccelBuildRequest: command.h:402-405 */
#define LC_HIP_BUILD_PREFER_UPDATE 0u
#define LC_HIP_BUILD_FORCE_BUILD 1u

const char *lc_hip_last_error(void);

/* ---- device: backend exports create/destroy/backend_device_names
   (src/backends/cuda/cuda_device.cpp:1192-1228, looked up at src/runtime/context.cpp:85-87);
   LCLibInterface::create_device (api_types.h:664) ---- */
int lc_hip_device_count(void);
/* index < 0 or SIZE_MAX-like values select device 0 (device_interface.h:85, cuda_device.cpp:1057). */
uint64_t lc_hip_device_create(int32_t index);
void lc_hip_device_destroy(uint64_t device);
/* DeviceInterface::compute_warp_size (device_interface.h:113): 64 on CDNA4. */
uint32_t lc_hip_device_warp_size(uint64_t device);
/* DeviceInterface::native_handle (device_interface.h:112): the HIP device ordinal. */
int lc_hip_device_native(uint64_t device);
/* DeviceInterface::query (device_interface.h:187): "device_name", "cu_count", "warp_size". */
int lc_hip_device_query(uint64_t device, const char *key, char *out, size_t capacity);

/* ---- buffers: DeviceInterface::create_buffer / destroy_buffer (device_interface.h:116-122),
   BufferCreationInfo {handle, native_handle, element_stride, total_size_bytes} ---- */
uint64_t lc_hip_buffer_create(uint64_t device, size_t element_stride, size_t element_count, void **native_handle);
void lc_hip_buffer_destroy(uint64_t device, uint64_t buffer);
/* create_buffer with external memory (device_interface.h:116-118, the `external_memory` argument):
   a buffer over existing device memory of element_stride * element_count bytes, which the
   caller keeps alive; destroy releases the handle only. */
uint64_t lc_hip_buffer_create_external(uint64_t device, size_t element_stride, size_t element_count, void *device_ptr);
/* BufferUploadCommand / BufferDownloadCommand / BufferCopyCommand (command.h:128-233). The
   upload source is consumed before the call returns; a download destination is valid after the
   stream is synchronized. */
int lc_hip_buffer_upload(uint64_t device, uint64_t stream, uint64_t buffer, size_t offset, size_t size,
                         const void *data);
int lc_hip_buffer_download(uint64_t device, uint64_t stream, uint64_t buffer, size_t offset, size_t size,
                           void *data);
int lc_hip_buffer_copy(uint64_t device, uint64_t stream, uint64_t src, size_t src_offset, uint64_t dst,
                       size_t dst_offset, size_t size);

/* ---- streams: create_stream / destroy_stream / synchronize_stream / dispatch callback
   (device_interface.h:136-139; CUDAStream::callback, cuda_stream.cpp:142-170) ---- */
uint64_t lc_hip_stream_create(uint64_t device, uint32_t stream_tag);
void lc_hip_stream_destroy(uint64_t device, uint64_t stream);
int lc_hip_stream_synchronize(uint64_t device, uint64_t stream);
/* Runs fn(user) on a backend thread after all work enqueued so far on the stream completed. */
int lc_hip_stream_add_callback(uint64_t device, uint64_t stream, void (*fn)(void *), void *user);
void *lc_hip_stream_native(uint64_t device, uint64_t stream);

/* DeviceInterface::set_name (device_interface.h:189; tags of Resource::Tag, resource.h:179-200):
   named meshes (3), procedural primitives (5), accels (7) and shaders (10) bracket their builds
   and dispatches with roctx marker ranges "<name>::build" / "::update" / "::compact" /
   "::dispatch" (rocprofv3 --marker-trace), as the CUDA backend does with NVTX
   (cuda_accel.cpp:28,68,101); other resources ignore the name. */
int lc_hip_set_name(uint64_t device, uint32_t resource_tag, uint64_t handle, const char *name);

/* ---- stream capture into HIP graphs. The reference's CUDA backend encodes each CommandList it is
   given (CUDAStream::dispatch, cuda_stream.cpp:194-203); a sequence submitted many times is
   recorded once here and replayed. Between begin and end the stream's shader dispatches, traces and
   copies are recorded, not run; traces run unsorted and within the working memory the stream already
   holds (run the sequence once before capturing it; otherwise LC_HIP_ERROR_UNSUPPORTED). Host
   waits, callbacks, events and builds are not capturable. A graph replays on any stream; its traces
   use the capturing stream's working memory, so replays must not overlap that stream's own traces.
   Destroy graphs before their device. ---- */
int lc_hip_stream_capture_begin(uint64_t device, uint64_t stream);
int lc_hip_stream_capture_end(uint64_t device, uint64_t stream, uint64_t *graph);
int lc_hip_graph_launch(uint64_t device, uint64_t stream, uint64_t graph);
void lc_hip_graph_destroy(uint64_t device, uint64_t graph);

/* ---- events: timeline fences (device_interface.h:163-168; CUDAEvent, cuda_event.cpp:21-63, a
   timeline semaphore). signal(stream, v) issues value v at the stream's current point; the
   timeline has reached v once a signal of a value >= v completed on the GPU (the initial value is
   0). wait(stream, v) orders the stream's later work after that point; synchronize(v) blocks the
   caller until it; is_completed(v) returns 1 if reached, else 0. A wait or synchronize for a
   value no signal >= v was issued for yet blocks the calling thread until another thread issues
   one (the GPU never spins on a timeline value). ---- */
uint64_t lc_hip_event_create(uint64_t device);
void lc_hip_event_destroy(uint64_t device, uint64_t event);
int lc_hip_event_signal(uint64_t device, uint64_t event, uint64_t stream, uint64_t value);
int lc_hip_event_wait(uint64_t device, uint64_t event, uint64_t stream, uint64_t value);
int lc_hip_event_synchronize(uint64_t device, uint64_t event, uint64_t value);
int lc_hip_event_is_completed(uint64_t device, uint64_t event, uint64_t value);

/* ---- meshes (BLAS): create_mesh / destroy_mesh (device_interface.h:171-172) and
   MeshBuildCommand (command.h:407-444): vertex buffer byte range and stride (a multiple of 16,
   src/runtime/rtx/mesh.cpp:42-49), triangle buffer byte range of uint32 triples. ---- */
uint64_t lc_hip_mesh_create(uint64_t device, const lc_hip_accel_option *option);
void lc_hip_mesh_destroy(uint64_t device, uint64_t mesh);
int lc_hip_mesh_build(uint64_t device, uint64_t stream, uint64_t mesh, uint32_t request, uint64_t vertex_buffer,
                      size_t vertex_offset, size_t vertex_size, size_t vertex_stride, uint64_t triangle_buffer,
                      size_t triangle_offset, size_t triangle_size);

/* ---- procedural primitives: create_procedural_primitive / destroy_procedural_primitive
   (device_interface.h:175-176) and ProceduralPrimitiveBuildCommand (command.h:489-510): a BLAS
   over a byte range of AABBs (6 floats, min then max: include/luisa/runtime/rtx/aabb.h:7-10).
   An accel instances it like a mesh (its handle goes where a mesh handle goes); plain traces do
   not see it (no intersection functions, cpu/accel.rs:476-484), ray queries report its
   candidates to the procedural callback. ---- */
uint64_t lc_hip_procedural_create(uint64_t device, const lc_hip_accel_option *option);
void lc_hip_procedural_destroy(uint64_t device, uint64_t primitive);
int lc_hip_procedural_build(uint64_t device, uint64_t stream, uint64_t primitive, uint32_t request,
                            uint64_t aabb_buffer, size_t aabb_offset, size_t aabb_size);

/* ---- accels (TLAS): create_accel / destroy_accel (device_interface.h:183-184) and
   AccelBuildCommand (command.h:531-616; semantics of AccelImpl::update,
   src/rust/luisa_compute_backend_impl/src/cpu/accel.rs:324-447). ---- */
uint64_t lc_hip_accel_create(uint64_t device, const lc_hip_accel_option *option);
void lc_hip_accel_destroy(uint64_t device, uint64_t accel);
int lc_hip_accel_build(uint64_t device, uint64_t stream, uint64_t accel, uint32_t request, uint32_t instance_count,
                       const lc_hip_accel_modification *modifications, size_t modification_count,
                       int update_instance_buffer_only);
/* Instance queries of the DSL ops RAY_TRACING_INSTANCE_TRANSFORM / _USER_ID / _VISIBILITY_MASK
   (include/luisa/ast/op.h:290-297), answered from the accel's instance table (host call; device
   writes issued since the last build are included). */
int lc_hip_accel_instance_info(uint64_t device, uint64_t accel, uint32_t index, float affine[12],
                               uint32_t *vis_mask, uint32_t *user_id);
/* The in-kernel instance ops RAY_TRACING_INSTANCE_TRANSFORM / _USER_ID / _VISIBILITY_MASK and
   RAY_TRACING_SET_INSTANCE_* (op.h:290-297; CPU backend cpu/stream.rs:582-660, cpu/accel.rs:537-579)
   as stream-ordered batches on the device: item i reads or writes `field` of instance
   indices[i] (uint32 buffer). Transforms are 12 floats per item (row-major 3x4 object -> world,
   as in lc_hip_accel_modification), user ids and visibility masks one uint32. Reads see the
   latest writes; writes reach traversal at the accel's next build (as the reference's dirty
   instances do at the next update). Out-of-range indices read zeros / are skipped and make the
   next build (or instance info) fail with LC_HIP_ERROR_INVALID_ARGUMENT. Writes of one batch to
   the same instance race, as in-kernel writes do. */
#define LC_HIP_INSTANCE_TRANSFORM 0u
#define LC_HIP_INSTANCE_USER_ID 1u
#define LC_HIP_INSTANCE_VISIBILITY 2u
int lc_hip_accel_instance_read(uint64_t device, uint64_t stream, uint64_t accel, uint32_t field, uint64_t index_buffer,
                               size_t index_offset, uint64_t out_buffer, size_t out_offset, uint32_t count);
int lc_hip_accel_instance_write(uint64_t device, uint64_t stream, uint64_t accel, uint32_t field,
                                uint64_t index_buffer, size_t index_offset, uint64_t in_buffer, size_t in_offset,
                                uint32_t count);

/* ---- traversal: the lowering of CallOp::RAY_TRACING_TRACE_CLOSEST / _TRACE_ANY
   (src/dsl/rtx/accel.cpp:136-150 -> lc_accel_trace_closest / lc_accel_trace_any,
   src/backends/cuda/cuda_builtin/cuda_device_resource.h:1783-1829) applied to a batch:
   hits[i] = accel.intersect(rays[i], {.visibility_mask = mask}). Byte offsets into the buffers;
   rays are lc_hip_ray, hits lc_hip_surface_hit, any-hit results one byte (bool) per ray. ---- */
int lc_hip_accel_trace_closest(uint64_t device, uint64_t stream, uint64_t accel, uint64_t ray_buffer,
                               size_t ray_offset, uint64_t hit_buffer, size_t hit_offset, uint32_t ray_count,
                               uint32_t vis_mask);
int lc_hip_accel_trace_any(uint64_t device, uint64_t stream, uint64_t accel, uint64_t ray_buffer, size_t ray_offset,
                           uint64_t hit_buffer, size_t hit_offset, uint32_t ray_count, uint32_t vis_mask);
/* Several ray batches traced by one persistent launch per 64 batches (the batches form one queue,
   sorted together when ray sorting is on); each batch has its own buffers and visibility mask.
   The same per-ray results as one trace call per batch. */
typedef struct lc_hip_ray_batch {
    uint64_t ray_buffer;
    size_t ray_offset;
    uint64_t hit_buffer; /* lc_hip_surface_hit per ray (closest) or one byte per ray (any) */
    size_t hit_offset;
    uint32_t ray_count;
    uint32_t vis_mask;
} lc_hip_ray_batch;
int lc_hip_accel_trace_closest_batched(uint64_t device, uint64_t stream, uint64_t accel,
                                       const lc_hip_ray_batch *batches, size_t batch_count);
int lc_hip_accel_trace_any_batched(uint64_t device, uint64_t stream, uint64_t accel, const lc_hip_ray_batch *batches,
                                   size_t batch_count);
/* Traces over a queue whose length a kernel produced on the device: rays and hit slots for
   `capacity` rays; the launch traces the first min(count, capacity), count being the uint32 at
   count_offset of count_buffer when the launch runs (a kernel dispatched over an
   IndirectDispatchBuffer calling accel.trace_closest / trace_any, include/luisa/runtime/
   dispatch_buffer.h; cuda_device_resource.h:1783-1829). With ray sorting on, the first
   min(count, capacity) rays are sorted on the device like any trace queue. */
int lc_hip_accel_trace_closest_indirect(uint64_t device, uint64_t stream, uint64_t accel, uint64_t ray_buffer,
                                        size_t ray_offset, uint64_t hit_buffer, size_t hit_offset, uint32_t capacity,
                                        uint64_t count_buffer, size_t count_offset, uint32_t vis_mask);
int lc_hip_accel_trace_any_indirect(uint64_t device, uint64_t stream, uint64_t accel, uint64_t ray_buffer,
                                    size_t ray_offset, uint64_t hit_buffer, size_t hit_offset, uint32_t capacity,
                                    uint64_t count_buffer, size_t count_offset, uint32_t vis_mask);
/* lcub DeviceRadixSort::SortPairs (include/luisa/backends/ext/cuda/lcub/device_radix_sort.h:15-27)
   for uint32 keys and uint32 values: a stable sort of the pairs by key bits [begin_bit, end_bit).
   The first call is the temp_storage_size query (num_items < 2^31). The sort reads keys_in and
   values_in, writes keys_out and values_out, and uses temp_buffer (all distinct buffers);
   values_in == 0 sorts the item indices 0..num_items-1. count_buffer != 0: only the first
   min(count, num_items) items are sorted, count being the uint32 at count_offset when the sort
   runs (later output words are left as they were). Allocates nothing; can be captured. */
int lc_hip_radix_sort_pairs_temp_bytes(uint32_t num_items, size_t *temp_bytes);
int lc_hip_radix_sort_pairs(uint64_t device, uint64_t stream, uint64_t temp_buffer, uint64_t keys_in,
                            uint64_t keys_out, uint64_t values_in, uint64_t values_out, uint32_t num_items,
                            int begin_bit, int end_bit, uint64_t count_buffer, size_t count_offset);
/* Counter-mode run of the same traversal: total BVH nodes visited and triangles tested. */
int lc_hip_accel_trace_count(uint64_t device, uint64_t stream, uint64_t accel, uint64_t ray_buffer, size_t ray_offset,
                             uint32_t ray_count, uint32_t vis_mask, int any, uint64_t *nodes_visited,
                             uint64_t *triangles_tested);
/* The same counter-mode run with its lane accounting (no reference counterpart: measurement).
   A lane takes one traversal step per loop iteration (a node visit, a triangle test, or nothing
   when the wave's step kind does not match what it holds); live_lane_steps counts the iterations
   lanes spent with a ray, wave_steps the iterations waves issued (x 64 = lane slots), and
   refill_rounds the times waves handed new rays to their idle lanes. */
typedef struct lc_hip_trace_steps {
    uint64_t nodes_visited;
    uint64_t triangles_tested;
    uint64_t live_lane_steps;
    uint64_t wave_steps;
    uint64_t refill_rounds;
} lc_hip_trace_steps;
int lc_hip_accel_trace_steps(uint64_t device, uint64_t stream, uint64_t accel, uint64_t ray_buffer, size_t ray_offset,
                             uint32_t ray_count, uint32_t vis_mask, int any, lc_hip_trace_steps *out);

/* Closest hits that allocate nothing. No counterpart in the reference: for callers that cannot
   allocate device memory (code under HIP graph capture, tight serving loops, harnesses that forbid
   hipMalloc). The same per-ray results as lc_hip_accel_trace_closest; its working memory (slot
   counters, traversal stack spill) is a static 32 MiB array per GPU held until the process ends,
   and rays are traced unsorted. Fails with LC_HIP_ERROR_UNSUPPORTED when one workgroup's stacks
   need more than that array. Calls that overlap on one GPU share the array and race. */
int lc_hip_accel_trace_closest_noalloc(uint64_t device, uint64_t stream, uint64_t accel, uint64_t ray_buffer,
                                       size_t ray_offset, uint64_t hit_buffer, size_t hit_offset, uint32_t ray_count,
                                       uint32_t vis_mask);

/* ---- ray queries: RayQueryAll / RayQueryAny (accel.traverse / traverse_any with
   on_surface_candidate / on_procedural_candidate, include/luisa/dsl/rtx/ray_query.h:11-243;
   semantics of AccelImpl::ray_query, src/rust/luisa_compute_backend_impl/src/cpu/accel.rs:582-800)
   applied to a batch with built-in candidate callbacks: hits[i] (lc_hip_committed_hit) =
   accel.traverse(rays[i], mask) (terminate_on_first = 0) or traverse_any (1). Surface candidates
   of opaque instances are committed without a callback; of other instances per surface_mode:
   COMMIT, IGNORE, or CUTOUT (commit iff with w = 1 - u - v the lengths |(w, u)|, |(u, v)|,
   |(v, w)| are all below `cutout`, src/tests/test_procedural.cpp:86-97). Procedural candidates
   per procedural_mode: IGNORE, or SPHERE (the sphere inscribed in the primitive's AABB, radius
   half its smallest extent, hit by the object-space ray, test_procedural.cpp:98-120); a committed
   distance counts inside [t_min, t_max). traverse returns the nearest committed candidate (ties:
   smaller instance, then primitive); traverse_any the first one found. HIP kernels of their own
   use the device template in ray_query.h with their own callbacks. ---- */
#define LC_HIP_QUERY_SURFACE_COMMIT 0u
#define LC_HIP_QUERY_SURFACE_IGNORE 1u
#define LC_HIP_QUERY_SURFACE_CUTOUT 2u
#define LC_HIP_QUERY_PROCEDURAL_IGNORE 0u
#define LC_HIP_QUERY_PROCEDURAL_SPHERE 1u
typedef struct lc_hip_query_policy {
    uint32_t surface_mode;
    float cutout;
    uint32_t procedural_mode;
    uint32_t terminate_on_first;
} lc_hip_query_policy;
int lc_hip_accel_query(uint64_t device, uint64_t stream, uint64_t accel, uint64_t ray_buffer, size_t ray_offset,
                       uint64_t hit_buffer, size_t hit_offset, uint32_t ray_count, uint32_t vis_mask,
                       const lc_hip_query_policy *policy);

/* ---- run-time shaders: create_shader / destroy_shader (device_interface.h:151-160) and
   ShaderDispatchCommand (command.h:99-160) for kernels given as HIP source -- the form a DSL
   codegen emits for this backend (the CUDA backend's cuda_codegen_ast.cpp:1291-1490 + NVRTC) --
   compiled by hiprtc for gfx950 against the device library lc_hip_device.h (trace_closest /
   trace_any, ray queries with the shader's own candidate callbacks, instance ops, make_ray,
   offset_ray_origin, triangle_interpolate). Arguments are encoded like the CUDA backend's
   (cuda_shader_native.cpp:73-184): in order, each aligned to 16 bytes -- buffer {pointer at the
   offset, size in bytes}, accel {descriptor pointer, 0}, uniform its bytes -- then uint4
   {dispatch x, y, z, 0}; the grid covers the dispatch size with blocks of the shader's block
   size. Compile errors return LC_HIP_INVALID_HANDLE / LC_HIP_ERROR_INVALID_ARGUMENT with the
   compiler's log in `log` (truncated to log_capacity). ---- */
#define LC_HIP_ARGUMENT_BUFFER 0u  /* Argument::Tag (include/luisa/runtime/rhi/argument.h:9-15) */
#define LC_HIP_ARGUMENT_TEXTURE 1u /* {pixel pointer, width, height, storage} (32 B) */
#define LC_HIP_ARGUMENT_UNIFORM 2u
#define LC_HIP_ARGUMENT_BINDLESS_ARRAY 3u /* {buffer slots, slot count, texture slots, 0} (32 B) */
#define LC_HIP_ARGUMENT_ACCEL 4u
typedef struct lc_hip_shader_argument {
    uint32_t tag;
    uint32_t _pad;
    uint64_t handle; /* buffer or accel */
    size_t offset;   /* buffer: byte offset; uniform: offset into the uniform bytes */
    size_t size;     /* buffer: bytes bound (0: to the end of the buffer); uniform: byte count */
} lc_hip_shader_argument;
/* Compile only (no device needed): 0 on success with the code object size in *code_size.
   Compiles go through the code-object cache (cuda_device.cpp:555-604 caches PTX the same way):
   a 64-bit hash of the compile options, the device library and the source names the entry, kept
   in memory and on disk (kernel_<hash>.hsaco under LUISA_HIP_CACHE_DIR, default
   $HOME/.cache/luisacompute_amd, "0" = no disk cache; LUISA_DUMP_SOURCE=1 writes the source
   beside it). */
int lc_hip_shader_compile(const char *source, char *log, size_t log_capacity, size_t *code_size);
/* The cache key of a source, and the process's cache hits (memory or disk) and misses (hiprtc runs). */
uint64_t lc_hip_shader_cache_key(const char *source);
int lc_hip_shader_cache_stats(uint64_t *hits, uint64_t *misses);
uint64_t lc_hip_shader_create(uint64_t device, const char *source, const char *entry, uint32_t block_x,
                              uint32_t block_y, uint32_t block_z, char *log, size_t log_capacity);
void lc_hip_shader_destroy(uint64_t device, uint64_t shader);
/* shader_argument_usage (device_interface.h:159): how the kernel uses argument `index`, Usage
   values of include/luisa/ast/usage.h (0 none, 1 read, 2 write, 3 read-write). A frontend that
   built the kernel's AST records them with set_argument_usage; unrecorded arguments answer 3. */
int lc_hip_shader_set_argument_usage(uint64_t device, uint64_t shader, const uint32_t *usage, size_t count);
int lc_hip_shader_argument_usage(uint64_t device, uint64_t shader, size_t index);
/* ShaderDispatchCommand with several dispatch sizes (3 uint32 per dispatch): launch k gets
   ls_kid = {size k, k}. */
int lc_hip_shader_dispatch_multiple(uint64_t device, uint64_t stream, uint64_t shader,
                                    const lc_hip_shader_argument *arguments, size_t argument_count,
                                    const void *uniforms, size_t uniform_size, const uint32_t *sizes, size_t count);
/* Indirect ShaderDispatchCommand: sizes from an indirect dispatch buffer written on the device,
   laid out as the CUDA backend's (cuda_device_resource.h:99-135): a 16-B header {count}, then
   32-B entries {block x, y, z, 0, dispatch x, y, z, kernel id}; entries [offset, offset +
   min(count, max_dispatch)) are launched (block sizes must be 0 or the shader's). The call waits
   for the stream to produce the sizes. */
int lc_hip_shader_dispatch_indirect(uint64_t device, uint64_t stream, uint64_t shader,
                                    const lc_hip_shader_argument *arguments, size_t argument_count,
                                    const void *uniforms, size_t uniform_size, uint64_t indirect_buffer,
                                    uint32_t offset, uint32_t max_dispatch);
int lc_hip_shader_dispatch(uint64_t device, uint64_t stream, uint64_t shader, const lc_hip_shader_argument *arguments,
                           size_t argument_count, const void *uniforms, size_t uniform_size, uint32_t dispatch_x,
                           uint32_t dispatch_y, uint32_t dispatch_z);

/* ---- images: create_texture / destroy_texture (device_interface.h:125-133) for 2-D images of one
   level, stored as linear rows of pixels, and TextureUploadCommand / TextureDownloadCommand
   (command.h:235-330) of the whole image (rows tightly packed). Storage values are the
   reference's PixelStorage (include/luisa/runtime/rhi/pixel.h:14-35): BYTE1..FLOAT4 (0..14);
   block-compressed and packed formats are not supported. Shaders read and write them with the
   reference's texel conversions (cuda_device_resource.h:225-300): unorm bytes / shorts for float
   images, half precision rounding to nearest, integer casts for int and uint images. ---- */
#define LC_HIP_PIXEL_BYTE1 0u
#define LC_HIP_PIXEL_BYTE2 1u
#define LC_HIP_PIXEL_BYTE4 2u
#define LC_HIP_PIXEL_SHORT1 3u
#define LC_HIP_PIXEL_SHORT2 4u
#define LC_HIP_PIXEL_SHORT4 5u
#define LC_HIP_PIXEL_INT1 6u
#define LC_HIP_PIXEL_INT2 7u
#define LC_HIP_PIXEL_INT4 8u
#define LC_HIP_PIXEL_HALF1 9u
#define LC_HIP_PIXEL_HALF2 10u
#define LC_HIP_PIXEL_HALF4 11u
#define LC_HIP_PIXEL_FLOAT1 12u
#define LC_HIP_PIXEL_FLOAT2 13u
#define LC_HIP_PIXEL_FLOAT4 14u
uint64_t lc_hip_texture_create(uint64_t device, uint32_t storage, uint32_t width, uint32_t height);
/* Volume of one level (Volume<T>, include/luisa/runtime/volume.h; create_texture with dimension 3,
   device_interface.h:125-128): width x height x depth pixels, x fastest, then y, then z. Upload,
   download, destroy and shader binding (LC_HIP_ARGUMENT_TEXTURE) as for 2-D images; bindless
   texture slots take 2-D images only. */
uint64_t lc_hip_volume_create(uint64_t device, uint32_t storage, uint32_t width, uint32_t height, uint32_t depth);
void lc_hip_texture_destroy(uint64_t device, uint64_t texture);
int lc_hip_texture_upload(uint64_t device, uint64_t stream, uint64_t texture, const void *data);
int lc_hip_texture_download(uint64_t device, uint64_t stream, uint64_t texture, void *data);
/* BufferToTextureCopyCommand / TextureToBufferCopyCommand (command.h:235-301) of the whole image:
   its pixels, tightly packed, at buffer_offset; to_texture != 0 copies buffer -> image. */
int lc_hip_texture_buffer_copy(uint64_t device, uint64_t stream, uint64_t texture, uint64_t buffer, size_t buffer_offset,
                               int to_texture);
/* TextureCopyCommand (command.h:303-336) of a whole image onto one of the same size and pixel size. */
int lc_hip_texture_copy(uint64_t device, uint64_t stream, uint64_t src, uint64_t dst);
/* Storage (PixelStorage) and size (width, height, depth) of an image. */
int lc_hip_texture_info(uint64_t device, uint64_t texture, uint32_t *storage, uint32_t size[3]);

/* ---- bindless arrays: create_bindless_array / destroy_bindless_array (device_interface.h) and
   BindlessArrayUpdateCommand (command.h:617-680): a buffer slot holds {pointer at the buffer
   offset, bytes from there to the buffer's end}, a 2-D texture slot an image (lc_hip_texture_*)
   with a sampler code (Sampler::code, include/luisa/runtime/rhi/sampler.h: filter << 2 |
   address; filters POINT, LINEAR_POINT, LINEAR_LINEAR, ANISOTROPIC, addresses EDGE, REPEAT,
   MIRROR, ZERO) (lc_hip_device.h dev::BindlessArray). Sampling is exact single-precision point or
   bilinear filtering on normalized coordinates (no mip levels: the linear filters are bilinear,
   anisotropic too). Operation values are the reference's Modification::Operation (NONE, EMPLACE,
   REMOVE). The update is complete when the call returns. ---- */
#define LC_HIP_BINDLESS_NONE 0u
#define LC_HIP_BINDLESS_EMPLACE 1u
#define LC_HIP_BINDLESS_REMOVE 2u
typedef struct lc_hip_bindless_modification {
    size_t slot;
    uint64_t buffer;     /* buffer EMPLACE: buffer handle */
    size_t offset_bytes; /* buffer EMPLACE: byte offset into the buffer */
    uint32_t buffer_op;
    uint32_t texture_op; /* 2-D texture slot operation */
    uint64_t texture;    /* texture EMPLACE: image handle */
    uint32_t sampler;    /* texture EMPLACE: Sampler::code */
    uint32_t _pad;
} lc_hip_bindless_modification;
uint64_t lc_hip_bindless_create(uint64_t device, size_t slots);
void lc_hip_bindless_destroy(uint64_t device, uint64_t array);
int lc_hip_bindless_update(uint64_t device, uint64_t stream, uint64_t array, const lc_hip_bindless_modification *mods,
                           size_t count);

/* ---- path tracing: the DSL program of src/tests/test_path_tracing.cpp:165-258 (camera ray,
   closest hit, emission with the balance heuristic, one area-light sample with a shadow ray
   through offset_ray_origin, cosine-weighted bounce, Russian roulette), run as a built-in
   kernel over an accel whose instance i is shaded with pt_instances[i]. The reference runs the
   same program through its AST codegen; here it is compiled into the library, in two forms:
   LC_HIP_PT_MEGAKERNEL (one launch, a lane per pixel) and LC_HIP_PT_WAVEFRONT (per-bounce
   compacted ray queues, persistent trace launches, optional queue sort and hipGraph replay).
   Both give the same image bit for bit. Hit points and normals come from the triangles the
   instance's mesh was built from (the reference program reads the same vertices through its
   bindless heap). ---- */
typedef struct lc_hip_pt_instance {
    float albedo[3];          /* Lambertian reflectance */
    uint32_t emissive;        /* 1: this instance is the light's geometry; a hit returns light_emission */
} lc_hip_pt_instance;
typedef struct lc_hip_pt_scene {
    float light_position[3];  /* the quad light: position + a u + b v, (a, b) in [0,1)^2 */
    float light_u[3];
    float light_v[3];
    float light_emission[3];  /* radiance, emitted toward normalize(cross(u, v)) */
    float camera_origin[3];
    float camera_right[3];    /* pixel (x, y) looks along normalize(fx right - fy up + forward), */
    float camera_up[3];       /* fx, fy in [-1, 1] along the shorter image side */
    float camera_forward[3];
} lc_hip_pt_scene;
enum { LC_HIP_PT_MEGAKERNEL = 0, LC_HIP_PT_WAVEFRONT = 1 };
typedef struct lc_hip_pt_options {
    uint32_t width, height;
    uint32_t spp;             /* samples per pixel of this call */
    uint32_t sample_offset;   /* index of the first sample (progressive rendering) */
    uint32_t max_depth;       /* closest-hit rays per path, 1..12 */
    uint32_t rr_depth;        /* Russian roulette from this bounce on (0 = the camera ray's hit) */
    uint64_t seed;            /* PCG32 initstate */
    uint64_t sequence;        /* pixel i draws from PCG32 stream initseq = sequence + i */
    uint32_t mode;            /* LC_HIP_PT_MEGAKERNEL or LC_HIP_PT_WAVEFRONT */
    uint32_t chunk_spp;       /* wavefront: samples per wave of paths (0 = about 16M paths) */
    uint32_t sort;            /* wavefront: bucket-sort continuation rays (origin cell, octant) */
    uint32_t graph;           /* wavefront: capture the launches in a hipGraph, replay when unchanged */
    uint32_t fetch_chunk;     /* wavefront traces: queue slots a wave claims per atomic (0: exact need) */
    uint32_t fetch_segments;  /* wavefront traces: 1, or 8 (a queue segment per XCD); 0 = default (128, 8) */
} lc_hip_pt_options;
/* image: float4 per pixel = (mean radiance of the samples, 1), row-major. counters: 0 or a buffer
   of two uint64 that receives (closest-hit rays traced, shadow rays traced). */
int lc_hip_pathtrace(uint64_t device, uint64_t stream, uint64_t accel, const lc_hip_pt_instance *pt_instances,
                     uint32_t instance_count, const lc_hip_pt_scene *scene, const lc_hip_pt_options *options,
                     uint64_t image_buffer, size_t image_offset, uint64_t counters_buffer);

/* ---- render serving (no reference counterpart: a serving loop over the path tracer above).
   Scenes (an accel + shading table + camera/light) stay resident; requests are admitted in
   arrival order while their accumulators fit the HBM budget, and the samples of all admitted
   requests of a scene are batched into shared wavefront waves (continuous batching: a request
   that arrives while others run joins the next wave). Each request's image equals a
   lc_hip_pathtrace render of it alone. The accel must not be rebuilt while the server uses it. */
enum { LC_HIP_REQUEST_QUEUED = 0, LC_HIP_REQUEST_ACTIVE = 1, LC_HIP_REQUEST_DONE = 2, LC_HIP_REQUEST_FAILED = 3 };
typedef struct lc_hip_server_options {
    uint64_t hbm_budget_bytes;  /* accumulator arena: admission budget (float4 per pixel per request) */
    uint64_t wave_paths;        /* path capacity of a scene's wave */
    uint32_t max_active;        /* admitted requests at once */
    uint32_t sort;              /* waves: bucket-sort continuation rays */
    uint32_t graph;             /* waves: replay each scene's wave from a hipGraph */
    uint32_t fetch_chunk;       /* waves' traces: see lc_hip_pt_options */
    uint32_t fetch_segments;
    uint32_t _pad;
} lc_hip_server_options;
typedef struct lc_hip_render_request {
    uint32_t scene, width, height, spp;
    uint64_t seed, sequence;    /* PCG32: pixel i draws from stream (seed, sequence + i) */
    uint64_t image_buffer;      /* float4 per pixel, written when the request completes */
    size_t image_offset;
} lc_hip_render_request;
typedef struct lc_hip_request_status {
    uint32_t state;             /* LC_HIP_REQUEST_* */
    uint32_t samples_done;
    double t_submit, t_admit, t_done; /* seconds since the server was created; -1 = not yet */
} lc_hip_request_status;
typedef struct lc_hip_serving_stats {
    uint64_t waves, rays_closest, rays_shadow, requests_done;
    uint32_t peak_active, active, queued, _pad;
    uint64_t peak_bytes;
} lc_hip_serving_stats;
uint64_t lc_hip_server_create(uint64_t device, const lc_hip_server_options *options);
void lc_hip_server_destroy(uint64_t device, uint64_t server);
int lc_hip_server_add_scene(uint64_t device, uint64_t server, uint64_t accel, const lc_hip_pt_instance *pt_instances,
                            uint32_t instance_count, const lc_hip_pt_scene *scene, uint32_t max_depth,
                            uint32_t rr_depth, uint32_t *scene_id);
int lc_hip_server_submit(uint64_t device, uint64_t server, const lc_hip_render_request *request, uint64_t *request_id);
int lc_hip_server_query(uint64_t device, uint64_t server, uint64_t request_id, lc_hip_request_status *status);
/* LC_HIP_OK when done; LC_HIP_ERROR_TIMEOUT when not finished within the timeout. */
int lc_hip_server_wait(uint64_t device, uint64_t server, uint64_t request_id, double timeout_seconds);
int lc_hip_server_drain(uint64_t device, uint64_t server, double timeout_seconds);
int lc_hip_server_stats(uint64_t device, uint64_t server, lc_hip_serving_stats *stats);

/* ---- statistics, tuning and timing (no reference counterpart: reporting for this backend) ---- */
typedef struct lc_hip_bvh_stats {
    uint32_t node_count;
    uint32_t leaf_primitive_count;
    uint32_t depth;
    uint32_t primitive_count;
    float bounds[6];
    uint64_t node_bytes;
    uint64_t primitive_bytes;
} lc_hip_bvh_stats;
int lc_hip_mesh_stats(uint64_t device, uint64_t mesh, lc_hip_bvh_stats *out);
/* Checks a built BLAS on the device. errors[k] counts violations of: 0 child index range,
   1 triangle range, 2 leaf triangle inside its quantised slot box, 3 child node inside its parent
   slot box, 4 each triangle record referenced once, 5 each primitive covered once, 6 empty slots
   empty. All zero for a well-formed tree. */
int lc_hip_mesh_validate(uint64_t device, uint64_t stream, uint64_t mesh, uint32_t errors[7]);
/* Device memory a built primitive holds (meshes and procedural primitives). With
   AccelOption::allow_compaction (include/luisa/runtime/rhi/resource.h:114) a build goes through the
   device's build arena and keeps only the used wide nodes, and no build scratch unless
   allow_update (the OptiX build compacts into a buffer that only grows,
   src/backends/cuda/cuda_mesh.cpp:42-89); without it the primitive keeps the worst-case node
   array and the build scratch. */
typedef struct lc_hip_memory_usage {
    uint64_t node_bytes;      /* wide-node array */
    uint64_t primitive_bytes; /* leaf records (procedural: + build boxes and leaf ids) */
    uint64_t scratch_bytes;   /* build scratch kept for refits */
    uint64_t total_bytes;
} lc_hip_memory_usage;
int lc_hip_mesh_memory(uint64_t device, uint64_t mesh, lc_hip_memory_usage *out);
/* The device's build arena (worst-case nodes, and the scratch of builds without allow_update) is
   kept for the next compacting build; these report its size and release it. No counterpart in
   the reference (its temporary build buffers come from a pool). */
uint64_t lc_hip_device_build_arena_bytes(uint64_t device);
int lc_hip_device_release_build_arena(uint64_t device);
int lc_hip_accel_stats(uint64_t device, uint64_t accel, lc_hip_bvh_stats *out);
/* kind 0 = closest, 1 = any. block in {64,128,256}; lds_stack in {8,16,24}; steps >= 1;
   blocks_per_cu 0 = occupancy limit. */
int lc_hip_trace_set_config(uint64_t device, int kind, int block, int lds_stack, int steps, int blocks_per_cu);
int lc_hip_trace_get_config(uint64_t device, int kind, int *block, int *lds_stack, int *steps, int *blocks_per_cu);
/* 1 = traverse rays in the order of a sort by (origin Morton code in the scene box, direction
   octant) so neighbouring lanes share subtrees; results still land in each ray's own slot.
   Applies to plain, batched, device-counted (indirect) and captured traces and to ray queries. */
int lc_hip_trace_set_ray_sort(uint64_t device, int kind, int enable);
int lc_hip_trace_get_ray_sort(uint64_t device, int kind, int *enable);
/* How a persistent trace launch hands out queue slots: a wave claims `chunk` slots per atomic
   (0: exactly what its idle lanes need, one atomic per refill round) from one queue counter
   (segments 1) or from a segment of the queue per XCD (segments 8). */
int lc_hip_trace_set_fetch(uint64_t device, int kind, uint32_t chunk, uint32_t segments);
int lc_hip_trace_get_fetch(uint64_t device, int kind, uint32_t *chunk, uint32_t *segments);
/* Traversal step policy: closest hit takes a triangle step when 4 x (lanes holding triangles) >=
   tri_quarters x (live lanes), tri_quarters in 1..4 (default 2); a wave stops stepping early to
   refill once fewer than refill_below of its 64 lanes hold a ray (0 = off, the default). */
int lc_hip_trace_set_steps(uint64_t device, int kind, uint32_t tri_quarters, uint32_t refill_below);
int lc_hip_trace_get_steps(uint64_t device, int kind, uint32_t *tri_quarters, uint32_t *refill_below);
/* Build knob: 1 = also open multi-triangle leaf candidates while a node has free slots. */
int lc_hip_build_set_split_small(uint64_t device, int enable);
/* Build knob: SAH refinement of FAST_TRACE builds (the radix tree's clusters of <= 64 primitives
   rebuilt by binned SAH). 1 = on (the default, unless the environment sets LUISA_HIP_BVH_SAH=0),
   0 = off. FAST_BUILD builds never refine. */
int lc_hip_build_set_sah(uint64_t device, int enable);
/* Build knob: children per wide BVH node, 2..8 (default 8, or LUISA_HIP_BVH_WIDTH); the node
   format keeps 8 slots, narrower trees leave slots empty. The getter returns the width. */
int lc_hip_build_set_width(uint64_t device, int width);
int lc_hip_build_get_width(uint64_t device);
/* Build knob: wide levels of the collapse launched over the whole GPU (0 = automatic:
   ceil(log_width n) + 4); deeper levels are walked by one workgroup. Same BVH either way (node
   numbering within a level may differ); a small value only moves work to that workgroup. */
int lc_hip_build_set_collapse_head(uint64_t device, int levels);
/* Kernel timing with HIP events recorded on the launch stream around each build/trace. */
int lc_hip_timing_enable(uint64_t device, int enable);
int lc_hip_timing_get(uint64_t device, const char *name, double *total_ms, uint64_t *launches);
int lc_hip_timing_reset(uint64_t device);

/* ---- synthetic workloads (PCG32, O'Neill 2014; initstate 42, initseq = workload number) ---- */
void lc_hip_workload_icosphere_counts(int level, uint32_t *vertex_count, uint32_t *triangle_count);
void lc_hip_workload_icosphere(int level, float *vertices4, uint32_t *triangles);
void lc_hip_workload_rays_orbit_to_ball(uint64_t initstate, uint64_t initseq, uint64_t ray_count, float orbit,
                                        lc_hip_ray *rays);
void lc_hip_workload_soup(uint64_t initstate, uint64_t initseq, uint64_t triangle_count, float extent, float sigma,
                          float *vertices4, uint32_t *triangles);
void lc_hip_workload_rays_box_sphere(uint64_t initstate, uint64_t initseq, uint64_t skip, uint64_t ray_count, float lo,
                                     float hi, lc_hip_ray *rays);
/* Instances: 8 draws each -- mesh id floor(u * mesh_count), uniform rotation (Shoemake 1992,
   3 draws), translation U[-extent, extent]^3, scale U[scale_lo, scale_hi]; affines row-major 3x4. */
void lc_hip_workload_instances(uint64_t initstate, uint64_t initseq, uint32_t instance_count, uint32_t mesh_count,
                               float translate_extent, float scale_lo, float scale_hi, float *affines,
                               uint32_t *mesh_ids);
/* Batch sizes log-uniform in [lo, hi] (one draw each, after `skip` draws). */
void lc_hip_workload_batch_sizes(uint64_t initstate, uint64_t initseq, uint64_t skip, uint32_t batch_count,
                                 uint32_t lo, uint32_t hi, uint32_t *sizes);

#ifdef __cplusplus
}
#endif
#endif
