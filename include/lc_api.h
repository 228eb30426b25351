/*
 * lc_api.h -- the pure-C backend entry point of libluisa_compute_backend_hip.so.
 *
 * The reference loads out-of-tree backends through one exported C function,
 * `luisa_compute_lib_interface`, which returns a LibInterface whose create_device hands back a
 * DeviceInterface function table; commands arrive as a C CommandList
 * (include/luisa/rust/api_types.h:416-485 for the commands, :597-664 for the two tables; the
 * loader is RustDevice in src/backends/common/rust_device_common.cpp:455-480, the export of the
 * reference CPU implementation src/rust/luisa_compute_backend_impl/src/lib.rs:233-245).
 *
 * The structures below are this library's own declarations of those C structures. They have the
 * same field order, types and therefore the same layout as the reference's (a test compiles both
 * headers together and compares every size and offset), so the reference loader calls this table
 * unchanged. Resource handles are the handles of the flat C ABI in lc_hip.h (one object, two
 * entry surfaces): an accel built through the table can be traced with lc_hip_accel_trace_*.
 *
 * Differences from the reference CPU implementation, by design:
 *   - create_device accepts the device name "hip"; the config string may hold {"index": N}.
 *   - create_shader takes the reference's IR KernelModule, which this backend does not compile
 *     (its kernels come from HIP source, lc_hip_shader_create): it returns the invalid handle,
 *     the reference's answer for an unsupported resource (include/luisa/runtime/rhi/resource.h).
 *   - curves, swapchains, mip levels above 0 and 3-D bindless slots are unsupported (invalid
 *     handles; a command naming them is an error).
 *   - errors are reported through the logger callback (level "E") and then abort the process,
 *     as the reference's panic_abort / LUISA_ERROR do.
 */
#ifndef LC_API_H
#define LC_API_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LCAPI_INVALID_HANDLE UINT64_MAX

/* Resource handles: one 64-bit word wrapped in a struct (LCBuffer, LCStream, ... in the
   reference, api_types.h:152-182,230-316). */
typedef struct lcapi_handle {
    uint64_t value;
} lcapi_handle;

/* enums travel as 32-bit ints (C enums of the reference header) */
typedef int32_t lcapi_enum;

/* AccelBuildRequest (api_types.h:12-15), AccelUsageHint (:17-20) */
#define LCAPI_BUILD_PREFER_UPDATE 0
#define LCAPI_BUILD_FORCE_BUILD 1
#define LCAPI_HINT_FAST_TRACE 0
#define LCAPI_HINT_FAST_BUILD 1

/* AccelBuildModification (api_types.h:193-200): flags as lc_hip.h's LC_HIP_MOD_* bits */
typedef struct lcapi_accel_modification {
    uint32_t index;
    uint32_t user_id;
    uint32_t flags;
    uint32_t visibility;
    uint64_t mesh;
    float affine[12];
} lcapi_accel_modification;

typedef struct lcapi_sampler {
    lcapi_enum filter;  /* POINT, LINEAR_POINT, LINEAR_LINEAR, ANISOTROPIC */
    lcapi_enum address; /* EDGE, REPEAT, MIRROR, ZERO */
} lcapi_sampler;

/* ---- commands (api_types.h:207-485) ---- */
typedef struct lcapi_buffer_upload {
    lcapi_handle buffer;
    size_t offset;
    size_t size;
    const uint8_t *data;
} lcapi_buffer_upload;

typedef struct lcapi_buffer_download {
    lcapi_handle buffer;
    size_t offset;
    size_t size;
    uint8_t *data;
} lcapi_buffer_download;

typedef struct lcapi_buffer_copy {
    lcapi_handle src;
    size_t src_offset;
    lcapi_handle dst;
    size_t dst_offset;
    size_t size;
} lcapi_buffer_copy;

/* BufferToTextureCopyCommand and TextureToBufferCopyCommand share this layout */
typedef struct lcapi_buffer_texture_copy {
    lcapi_handle buffer;
    size_t buffer_offset;
    lcapi_handle texture;
    lcapi_enum storage; /* PixelStorage */
    uint32_t texture_level;
    uint32_t texture_size[3];
} lcapi_buffer_texture_copy;

/* TextureUploadCommand and TextureDownloadCommand share this layout */
typedef struct lcapi_texture_transfer {
    lcapi_handle texture;
    lcapi_enum storage;
    uint32_t level;
    uint32_t size[3];
    uint8_t *data; /* const for uploads */
} lcapi_texture_transfer;

typedef struct lcapi_texture_copy {
    lcapi_enum storage;
    lcapi_handle src;
    lcapi_handle dst;
    uint32_t size[3];
    uint32_t src_level;
    uint32_t dst_level;
} lcapi_texture_copy;

/* Argument (api_types.h:296-316): tag 0 buffer, 1 texture, 2 uniform, 3 bindless array, 4 accel */
typedef struct lcapi_argument {
    lcapi_enum tag;
    union {
        struct {
            lcapi_handle buffer;
            size_t offset;
            size_t size;
        } buffer;
        struct {
            lcapi_handle texture;
            uint32_t level;
        } texture;
        struct {
            const uint8_t *data;
            size_t size;
        } uniform;
        lcapi_handle bindless_array;
        lcapi_handle accel;
    } u;
} lcapi_argument;

typedef struct lcapi_shader_dispatch {
    lcapi_handle shader;
    uint32_t dispatch_size[3];
    const lcapi_argument *args;
    size_t args_count;
} lcapi_shader_dispatch;

typedef struct lcapi_mesh_build {
    lcapi_handle mesh;
    lcapi_enum request;
    lcapi_handle vertex_buffer;
    size_t vertex_buffer_offset;
    size_t vertex_buffer_size;
    size_t vertex_stride;
    lcapi_handle index_buffer;
    size_t index_buffer_offset;
    size_t index_buffer_size;
    size_t index_stride;
} lcapi_mesh_build;

typedef struct lcapi_curve_build {
    lcapi_handle curve;
    lcapi_enum request;
    lcapi_enum basis;
    size_t cp_count;
    size_t seg_count;
    lcapi_handle cp_buffer;
    size_t cp_buffer_offset;
    size_t cp_buffer_stride;
    lcapi_handle seg_buffer;
    size_t seg_buffer_offset;
} lcapi_curve_build;

typedef struct lcapi_procedural_build {
    lcapi_handle primitive;
    lcapi_enum request;
    lcapi_handle aabb_buffer;
    size_t aabb_buffer_offset;
    size_t aabb_count;
} lcapi_procedural_build;

typedef struct lcapi_accel_build {
    lcapi_handle accel;
    lcapi_enum request;
    uint32_t instance_count;
    const lcapi_accel_modification *modifications;
    size_t modifications_count;
    bool update_instance_buffer_only;
} lcapi_accel_build;

/* BindlessArrayUpdateModification (api_types.h:401-418): op 0 NONE, 1 EMPLACE, 2 REMOVE */
typedef struct lcapi_bindless_buffer_slot {
    lcapi_enum op;
    lcapi_handle buffer;
    size_t offset;
} lcapi_bindless_buffer_slot;

typedef struct lcapi_bindless_texture_slot {
    lcapi_enum op;
    lcapi_handle texture;
    lcapi_sampler sampler;
} lcapi_bindless_texture_slot;

typedef struct lcapi_bindless_modification {
    size_t slot;
    lcapi_bindless_buffer_slot buffer;
    lcapi_bindless_texture_slot tex2d;
    lcapi_bindless_texture_slot tex3d;
} lcapi_bindless_modification;

typedef struct lcapi_bindless_update {
    lcapi_handle array;
    const lcapi_bindless_modification *modifications;
    size_t modifications_count;
} lcapi_bindless_update;

/* Command tags in the reference's order (api_types.h:421-436) */
enum {
    LCAPI_CMD_BUFFER_UPLOAD = 0,
    LCAPI_CMD_BUFFER_DOWNLOAD,
    LCAPI_CMD_BUFFER_COPY,
    LCAPI_CMD_BUFFER_TO_TEXTURE_COPY,
    LCAPI_CMD_TEXTURE_TO_BUFFER_COPY,
    LCAPI_CMD_TEXTURE_UPLOAD,
    LCAPI_CMD_TEXTURE_DOWNLOAD,
    LCAPI_CMD_TEXTURE_COPY,
    LCAPI_CMD_SHADER_DISPATCH,
    LCAPI_CMD_MESH_BUILD,
    LCAPI_CMD_CURVE_BUILD,
    LCAPI_CMD_PROCEDURAL_BUILD,
    LCAPI_CMD_ACCEL_BUILD,
    LCAPI_CMD_BINDLESS_UPDATE
};

typedef struct lcapi_command {
    lcapi_enum tag;
    union {
        lcapi_buffer_upload buffer_upload;
        lcapi_buffer_download buffer_download;
        lcapi_buffer_copy buffer_copy;
        lcapi_buffer_texture_copy buffer_to_texture_copy;
        lcapi_buffer_texture_copy texture_to_buffer_copy;
        lcapi_texture_transfer texture_upload;
        lcapi_texture_transfer texture_download;
        lcapi_texture_copy texture_copy;
        lcapi_shader_dispatch shader_dispatch;
        lcapi_mesh_build mesh_build;
        lcapi_curve_build curve_build;
        lcapi_procedural_build procedural_build;
        lcapi_accel_build accel_build;
        lcapi_bindless_update bindless_update;
    } u;
} lcapi_command;

typedef struct lcapi_command_list {
    const lcapi_command *commands;
    size_t commands_count;
} lcapi_command_list;

/* ---- creation results and options (api_types.h:492-560) ---- */
typedef struct lcapi_created_resource {
    uint64_t handle;
    void *native_handle;
} lcapi_created_resource;

typedef struct lcapi_created_buffer {
    lcapi_created_resource resource;
    size_t element_stride;
    size_t total_size_bytes;
} lcapi_created_buffer;

typedef struct lcapi_created_shader {
    lcapi_created_resource resource;
    uint32_t block_size[3];
} lcapi_created_shader;

typedef struct lcapi_created_swapchain {
    lcapi_created_resource resource;
    lcapi_enum storage;
} lcapi_created_swapchain;

typedef struct lcapi_accel_option {
    lcapi_enum hint;
    bool allow_compaction;
    bool allow_update;
} lcapi_accel_option;

typedef struct lcapi_shader_option {
    bool enable_cache;
    bool enable_fast_math;
    bool enable_debug_info;
    bool compile_only;
    bool time_trace;
    uint32_t max_registers;
    const char *name;
    const char *native_include;
} lcapi_shader_option;

typedef void (*lcapi_dispatch_callback)(uint8_t *);

typedef struct lcapi_swapchain_option {
    uint64_t display;
    uint64_t window;
    uint32_t width;
    uint32_t height;
    bool wants_hdr;
    bool wants_vsync;
    uint32_t back_buffer_count;
} lcapi_swapchain_option;

typedef struct lcapi_pinned_memory_option {
    bool write_combined;
} lcapi_pinned_memory_option;

typedef struct lcapi_pinned_memory_ext {
    void *data;
    void (*pin_host_memory)(struct lcapi_pinned_memory_ext *, const void *, size_t, void *,
                            const lcapi_pinned_memory_option *);
    void (*allocate_pinned_memory)(struct lcapi_pinned_memory_ext *, size_t, void *);
} lcapi_pinned_memory_ext;

/* DenoiserExt (api_types.h:586-595): not provided (all members null). */
typedef struct lcapi_denoiser_ext {
    void *data;
    void *(*create)(const struct lcapi_denoiser_ext *, uint64_t stream);
    void (*init)(const struct lcapi_denoiser_ext *, void *, const void *);
    void (*execute)(const struct lcapi_denoiser_ext *, void *, bool);
    void (*destroy)(const struct lcapi_denoiser_ext *, void *);
} lcapi_denoiser_ext;

/* ---- the device table (LCDeviceInterface, api_types.h:597-654), members in its order.
   `device` is the flat ABI's device handle. create_buffer's `element` is the reference's
   `const CArc<ir::Type> *` (include/luisa/rust/ir_common.h:24-46, ir.hpp:138-222): the element
   size follows the IR's size rules (src/rust/luisa_compute_ir/src/ir.rs:207-335); a void element
   makes a buffer of `count` bytes. ---- */
typedef struct lcapi_device_interface {
    lcapi_handle device;
    void (*destroy_device)(struct lcapi_device_interface);
    lcapi_created_buffer (*create_buffer)(lcapi_handle, const void *element, size_t count, void *ext_mem);
    void (*destroy_buffer)(lcapi_handle, lcapi_handle);
    lcapi_created_resource (*create_texture)(lcapi_handle, lcapi_enum format, uint32_t dimension, uint32_t width,
                                             uint32_t height, uint32_t depth, uint32_t mipmap_levels,
                                             bool simultaneous_access, bool allow_raster_target);
    void *(*native_handle)(lcapi_handle);
    uint32_t (*compute_warp_size)(lcapi_handle);
    void (*destroy_texture)(lcapi_handle, lcapi_handle);
    lcapi_created_resource (*create_bindless_array)(lcapi_handle, size_t);
    void (*destroy_bindless_array)(lcapi_handle, lcapi_handle);
    lcapi_created_resource (*create_stream)(lcapi_handle, lcapi_enum tag);
    void (*destroy_stream)(lcapi_handle, lcapi_handle);
    void (*synchronize_stream)(lcapi_handle, lcapi_handle);
    void (*dispatch)(lcapi_handle, lcapi_handle stream, lcapi_command_list, lcapi_dispatch_callback, uint8_t *);
    lcapi_created_swapchain (*create_swapchain)(lcapi_handle, const lcapi_swapchain_option *, lcapi_handle stream);
    void (*present_display_in_stream)(lcapi_handle, lcapi_handle stream, lcapi_handle swapchain, lcapi_handle image);
    void (*destroy_swapchain)(lcapi_handle, lcapi_handle);
    lcapi_created_shader (*create_shader)(lcapi_handle, lcapi_handle kernel_module, const lcapi_shader_option *);
    void (*destroy_shader)(lcapi_handle, lcapi_handle);
    lcapi_created_resource (*create_event)(lcapi_handle);
    void (*destroy_event)(lcapi_handle, lcapi_handle);
    void (*signal_event)(lcapi_handle, lcapi_handle event, lcapi_handle stream, uint64_t value);
    void (*synchronize_event)(lcapi_handle, lcapi_handle event, uint64_t value);
    void (*wait_event)(lcapi_handle, lcapi_handle event, lcapi_handle stream, uint64_t value);
    bool (*is_event_completed)(lcapi_handle, lcapi_handle event, uint64_t value);
    lcapi_created_resource (*create_mesh)(lcapi_handle, const lcapi_accel_option *);
    void (*destroy_mesh)(lcapi_handle, lcapi_handle);
    lcapi_created_resource (*create_curve)(lcapi_handle, const lcapi_accel_option *);
    void (*destroy_curve)(lcapi_handle, lcapi_handle);
    lcapi_created_resource (*create_procedural_primitive)(lcapi_handle, const lcapi_accel_option *);
    void (*destroy_procedural_primitive)(lcapi_handle, lcapi_handle);
    lcapi_created_resource (*create_accel)(lcapi_handle, const lcapi_accel_option *);
    void (*destroy_accel)(lcapi_handle, lcapi_handle);
    char *(*query)(lcapi_handle, const char *key);
    lcapi_pinned_memory_ext (*pinned_memory_ext)(lcapi_handle);
    lcapi_denoiser_ext (*denoiser_ext)(lcapi_handle);
} lcapi_device_interface;

typedef struct lcapi_logger_message {
    const char *target;
    const char *level; /* "E", "W", "I", "V" */
    const char *message;
} lcapi_logger_message;

/* LibInterface (api_types.h:657-664) */
typedef struct lcapi_lib_interface {
    void *inner;
    void (*set_logger_callback)(void (*)(lcapi_logger_message));
    lcapi_handle (*create_context)(const char *runtime_path);
    void (*destroy_context)(lcapi_handle);
    lcapi_device_interface (*create_device)(lcapi_handle context, const char *device_name, const char *config_json);
    void (*free_string)(char *);
} lcapi_lib_interface;

/* The export the reference's loader looks up by name (rust_device_common.cpp:466). */
lcapi_lib_interface luisa_compute_lib_interface(void);

#ifdef __cplusplus
}
#endif

#endif /* LC_API_H */
