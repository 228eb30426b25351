// lc_api.cpp -- luisa_compute_lib_interface: the reference's pure-C backend entry point
// (LibInterface -> DeviceInterface function table -> CommandList dispatch,
// include/luisa/rust/api_types.h:416-664; loader src/backends/common/rust_device_common.cpp:455-480)
// over this library's flat C ABI (lc_hip.h). Every handle in the table is the flat ABI's handle.
//
// Command semantics follow the reference's C++ commands (include/luisa/runtime/rhi/command.h):
// upload sources are consumed before dispatch returns, download destinations are valid once the
// stream has passed the command, and the dispatch callback runs on a backend thread after all of
// the list's work has completed (CUDAStream::callback, src/backends/cuda/cuda_stream.cpp:142-170).
#include "lc_api.h"
#include "lc_hip.h"

#include <atomic>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

std::atomic<void (*)(lcapi_logger_message)> g_logger{nullptr};

void log_message(const char *level, const std::string &msg) {
    auto cb = g_logger.load();
    if (cb) {
        cb(lcapi_logger_message{"luisa_compute_backend_hip", level, msg.c_str()});
    } else {
        std::fprintf(stderr, "[luisa_compute_backend_hip] [%s] %s\n", level, msg.c_str());
        std::fflush(stderr);
    }
}

// Fatal errors: logged at level "E", then the process aborts, as the reference's LUISA_ERROR and
// panic_abort! do (include/luisa/core/logging.h:52-63).
[[noreturn]] void fatal(const std::string &what) {
    log_message("E", what + (lc_hip_last_error()[0] ? std::string(": ") + lc_hip_last_error() : std::string()));
    std::abort();
}

void check(int rc, const char *what) {
    if (rc != LC_HIP_OK) fatal(what);
}

lcapi_created_resource invalid_resource() { return lcapi_created_resource{LCAPI_INVALID_HANDLE, nullptr}; }

lcapi_created_resource resource(uint64_t h, void *native = nullptr) {
    if (h == LC_HIP_INVALID_HANDLE) return invalid_resource();
    return lcapi_created_resource{h, native};
}

lc_hip_accel_option accel_option(const lcapi_accel_option *o) {
    lc_hip_accel_option r{};
    if (o) {
        r.hint = o->hint == LCAPI_HINT_FAST_BUILD ? LC_HIP_ACCEL_FAST_BUILD : LC_HIP_ACCEL_FAST_TRACE;
        r.allow_compaction = o->allow_compaction ? 1 : 0;
        r.allow_update = o->allow_update ? 1 : 0;
    }
    return r;
}

// ---- element sizes of the reference's IR types. `carc` points at a CArc<ir::Type>: one pointer
// to a shared block whose first member points at the Type (include/luisa/rust/ir_common.h:24-46).
// A Type is a 32-bit tag followed by its body at offset 8 (ir.hpp:187-222); the size rules are
// those of src/rust/luisa_compute_ir/src/ir.rs:207-335 (3-vectors of scalars padded to 4 lanes,
// float matrices of column vectors, struct sizes stored in the type, arrays element * length).
const char *arc_target(const void *carc) {
    if (!carc) return nullptr;
    const void *block = *static_cast<const void *const *>(carc);
    if (!block) return nullptr;
    return *static_cast<const char *const *>(block);
}

template<typename T>
T field(const char *base, size_t offset) {
    T v;
    std::memcpy(&v, base + offset, sizeof(T));
    return v;
}

size_t primitive_size(int32_t p) {
    static const size_t sizes[12] = {1, 1, 1, 2, 2, 4, 4, 8, 8, 2, 4, 8}; // Bool .. Float64 (ir.hpp:20-33)
    return p >= 0 && p < 12 ? sizes[p] : 0;
}

// VectorType at `v`: VectorElementType {tag, body at 8} (16 B), then the length (uint32 at 16).
bool vector_size(const char *v, size_t &size, int depth) {
    if (!v || depth > 8) return false;
    const int32_t etag = field<int32_t>(v, 0);
    const uint32_t length = field<uint32_t>(v, 16);
    if (etag == 0) { // scalar lanes: lengths 4k + 3 take 4k + 4 lanes
        const size_t ps = primitive_size(field<int32_t>(v, 8));
        if (ps == 0) return false;
        const uint32_t lanes = length % 4 <= 2 ? length : length / 4 * 4 + 4;
        size = ps * lanes;
        return true;
    }
    if (etag == 1) { // vector lanes: CArc<VectorType> at 8, no padding
        size_t inner = 0;
        if (!vector_size(arc_target(v + 8), inner, depth + 1)) return false;
        size = inner * length;
        return true;
    }
    return false;
}

bool type_size(const void *carc, size_t &size, bool &is_void, int depth = 0) {
    const char *t = arc_target(carc);
    if (!t || depth > 16) return false;
    is_void = false;
    switch (field<int32_t>(t, 0)) {
        case 0: // Void
        case 1: // UserData
            size = 0;
            is_void = true;
            return true;
        case 2: // Primitive
            size = primitive_size(field<int32_t>(t, 8));
            return size != 0;
        case 3: // Vector
            return vector_size(t + 8, size, depth);
        case 4: { // Matrix: float32 scalar elements only; columns of 2, 4, 4 lanes
            if (field<int32_t>(t, 8) != 0 || field<int32_t>(t, 16) != 10) return false;
            const uint32_t dim = field<uint32_t>(t, 24);
            if (dim < 2 || dim > 4) return false;
            size = 4u * (dim == 2 ? 2u : 4u) * dim;
            return true;
        }
        case 5: // Struct: fields (24 B slice) at 8, alignment at 32, size at 40
            size = field<size_t>(t, 40);
            return true;
        case 6: { // Array: CArc<Type> element at 8, length at 16
            size_t es = 0;
            bool ev = false;
            if (!type_size(t + 8, es, ev, depth + 1) || ev) return false;
            size = es * field<size_t>(t, 16);
            return true;
        }
        default: // Opaque
            return false;
    }
}

// PixelFormat (api_types.h:59-100) -> PixelStorage, as pixel_format_to_storage
// (include/luisa/runtime/rhi/pixel.h:120-180); packed and block-compressed formats: ~0u.
uint32_t format_storage(int32_t f) {
    if (f < 0) return ~0u;
    if (f < 18) return (uint32_t)f / 3u;                 // 8- and 16-bit integer formats
    if (f < 24) return LC_HIP_PIXEL_INT1 + (uint32_t)(f - 18) / 2u; // 32-bit integer formats
    if (f < 27) return LC_HIP_PIXEL_HALF1 + (uint32_t)(f - 24);
    if (f < 30) return LC_HIP_PIXEL_FLOAT1 + (uint32_t)(f - 27);
    return ~0u;
}

// ---- device table ----
void api_destroy_device(lcapi_device_interface d) { lc_hip_device_destroy(d.device.value); }

lcapi_created_buffer api_create_buffer(lcapi_handle d, const void *element, size_t count, void *ext_mem) {
    size_t es = 0;
    bool is_void = false;
    if (!type_size(element, es, is_void)) fatal("create_buffer: unsupported element type");
    const size_t stride = is_void ? 1 : es;
    if (stride == 0) fatal("create_buffer: element of size 0");
    void *native = ext_mem;
    const uint64_t h = ext_mem ? lc_hip_buffer_create_external(d.value, stride, count, ext_mem)
                               : lc_hip_buffer_create(d.value, stride, count, &native);
    if (h == LC_HIP_INVALID_HANDLE) fatal("create_buffer");
    // a void element is a byte buffer: element_stride reports the type's size (0), as the
    // reference implementation of this table does (cpu/mod.rs:51-83)
    return lcapi_created_buffer{lcapi_created_resource{h, native}, es, stride * count};
}

void api_destroy_buffer(lcapi_handle d, lcapi_handle b) { lc_hip_buffer_destroy(d.value, b.value); }

lcapi_created_resource api_create_texture(lcapi_handle d, lcapi_enum format, uint32_t dimension, uint32_t w, uint32_t h,
                                          uint32_t depth, uint32_t mips, bool, bool) {
    const uint32_t storage = format_storage(format);
    if (storage == ~0u || mips != 1 || (dimension != 2 && dimension != 3)) {
        log_message("W", "create_texture: only 2-D and 3-D images of one level in the uncompressed formats are supported");
        return invalid_resource();
    }
    return resource(dimension == 2 ? lc_hip_texture_create(d.value, storage, w, h)
                                   : lc_hip_volume_create(d.value, storage, w, h, depth));
}

void *api_native_handle(lcapi_handle d) { return reinterpret_cast<void *>((intptr_t)lc_hip_device_native(d.value)); }
uint32_t api_warp_size(lcapi_handle d) { return lc_hip_device_warp_size(d.value); }
void api_destroy_texture(lcapi_handle d, lcapi_handle t) { lc_hip_texture_destroy(d.value, t.value); }

lcapi_created_resource api_create_bindless(lcapi_handle d, size_t n) { return resource(lc_hip_bindless_create(d.value, n)); }
void api_destroy_bindless(lcapi_handle d, lcapi_handle a) { lc_hip_bindless_destroy(d.value, a.value); }

lcapi_created_resource api_create_stream(lcapi_handle d, lcapi_enum tag) {
    const uint64_t s = lc_hip_stream_create(d.value, (uint32_t)tag);
    if (s == LC_HIP_INVALID_HANDLE) return invalid_resource();
    return resource(s, lc_hip_stream_native(d.value, s));
}

void api_destroy_stream(lcapi_handle d, lcapi_handle s) { lc_hip_stream_destroy(d.value, s.value); }
void api_synchronize_stream(lcapi_handle d, lcapi_handle s) { check(lc_hip_stream_synchronize(d.value, s.value), "synchronize_stream"); }

struct CallbackCall {
    lcapi_dispatch_callback fn;
    uint8_t *user;
};

void run_callback(void *p) {
    auto *c = static_cast<CallbackCall *>(p);
    c->fn(c->user);
    delete c;
}

void dispatch_one(uint64_t d, uint64_t s, const lcapi_command &c) {
    switch (c.tag) {
        case LCAPI_CMD_BUFFER_UPLOAD: {
            const auto &u = c.u.buffer_upload;
            check(lc_hip_buffer_upload(d, s, u.buffer.value, u.offset, u.size, u.data), "BufferUploadCommand");
            break;
        }
        case LCAPI_CMD_BUFFER_DOWNLOAD: {
            const auto &u = c.u.buffer_download;
            check(lc_hip_buffer_download(d, s, u.buffer.value, u.offset, u.size, u.data), "BufferDownloadCommand");
            break;
        }
        case LCAPI_CMD_BUFFER_COPY: {
            const auto &u = c.u.buffer_copy;
            check(lc_hip_buffer_copy(d, s, u.src.value, u.src_offset, u.dst.value, u.dst_offset, u.size), "BufferCopyCommand");
            break;
        }
        case LCAPI_CMD_BUFFER_TO_TEXTURE_COPY:
        case LCAPI_CMD_TEXTURE_TO_BUFFER_COPY: {
            const bool to_tex = c.tag == LCAPI_CMD_BUFFER_TO_TEXTURE_COPY;
            const auto &u = to_tex ? c.u.buffer_to_texture_copy : c.u.texture_to_buffer_copy;
            uint32_t storage = 0, size[3] = {0, 0, 0};
            check(lc_hip_texture_info(d, u.texture.value, &storage, size), "buffer / texture copy");
            if (u.texture_level != 0 || (uint32_t)u.storage != storage || u.texture_size[0] != size[0] ||
                u.texture_size[1] != size[1] || u.texture_size[2] != size[2])
                fatal("buffer / texture copy: whole level-0 images only");
            check(lc_hip_texture_buffer_copy(d, s, u.texture.value, u.buffer.value, u.buffer_offset, to_tex ? 1 : 0),
                  to_tex ? "BufferToTextureCopyCommand" : "TextureToBufferCopyCommand");
            break;
        }
        case LCAPI_CMD_TEXTURE_UPLOAD:
        case LCAPI_CMD_TEXTURE_DOWNLOAD: {
            const bool up = c.tag == LCAPI_CMD_TEXTURE_UPLOAD;
            const auto &u = up ? c.u.texture_upload : c.u.texture_download;
            uint32_t storage = 0, size[3] = {0, 0, 0};
            check(lc_hip_texture_info(d, u.texture.value, &storage, size), "texture transfer");
            if (u.level != 0 || (uint32_t)u.storage != storage || u.size[0] != size[0] || u.size[1] != size[1] ||
                u.size[2] != size[2])
                fatal("texture upload / download: whole level-0 images only");
            check(up ? lc_hip_texture_upload(d, s, u.texture.value, u.data) : lc_hip_texture_download(d, s, u.texture.value, u.data),
                  up ? "TextureUploadCommand" : "TextureDownloadCommand");
            break;
        }
        case LCAPI_CMD_TEXTURE_COPY: {
            const auto &u = c.u.texture_copy;
            if (u.src_level != 0 || u.dst_level != 0) fatal("TextureCopyCommand: level-0 images only");
            check(lc_hip_texture_copy(d, s, u.src.value, u.dst.value), "TextureCopyCommand");
            break;
        }
        case LCAPI_CMD_SHADER_DISPATCH:
            // create_shader never returns a valid handle for an IR kernel module
            fatal("ShaderDispatchCommand: IR kernel modules are not compiled by this backend");
        case LCAPI_CMD_MESH_BUILD: {
            const auto &u = c.u.mesh_build;
            if (u.index_stride != 12) fatal("MeshBuildCommand: triangles must be three uint32 indices (stride 12)");
            check(lc_hip_mesh_build(d, s, u.mesh.value, (uint32_t)u.request, u.vertex_buffer.value, u.vertex_buffer_offset,
                                    u.vertex_buffer_size, u.vertex_stride, u.index_buffer.value, u.index_buffer_offset,
                                    u.index_buffer_size),
                  "MeshBuildCommand");
            break;
        }
        case LCAPI_CMD_CURVE_BUILD:
            fatal("CurveBuildCommand: curves are not supported");
        case LCAPI_CMD_PROCEDURAL_BUILD: {
            const auto &u = c.u.procedural_build;
            check(lc_hip_procedural_build(d, s, u.primitive.value, (uint32_t)u.request, u.aabb_buffer.value,
                                          u.aabb_buffer_offset, u.aabb_count * 24u),
                  "ProceduralPrimitiveBuildCommand");
            break;
        }
        case LCAPI_CMD_ACCEL_BUILD: {
            const auto &u = c.u.accel_build;
            std::vector<lc_hip_accel_modification> mods(u.modifications_count);
            for (size_t i = 0; i < u.modifications_count; i++) {
                const lcapi_accel_modification &m = u.modifications[i];
                lc_hip_accel_modification &o = mods[i];
                std::memset(&o, 0, sizeof(o));
                o.index = m.index;
                o.user_id = m.user_id;
                o.flags = m.flags; // same bit values (api_types.h:186-191, command.h:536-545)
                o.vis_mask = m.visibility;
                std::memcpy(o.affine, m.affine, sizeof(o.affine));
                o.primitive = m.mesh;
            }
            check(lc_hip_accel_build(d, s, u.accel.value, (uint32_t)u.request, u.instance_count, mods.data(), mods.size(),
                                     u.update_instance_buffer_only ? 1 : 0),
                  "AccelBuildCommand");
            break;
        }
        case LCAPI_CMD_BINDLESS_UPDATE: {
            const auto &u = c.u.bindless_update;
            std::vector<lc_hip_bindless_modification> mods(u.modifications_count);
            for (size_t i = 0; i < u.modifications_count; i++) {
                const lcapi_bindless_modification &m = u.modifications[i];
                if (m.tex3d.op != 0) fatal("BindlessArrayUpdateCommand: 3-D texture slots are not supported");
                lc_hip_bindless_modification &o = mods[i];
                std::memset(&o, 0, sizeof(o));
                o.slot = m.slot;
                o.buffer_op = (uint32_t)m.buffer.op;
                o.buffer = m.buffer.buffer.value;
                o.offset_bytes = m.buffer.offset;
                o.texture_op = (uint32_t)m.tex2d.op;
                o.texture = m.tex2d.texture.value;
                o.sampler = ((uint32_t)m.tex2d.sampler.filter << 2) | (uint32_t)m.tex2d.sampler.address; // Sampler::code
            }
            check(lc_hip_bindless_update(d, s, u.array.value, mods.data(), mods.size()), "BindlessArrayUpdateCommand");
            break;
        }
        default:
            fatal("dispatch: unknown command tag " + std::to_string(c.tag));
    }
}

void api_dispatch(lcapi_handle d, lcapi_handle s, lcapi_command_list list, lcapi_dispatch_callback cb, uint8_t *user) {
    for (size_t i = 0; i < list.commands_count; i++) dispatch_one(d.value, s.value, list.commands[i]);
    if (cb) {
        auto *c = new CallbackCall{cb, user};
        if (lc_hip_stream_add_callback(d.value, s.value, run_callback, c) != LC_HIP_OK) {
            delete c;
            fatal("dispatch callback");
        }
    }
}

lcapi_created_swapchain api_create_swapchain(lcapi_handle, const lcapi_swapchain_option *, lcapi_handle) {
    log_message("W", "create_swapchain: no display support");
    return lcapi_created_swapchain{invalid_resource(), 0};
}

void api_present(lcapi_handle, lcapi_handle, lcapi_handle, lcapi_handle) { fatal("present_display_in_stream: no display support"); }
void api_destroy_swapchain(lcapi_handle, lcapi_handle) {}

lcapi_created_shader api_create_shader(lcapi_handle, lcapi_handle, const lcapi_shader_option *) {
    log_message("W", "create_shader: IR kernel modules are not compiled by this backend (kernels come from HIP source "
                     "through lc_hip_shader_create)");
    return lcapi_created_shader{invalid_resource(), {0, 0, 0}};
}

void api_destroy_shader(lcapi_handle d, lcapi_handle s) {
    if (s.value != LCAPI_INVALID_HANDLE) lc_hip_shader_destroy(d.value, s.value);
}

lcapi_created_resource api_create_event(lcapi_handle d) { return resource(lc_hip_event_create(d.value)); }
void api_destroy_event(lcapi_handle d, lcapi_handle e) { lc_hip_event_destroy(d.value, e.value); }
void api_signal_event(lcapi_handle d, lcapi_handle e, lcapi_handle s, uint64_t v) {
    check(lc_hip_event_signal(d.value, e.value, s.value, v), "signal_event");
}
void api_synchronize_event(lcapi_handle d, lcapi_handle e, uint64_t v) {
    check(lc_hip_event_synchronize(d.value, e.value, v), "synchronize_event");
}
void api_wait_event(lcapi_handle d, lcapi_handle e, lcapi_handle s, uint64_t v) {
    check(lc_hip_event_wait(d.value, e.value, s.value, v), "wait_event");
}
bool api_is_event_completed(lcapi_handle d, lcapi_handle e, uint64_t v) {
    const int r = lc_hip_event_is_completed(d.value, e.value, v);
    if (r < 0) fatal("is_event_completed");
    return r != 0;
}

lcapi_created_resource api_create_mesh(lcapi_handle d, const lcapi_accel_option *o) {
    const lc_hip_accel_option opt = accel_option(o);
    return resource(lc_hip_mesh_create(d.value, &opt));
}
void api_destroy_mesh(lcapi_handle d, lcapi_handle m) { lc_hip_mesh_destroy(d.value, m.value); }

lcapi_created_resource api_create_curve(lcapi_handle, const lcapi_accel_option *) {
    log_message("W", "create_curve: curves are not supported");
    return invalid_resource();
}
void api_destroy_curve(lcapi_handle, lcapi_handle) {}

lcapi_created_resource api_create_procedural(lcapi_handle d, const lcapi_accel_option *o) {
    const lc_hip_accel_option opt = accel_option(o);
    return resource(lc_hip_procedural_create(d.value, &opt));
}
void api_destroy_procedural(lcapi_handle d, lcapi_handle p) { lc_hip_procedural_destroy(d.value, p.value); }

lcapi_created_resource api_create_accel(lcapi_handle d, const lcapi_accel_option *o) {
    const lc_hip_accel_option opt = accel_option(o);
    return resource(lc_hip_accel_create(d.value, &opt));
}
void api_destroy_accel(lcapi_handle d, lcapi_handle a) { lc_hip_accel_destroy(d.value, a.value); }

// DeviceInterface::query (device_interface.h:187): the answer is malloc'ed and released by the
// caller through LibInterface::free_string; unknown keys give the empty string.
char *api_query(lcapi_handle d, const char *key) {
    char buf[256] = {0};
    if (!key || lc_hip_device_query(d.value, key, buf, sizeof(buf)) != LC_HIP_OK) buf[0] = 0;
    return strdup(buf);
}

lcapi_pinned_memory_ext api_pinned_memory_ext(lcapi_handle) { return lcapi_pinned_memory_ext{nullptr, nullptr, nullptr}; }
lcapi_denoiser_ext api_denoiser_ext(lcapi_handle) { return lcapi_denoiser_ext{nullptr, nullptr, nullptr, nullptr, nullptr}; }

lcapi_device_interface device_table(uint64_t device) {
    lcapi_device_interface t;
    std::memset(&t, 0, sizeof(t));
    t.device = lcapi_handle{device};
    t.destroy_device = api_destroy_device;
    t.create_buffer = api_create_buffer;
    t.destroy_buffer = api_destroy_buffer;
    t.create_texture = api_create_texture;
    t.native_handle = api_native_handle;
    t.compute_warp_size = api_warp_size;
    t.destroy_texture = api_destroy_texture;
    t.create_bindless_array = api_create_bindless;
    t.destroy_bindless_array = api_destroy_bindless;
    t.create_stream = api_create_stream;
    t.destroy_stream = api_destroy_stream;
    t.synchronize_stream = api_synchronize_stream;
    t.dispatch = api_dispatch;
    t.create_swapchain = api_create_swapchain;
    t.present_display_in_stream = api_present;
    t.destroy_swapchain = api_destroy_swapchain;
    t.create_shader = api_create_shader;
    t.destroy_shader = api_destroy_shader;
    t.create_event = api_create_event;
    t.destroy_event = api_destroy_event;
    t.signal_event = api_signal_event;
    t.synchronize_event = api_synchronize_event;
    t.wait_event = api_wait_event;
    t.is_event_completed = api_is_event_completed;
    t.create_mesh = api_create_mesh;
    t.destroy_mesh = api_destroy_mesh;
    t.create_curve = api_create_curve;
    t.destroy_curve = api_destroy_curve;
    t.create_procedural_primitive = api_create_procedural;
    t.destroy_procedural_primitive = api_destroy_procedural;
    t.create_accel = api_create_accel;
    t.destroy_accel = api_destroy_accel;
    t.query = api_query;
    t.pinned_memory_ext = api_pinned_memory_ext;
    t.denoiser_ext = api_denoiser_ext;
    return t;
}

// ---- library table ----
struct ApiContext {
    std::string runtime_path;
};

void lib_set_logger(void (*cb)(lcapi_logger_message)) { g_logger.store(cb); }

lcapi_handle lib_create_context(const char *path) {
    return lcapi_handle{reinterpret_cast<uint64_t>(new ApiContext{path ? path : ""})};
}

void lib_destroy_context(lcapi_handle ctx) { delete reinterpret_cast<ApiContext *>(ctx.value); }

// The device index from a JSON config such as {"index": 1} (absent or unparsable: device 0).
int32_t config_index(const char *config) {
    if (!config) return 0;
    const char *k = std::strstr(config, "\"index\"");
    if (!k) return 0;
    k = std::strchr(k + 7, ':');
    if (!k) return 0;
    return (int32_t)std::strtol(k + 1, nullptr, 10);
}

lcapi_device_interface lib_create_device(lcapi_handle, const char *name, const char *config) {
    if (!name || std::strcmp(name, "hip") != 0) fatal(std::string("unknown device ") + (name ? name : "(null)"));
    const uint64_t d = lc_hip_device_create(config_index(config));
    if (d == LC_HIP_INVALID_HANDLE) fatal("create_device(\"hip\")");
    return device_table(d);
}

void lib_free_string(char *s) { std::free(s); }

} // namespace

extern "C" __attribute__((visibility("default"))) lcapi_lib_interface luisa_compute_lib_interface(void) {
    return lcapi_lib_interface{nullptr, lib_set_logger, lib_create_context, lib_destroy_context, lib_create_device,
                               lib_free_string};
}
