// lc_hip_device.h -- device library of shaders compiled at run time (hiprtc, gfx950): what the
// kernels a LuisaCompute codegen emits call for the ray-tracing CallOps, and the layouts of their
// arguments (the CUDA backend's counterparts: cuda_device_resource.h:1279-1510,1776-2241).
//
// A shader is HIP source with one entry `extern "C" __global__ void <name>(const Params p)` whose
// Params struct lists the dispatch arguments in order, each aligned to 16 bytes, followed by
// `alignas(16) uint4 ls_kid` = {dispatch size x, y, z, 0} (the argument encoding of
// src/backends/cuda/cuda_shader_native.cpp:73-184). Buffers arrive as lchip::dev::Buffer<T>
// {pointer at the bound offset, bound size in bytes}, accels as lchip::dev::Accel, uniforms as
// their bytes.
//
// Semantics: trace_closest / trace_any are AccelImpl::trace_closest / trace_any
// (src/rust/luisa_compute_backend_impl/src/cpu/accel.rs:449-535); query() is accel.traverse /
// traverse_any with the shader's own candidate callbacks (ray_query.h); the instance ops read and
// write the accel's device instance table (reads see the latest writes, writes reach traversal
// at the accel's next build, cpu/accel.rs:537-579). The walks keep their stack in private memory
// (kStack entries; the host checks the accel's bound at dispatch).
#pragma once

#include "lc_hip_bvh.h"
#include "lc_hip_rtx.h"
#include "ray_query.h"
#include "traverse.h"

namespace lchip {

namespace dev {

constexpr int kStack = 64;

// array<T, N> of the AST (local and shared arrays; type.h's ARRAY)
template<class T, uint32_t N>
struct Array {
    T v[N];
    __device__ __forceinline__ T &operator[](uint32_t i) { return v[i]; }
    __device__ __forceinline__ const T &operator[](uint32_t i) const { return v[i]; }
};

struct Ray { // lc_hip_ray / Ray (include/luisa/runtime/rtx/ray.h:10-16)
    float origin[3];
    float t_min;
    float direction[3];
    float t_max;
};

struct SurfaceHit { // SurfaceHit (hit.h:30-37); miss: inst = prim = ~0u, bary (0, 0), t = t_max
    uint32_t inst, prim;
    float bary[2];
    float committed_ray_t;
    uint32_t _pad;
};

template<class T>
struct alignas(16) Buffer {
    T *ptr;
    uint64_t size_bytes;
    __device__ __forceinline__ T &operator[](uint64_t i) const { return ptr[i]; }
    __device__ __forceinline__ uint64_t size() const { return size_bytes / sizeof(T); }
};

struct alignas(16) Accel {
    const DeviceAccel *desc;
    uint64_t _pad;
};

// 2-D image (Image<T>, include/luisa/runtime/image.h): linear rows of pixels of a PixelStorage
// (BYTE1..FLOAT4 = 0..14; channel type = storage / 3, channel count 1 / 2 / 4 = storage % 3).
// Reads and writes convert like the reference's surfaces (cuda_device_resource.h:225-300): float
// images see bytes / shorts as unorm (x / 255, round(saturate(x) * 255) on write) and halves
// rounded to nearest; int / uint images cast. Channels a storage lacks read as 0. Coordinates are
// not checked.
struct alignas(16) Image {
    void *data;
    uint32_t width, height;
    uint32_t storage;
    uint32_t depth; // 1 for 2-D images
    uint32_t _pad[2];
};
// Volume (Volume<T>, include/luisa/runtime/volume.h): the same layout with depth slices of
// height rows; pixel (x, y, z) sits at (z * height + y) * width + x.
using Volume = Image;

__device__ __forceinline__ uint32_t image_channels(const Image &img) { return img.storage % 3u == 2u ? 4u : img.storage % 3u + 1u; }

// raw channel c of pixel (x, y) as float (unorm / half / float) or as a 32-bit integer pattern
__device__ __forceinline__ float image_channel_float(const Image &img, uint64_t p, uint32_t c) {
    const uint32_t kind = img.storage / 3u, n = image_channels(img);
    if (kind == 0u) return (float)static_cast<const unsigned char *>(img.data)[p * n + c] * (1.0f / 255.0f);
    if (kind == 1u) return (float)static_cast<const unsigned short *>(img.data)[p * n + c] * (1.0f / 65535.0f);
    if (kind == 2u) return (float)static_cast<const int *>(img.data)[p * n + c];
    if (kind == 3u) return (float)static_cast<const _Float16 *>(img.data)[p * n + c];
    return static_cast<const float *>(img.data)[p * n + c];
}
__device__ __forceinline__ int image_channel_int(const Image &img, uint64_t p, uint32_t c) {
    const uint32_t kind = img.storage / 3u, n = image_channels(img);
    if (kind == 0u) return (int)static_cast<const unsigned char *>(img.data)[p * n + c];
    if (kind == 1u) return (int)static_cast<const unsigned short *>(img.data)[p * n + c];
    return static_cast<const int *>(img.data)[p * n + c];
}

__device__ __forceinline__ float4 pixel_read_float(const Image img, uint64_t p) {
    const uint32_t n = image_channels(img);
    float v[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    for (uint32_t c = 0; c < n; c++) v[c] = image_channel_float(img, p, c);
    return make_float4(v[0], v[1], v[2], v[3]);
}
__device__ __forceinline__ uint4 pixel_read_uint(const Image img, uint64_t p) {
    const uint32_t n = image_channels(img);
    uint32_t v[4] = {0u, 0u, 0u, 0u};
    for (uint32_t c = 0; c < n; c++) v[c] = (uint32_t)image_channel_int(img, p, c);
    return make_uint4(v[0], v[1], v[2], v[3]);
}
__device__ __forceinline__ int4 pixel_read_int(const Image img, uint64_t p) {
    const uint4 u = pixel_read_uint(img, p);
    return make_int4((int)u.x, (int)u.y, (int)u.z, (int)u.w);
}

__device__ __forceinline__ void pixel_write_float(const Image img, uint64_t p, float4 value) {
    const uint32_t kind = img.storage / 3u, n = image_channels(img);
    const float v[4] = {value.x, value.y, value.z, value.w};
    for (uint32_t c = 0; c < n; c++) {
        const float s = fminf(fmaxf(v[c], 0.0f), 1.0f); // saturate (NaN -> 0)
        if (kind == 0u) static_cast<unsigned char *>(img.data)[p * n + c] = (unsigned char)roundf(s * 255.0f);
        else if (kind == 1u) static_cast<unsigned short *>(img.data)[p * n + c] = (unsigned short)roundf(s * 65535.0f);
        else if (kind == 2u) static_cast<int *>(img.data)[p * n + c] = (int)v[c];
        else if (kind == 3u) static_cast<_Float16 *>(img.data)[p * n + c] = (_Float16)v[c];
        else static_cast<float *>(img.data)[p * n + c] = v[c];
    }
}
__device__ __forceinline__ void pixel_write_uint(const Image img, uint64_t p, uint4 value) {
    const uint32_t kind = img.storage / 3u, n = image_channels(img);
    const uint32_t v[4] = {value.x, value.y, value.z, value.w};
    for (uint32_t c = 0; c < n; c++) {
        if (kind == 0u) static_cast<unsigned char *>(img.data)[p * n + c] = (unsigned char)v[c];
        else if (kind == 1u) static_cast<unsigned short *>(img.data)[p * n + c] = (unsigned short)v[c];
        else static_cast<uint32_t *>(img.data)[p * n + c] = v[c];
    }
}
__device__ __forceinline__ void pixel_write_int(const Image img, uint64_t p, int4 value) {
    pixel_write_uint(img, p, make_uint4((uint32_t)value.x, (uint32_t)value.y, (uint32_t)value.z, (uint32_t)value.w));
}

__device__ __forceinline__ uint64_t image_pixel(const Image &img, uint32_t x, uint32_t y) { return (uint64_t)y * img.width + x; }
__device__ __forceinline__ uint64_t volume_pixel(const Volume &v, uint32_t x, uint32_t y, uint32_t z) {
    return ((uint64_t)z * v.height + y) * v.width + x;
}
__device__ __forceinline__ float4 image_read_float(const Image img, uint32_t x, uint32_t y) { return pixel_read_float(img, image_pixel(img, x, y)); }
__device__ __forceinline__ uint4 image_read_uint(const Image img, uint32_t x, uint32_t y) { return pixel_read_uint(img, image_pixel(img, x, y)); }
__device__ __forceinline__ int4 image_read_int(const Image img, uint32_t x, uint32_t y) { return pixel_read_int(img, image_pixel(img, x, y)); }
__device__ __forceinline__ void image_write_float(const Image img, uint32_t x, uint32_t y, float4 value) { pixel_write_float(img, image_pixel(img, x, y), value); }
__device__ __forceinline__ void image_write_uint(const Image img, uint32_t x, uint32_t y, uint4 value) { pixel_write_uint(img, image_pixel(img, x, y), value); }
__device__ __forceinline__ void image_write_int(const Image img, uint32_t x, uint32_t y, int4 value) { pixel_write_int(img, image_pixel(img, x, y), value); }
__device__ __forceinline__ float4 volume_read_float(const Volume v, uint32_t x, uint32_t y, uint32_t z) { return pixel_read_float(v, volume_pixel(v, x, y, z)); }
__device__ __forceinline__ uint4 volume_read_uint(const Volume v, uint32_t x, uint32_t y, uint32_t z) { return pixel_read_uint(v, volume_pixel(v, x, y, z)); }
__device__ __forceinline__ int4 volume_read_int(const Volume v, uint32_t x, uint32_t y, uint32_t z) { return pixel_read_int(v, volume_pixel(v, x, y, z)); }
__device__ __forceinline__ void volume_write_float(const Volume v, uint32_t x, uint32_t y, uint32_t z, float4 value) { pixel_write_float(v, volume_pixel(v, x, y, z), value); }
__device__ __forceinline__ void volume_write_uint(const Volume v, uint32_t x, uint32_t y, uint32_t z, uint4 value) { pixel_write_uint(v, volume_pixel(v, x, y, z), value); }
__device__ __forceinline__ void volume_write_int(const Volume v, uint32_t x, uint32_t y, uint32_t z, int4 value) { pixel_write_int(v, volume_pixel(v, x, y, z), value); }

// Bindless array of buffer slots (BindlessArray, include/luisa/runtime/bindless_array.h; the
// CUDA backend's buffer reads cuda_device_resource.h): slot s = {pointer at the emplaced offset,
// bytes to the buffer's end}; reads and writes are unchecked, like the reference's.
struct alignas(16) BindlessArray {
    const BindlessSlot *slots;
    uint64_t count;
    const BindlessTex2D *tex;
    uint64_t _pad;
};
template<class T>
__device__ __forceinline__ T bindless_buffer_read(const BindlessArray a, uint32_t slot, uint64_t i) {
    return static_cast<const T *>(a.slots[slot].ptr)[i];
}
template<class T>
__device__ __forceinline__ void bindless_buffer_write(const BindlessArray a, uint32_t slot, uint64_t i, const T &v) {
    static_cast<T *>(a.slots[slot].ptr)[i] = v;
}
template<class T>
__device__ __forceinline__ uint64_t bindless_buffer_size(const BindlessArray a, uint32_t slot) {
    return a.slots[slot].size_bytes / sizeof(T);
}

// Indirect dispatch buffer (IndirectDispatchBuffer; the CUDA backend's LCIndirectBuffer layout,
// cuda_device_resource.h:99-135): a 16-B header {count}, then 32-B entries {block size, 0,
// dispatch size, kernel id}, filled by a shader and launched by lc_hip_shader_dispatch_indirect.
__device__ __forceinline__ void indirect_set_dispatch_count(const Buffer<uint32_t> b, uint32_t count) {
    b.ptr[0] = count;
}
__device__ __forceinline__ void indirect_set_dispatch_kernel(const Buffer<uint32_t> b, uint32_t index, uint3 block,
                                                             uint3 size, uint32_t kernel_id) {
    uint32_t *e = b.ptr + 4 + 8 * (uint64_t)index;
    e[0] = block.x; e[1] = block.y; e[2] = block.z; e[3] = 0u;
    e[4] = size.x; e[5] = size.y; e[6] = size.z; e[7] = kernel_id;
}
// kernel id of a launch of a multiple or indirect dispatch (ls_kid.w)
__device__ __forceinline__ uint32_t kernel_id(uint4 ls_kid) { return ls_kid.w; }

// 2-D texture slots: reads by texel, and sampling on normalized coordinates with the slot's sampler
// (Sampler::code = filter << 2 | address, include/luisa/runtime/rhi/sampler.h): POINT takes the
// texel floor(uv * size); the linear filters (no mip levels: LINEAR_POINT, LINEAR_LINEAR and
// ANISOTROPIC alike) blend the four texels around uv * size - 0.5 with exact single-precision
// weights, rows first. Addresses: EDGE clamps, REPEAT wraps, MIRROR reflects every other period,
// ZERO reads 0 outside. Float images only (the conversions of image_read_float).
__device__ __forceinline__ int tex_address(int i, int n, uint32_t address) {
    if (address == 0u) return i < 0 ? 0 : (i >= n ? n - 1 : i);
    if (address == 1u) { const int m = i % n; return m < 0 ? m + n : m; }
    if (address == 2u) {
        int m = i % (2 * n);
        if (m < 0) m += 2 * n;
        return m < n ? m : 2 * n - 1 - m;
    }
    return (i < 0 || i >= n) ? -1 : i;
}
__device__ __forceinline__ float4 tex_texel(const BindlessTex2D &t, int x, int y) {
    if (x < 0 || y < 0) return make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    const Image img{t.data, t.width, t.height, t.storage, 1u, {0u, 0u}};
    return image_read_float(img, (uint32_t)x, (uint32_t)y);
}
__device__ __forceinline__ float4 bindless_tex2d_read(const BindlessArray a, uint32_t slot, uint32_t x, uint32_t y) {
    return tex_texel(a.tex[slot], (int)x, (int)y);
}
__device__ __forceinline__ uint2 bindless_tex2d_size(const BindlessArray a, uint32_t slot) {
    return make_uint2(a.tex[slot].width, a.tex[slot].height);
}
__device__ __forceinline__ float4 bindless_tex2d_sample(const BindlessArray a, uint32_t slot, float u, float v) {
    const BindlessTex2D t = a.tex[slot];
    const int w = (int)t.width, h = (int)t.height;
    const uint32_t filter = t.sampler >> 2, address = t.sampler & 3u;
    if (filter == 0u) {
        const int x = tex_address((int)floorf(u * (float)w), w, address);
        const int y = tex_address((int)floorf(v * (float)h), h, address);
        return tex_texel(t, y < 0 ? -1 : x, y);
    }
    const float fx = u * (float)w - 0.5f, fy = v * (float)h - 0.5f;
    const float x0 = floorf(fx), y0 = floorf(fy);
    const float ax = fx - x0, ay = fy - y0;
    const int i0 = (int)x0, j0 = (int)y0;
    const int xa = tex_address(i0, w, address), xb = tex_address(i0 + 1, w, address);
    const int ya = tex_address(j0, h, address), yb = tex_address(j0 + 1, h, address);
    const float4 t00 = tex_texel(t, ya < 0 ? -1 : xa, ya), t10 = tex_texel(t, ya < 0 ? -1 : xb, ya);
    const float4 t01 = tex_texel(t, yb < 0 ? -1 : xa, yb), t11 = tex_texel(t, yb < 0 ? -1 : xb, yb);
    const float bx = 1.0f - ax, by = 1.0f - ay;
    const float c00[4] = {t00.x, t00.y, t00.z, t00.w}, c10[4] = {t10.x, t10.y, t10.z, t10.w};
    const float c01[4] = {t01.x, t01.y, t01.z, t01.w}, c11[4] = {t11.x, t11.y, t11.z, t11.w};
    float r[4];
    for (int c = 0; c < 4; c++) {
        const float top = c00[c] * bx + c10[c] * ax;
        const float bot = c01[c] * bx + c11[c] * ax;
        r[c] = top * by + bot * ay;
    }
    return make_float4(r[0], r[1], r[2], r[3]);
}

// dispatch_id / dispatch_size of a 1-3D dispatch (ls_kid.xyz = dispatch size)
__device__ __forceinline__ uint3 dispatch_id() {
    return make_uint3(blockIdx.x * blockDim.x + threadIdx.x, blockIdx.y * blockDim.y + threadIdx.y,
                      blockIdx.z * blockDim.z + threadIdx.z);
}
__device__ __forceinline__ bool in_dispatch(uint4 ls_kid) {
    const uint3 i = dispatch_id();
    return i.x < ls_kid.x && i.y < ls_kid.y && i.z < ls_kid.z;
}

// make_ray (src/dsl/rtx/ray.cpp:7-14): t_max defaults to FLT_MAX
__device__ __forceinline__ Ray make_ray(rtx::V3 o, rtx::V3 d, float t_min = 0.0f, float t_max = FLT_MAX) {
    return Ray{{o.x, o.y, o.z}, t_min, {d.x, d.y, d.z}, t_max};
}

template<class S, class P>
struct CallbackPolicy {
    S on_surface;
    P on_procedural;
    __device__ __forceinline__ bool surface(uint32_t inst, uint32_t prim, float u, float v, float t) const {
        return on_surface(inst, prim, u, v, t);
    }
    __device__ __forceinline__ bool procedural(uint32_t inst, uint32_t prim, const float *o, const float *d,
                                               const float *wo, const float *wd, float4 lo, float4 hi, float tmax,
                                               float &t) const {
        return on_procedural(inst, prim, o, d, wo, wd, lo, hi, tmax, t);
    }
};

struct PlainTrace { // every triangle commits, procedural primitives never do
    __device__ __forceinline__ bool surface(uint32_t, uint32_t, float, float, float) const { return true; }
    __device__ __forceinline__ bool procedural(uint32_t, uint32_t, const float *, const float *, const float *,
                                               const float *, float4, float4, float, float &) const {
        return false;
    }
};

template<class Policy>
__device__ __forceinline__ rq::CommittedHit walk(const Accel a, const Ray &r, uint32_t mask, bool first,
                                                 const Policy &pol) {
    uint2 local[kStack];
    rq::QueryWalker<Policy, 1, 0> W;
    W.S = rq::LaneStack<1, 0>{nullptr, 0u, local, (uint32_t)kStack, a.desc->error, 0};
    const float4 r0 = make_float4(r.origin[0], r.origin[1], r.origin[2], r.t_min);
    const float4 r1 = make_float4(r.direction[0], r.direction[1], r.direction[2], r.t_max);
    const QueryGeometry Q = a.desc->geo;
    W.start(Q, r0, r1, mask, first);
    while (!W.run(Q, pol, 1 << 30)) {
    }
    return W.hit;
}

// CallOp::RAY_TRACING_TRACE_CLOSEST
__device__ __forceinline__ SurfaceHit trace_closest(const Accel a, const Ray &r, uint32_t mask = 0xFFu) {
    const rq::CommittedHit h = walk(a, r, mask, false, PlainTrace{});
    return SurfaceHit{h.inst, h.prim, {h.u, h.v}, h.t, 0u};
}

// CallOp::RAY_TRACING_TRACE_ANY
__device__ __forceinline__ bool trace_any(const Accel a, const Ray &r, uint32_t mask = 0xFFu) {
    return walk(a, r, mask, true, PlainTrace{}).type != rq::kHitMiss;
}

// accel.traverse (any = false) / traverse_any (any = true) with candidate callbacks
//   on_surface(inst, prim, u, v, t) -> commit?
//   on_procedural(inst, prim, obj_o, obj_d, world_o, world_d, aabb_lo, aabb_hi, t_max, &t) -> commit t?
template<class S, class P>
__device__ __forceinline__ rq::CommittedHit query(const Accel a, const Ray &r, uint32_t mask, bool any, S on_surface,
                                                  P on_procedural) {
    return walk(a, r, mask, any, CallbackPolicy<S, P>{on_surface, on_procedural});
}

// Instance ops (RAY_TRACING_INSTANCE_TRANSFORM / _USER_ID / _VISIBILITY_MASK and SET_ forms,
// include/luisa/ast/op.h:290-297). Out-of-range indices read zeros and are not written; they
// set the table's error word, which fails the accel's next build.
__device__ __forceinline__ bool instance_ok(const Accel a, uint32_t i) {
    if (i < a.desc->state_count) return true;
    atomicOr(reinterpret_cast<uint32_t *>(a.desc->state + a.desc->state_count), 1u);
    return false;
}
__device__ __forceinline__ void instance_transform(const Accel a, uint32_t i, float out[12]) {
    const bool ok = instance_ok(a, i);
    for (int c = 0; c < 12; c++) out[c] = ok ? a.desc->state[i].affine[c] : 0.0f;
}
__device__ __forceinline__ uint32_t instance_user_id(const Accel a, uint32_t i) {
    return instance_ok(a, i) ? a.desc->state[i].user_id : 0u;
}
__device__ __forceinline__ uint32_t instance_visibility_mask(const Accel a, uint32_t i) {
    return instance_ok(a, i) ? a.desc->state[i].vis_mask : 0u;
}
__device__ __forceinline__ void set_instance_transform(const Accel a, uint32_t i, const float m[12]) {
    if (!instance_ok(a, i)) return;
    for (int c = 0; c < 12; c++) a.desc->state[i].affine[c] = m[c];
    atomicOr(&a.desc->state[i].dirty, 1u << kInstanceTransform);
}
__device__ __forceinline__ void set_instance_user_id(const Accel a, uint32_t i, uint32_t id) {
    if (!instance_ok(a, i)) return;
    a.desc->state[i].user_id = id;
    atomicOr(&a.desc->state[i].dirty, 1u << kInstanceUserId);
}
__device__ __forceinline__ void set_instance_visibility_mask(const Accel a, uint32_t i, uint32_t m) {
    if (!instance_ok(a, i)) return;
    a.desc->state[i].vis_mask = m;
    atomicOr(&a.desc->state[i].dirty, 1u << kInstanceVisibility);
}

} // namespace dev
} // namespace lchip
