"""Shader sources of the run-time shader tests: the kernels a DSL codegen would emit for the
reference's ray-tracing programs, written against the device library (lc_hip_device.h)."""
from luisacompute_amd.programs import QUERY  # noqa: F401  (the benchmark's query shader)

# hit = accel.intersect(ray, {mask}) / hit_any = accel.intersect_any(ray, {mask}) per thread
# (src/dsl/rtx/accel.cpp:136-150)
TRACE = r'''
#include "lc_hip_device.h"
using namespace lchip;
struct Params {
    alignas(16) dev::Buffer<dev::Ray> rays;
    alignas(16) dev::Buffer<dev::SurfaceHit> hits;
    alignas(16) dev::Buffer<uint8_t> any;
    alignas(16) dev::Accel accel;
    alignas(16) uint32_t mask;
    alignas(16) uint4 ls_kid;
};
extern "C" __global__ void __launch_bounds__(128) trace_kernel(const Params p) {
    if (!dev::in_dispatch(p.ls_kid)) return;
    const uint32_t i = dev::dispatch_id().x;
    const dev::Ray r = p.rays[i];
    p.hits[i] = dev::trace_closest(p.accel, r, p.mask);
    p.any[i] = dev::trace_any(p.accel, r, p.mask) ? 1u : 0u;
}
'''

# a surface callback of the shader's own: commit only even primitive ids; procedural ignored
QUERY_EVEN = r'''
#include "lc_hip_device.h"
using namespace lchip;
struct Params {
    alignas(16) dev::Buffer<dev::Ray> rays;
    alignas(16) dev::Buffer<rq::CommittedHit> out;
    alignas(16) dev::Accel accel;
    alignas(16) uint4 ls_kid;
};
extern "C" __global__ void __launch_bounds__(128) even_kernel(const Params p) {
    if (!dev::in_dispatch(p.ls_kid)) return;
    const uint32_t i = dev::dispatch_id().x;
    auto even = [](uint32_t, uint32_t prim, float, float, float) { return (prim & 1u) == 0u; };
    auto none = [](uint32_t, uint32_t, const float *, const float *, const float *, const float *, float4, float4,
                   float, float &) { return false; };
    p.out[i] = dev::query(p.accel, p.rays[i], 0xFFu, false, even, none);
}
'''

# RAY_TRACING_INSTANCE_TRANSFORM / _USER_ID / _VISIBILITY_MASK reads, then SET_INSTANCE_USER_ID
# and SET_INSTANCE_VISIBILITY_MASK on the instances named by `writes` (include/luisa/ast/op.h:290-297)
INSTANCES = r'''
#include "lc_hip_device.h"
using namespace lchip;
struct Params {
    alignas(16) dev::Buffer<float> transforms;
    alignas(16) dev::Buffer<uint32_t> ids;
    alignas(16) dev::Accel accel;
    alignas(16) uint32_t new_id_base;
    alignas(16) uint32_t new_mask;
    alignas(16) uint4 ls_kid;
};
extern "C" __global__ void __launch_bounds__(64) instance_kernel(const Params p) {
    if (!dev::in_dispatch(p.ls_kid)) return;
    const uint32_t i = dev::dispatch_id().x;
    float m[12];
    dev::instance_transform(p.accel, i, m);
    for (int c = 0; c < 12; c++) p.transforms[12 * (uint64_t)i + c] = m[c];
    p.ids[2 * i] = dev::instance_user_id(p.accel, i);
    p.ids[2 * i + 1] = dev::instance_visibility_mask(p.accel, i);
    if (i % 3 == 0) {
        dev::set_instance_user_id(p.accel, i, p.new_id_base + i);
        dev::set_instance_visibility_mask(p.accel, i, p.new_mask);
    }
}
'''

# make_ray / offset_ray_origin / triangle_interpolate (src/dsl/rtx/ray.cpp:7-28, hit.cpp:15-17)
HELPERS = r'''
#include "lc_hip_device.h"
using namespace lchip;
struct Params {
    alignas(16) dev::Buffer<float> in;   // per item: p, n, w, a, b, c (3 floats each), u, v
    alignas(16) dev::Buffer<float> out;  // per item: offset(p, n), offset(p, n, w), interpolate, make_ray (8)
    alignas(16) uint4 ls_kid;
};
extern "C" __global__ void __launch_bounds__(64) helper_kernel(const Params p) {
    if (!dev::in_dispatch(p.ls_kid)) return;
    const uint64_t i = dev::dispatch_id().x;
    const float *x = &p.in[20 * i];
    const rtx::V3 P = rtx::v3(x[0], x[1], x[2]), N = rtx::v3(x[3], x[4], x[5]), W = rtx::v3(x[6], x[7], x[8]);
    const rtx::V3 A = rtx::v3(x[9], x[10], x[11]), B = rtx::v3(x[12], x[13], x[14]), C = rtx::v3(x[15], x[16], x[17]);
    const rtx::V3 o1 = rtx::offset_ray_origin(P, N), o2 = rtx::offset_ray_origin(P, N, W);
    const rtx::V3 q = rtx::triangle_interpolate(x[18], x[19], A, B, C);
    const dev::Ray r = dev::make_ray(o1, N);
    float *y = &p.out[17 * i];
    y[0] = o1.x; y[1] = o1.y; y[2] = o1.z;
    y[3] = o2.x; y[4] = o2.y; y[5] = o2.z;
    y[6] = q.x; y[7] = q.y; y[8] = q.z;
    y[9] = r.origin[0]; y[10] = r.origin[1]; y[11] = r.origin[2]; y[12] = r.t_min;
    y[13] = r.direction[0]; y[14] = r.direction[1]; y[15] = r.direction[2]; y[16] = r.t_max;
}
'''

ALL = {"trace_kernel": TRACE, "query_kernel": QUERY, "even_kernel": QUERY_EVEN, "instance_kernel": INSTANCES,
       "helper_kernel": HELPERS}

# a path tracer's heap access (src/tests/test_path_tracing.cpp:212-222): per hit, the instance's
# vertex and triangle buffers come from a bindless array (slots 2i, 2i + 1), the hit point from
# triangle_interpolate of the three vertices
BINDLESS = r'''
#include "lc_hip_device.h"
using namespace lchip;
struct Params {
    alignas(16) dev::Buffer<dev::Ray> rays;
    alignas(16) dev::Buffer<float4> points;   // hit point, w = the slot's triangle count; zeros on a miss
    alignas(16) dev::Accel accel;
    alignas(16) dev::BindlessArray heap;
    alignas(16) uint4 ls_kid;
};
extern "C" __global__ void __launch_bounds__(128) bindless_kernel(const Params p) {
    if (!dev::in_dispatch(p.ls_kid)) return;
    const uint32_t i = dev::dispatch_id().x;
    const dev::SurfaceHit h = dev::trace_closest(p.accel, p.rays[i]);
    if (h.inst == 0xFFFFFFFFu) { p.points[i] = make_float4(0.0f, 0.0f, 0.0f, 0.0f); return; }
    const uint64_t ntri = dev::bindless_buffer_size<uint3>(p.heap, 2 * h.inst + 1);
    if (h.prim >= ntri) { p.points[i] = make_float4(0.0f, 0.0f, 0.0f, -1.0f); return; } // reads stay in bounds
    const uint3 t = dev::bindless_buffer_read<uint3>(p.heap, 2 * h.inst + 1, h.prim);
    const float4 a = dev::bindless_buffer_read<float4>(p.heap, 2 * h.inst, t.x);
    const float4 b = dev::bindless_buffer_read<float4>(p.heap, 2 * h.inst, t.y);
    const float4 c = dev::bindless_buffer_read<float4>(p.heap, 2 * h.inst, t.z);
    const rtx::V3 q = rtx::triangle_interpolate(h.bary[0], h.bary[1], rtx::v3(a.x, a.y, a.z), rtx::v3(b.x, b.y, b.z),
                                                rtx::v3(c.x, c.y, c.z));
    p.points[i] = make_float4(q.x, q.y, q.z, (float)ntri);
}
'''

ALL["bindless_kernel"] = BINDLESS

# image reads and writes in several storages (the accumulate / hdr2ldr kernels of
# src/tests/test_path_tracing.cpp:264-288 move pixels between FLOAT4, HALF4 and BYTE4 images)
IMAGES = r'''
#include "lc_hip_device.h"
using namespace lchip;
struct Params {
    alignas(16) dev::Image src;   // FLOAT4
    alignas(16) dev::Image h4;    // HALF4
    alignas(16) dev::Image b4;    // BYTE4
    alignas(16) dev::Image u1;    // INT1 as a uint image
    alignas(16) dev::Image s2;    // SHORT2
    alignas(16) dev::Buffer<float4> back;
    alignas(16) uint4 ls_kid;
};
extern "C" __global__ void __launch_bounds__(64, 1) image_kernel(const Params p) {
    if (!dev::in_dispatch(p.ls_kid)) return;
    const uint3 id = dev::dispatch_id();
    const float4 v = dev::image_read_float(p.src, id.x, id.y);
    dev::image_write_float(p.h4, id.x, id.y, v);
    dev::image_write_float(p.b4, id.x, id.y, v);
    dev::image_write_float(p.s2, id.x, id.y, v);
    dev::image_write_uint(p.u1, id.x, id.y, make_uint4(id.x * 7u + id.y, 0u, 0u, 0u));
    const float4 b = dev::image_read_float(p.b4, id.x, id.y);
    const uint4 u = dev::image_read_uint(p.u1, id.x, id.y);
    const float4 s = dev::image_read_float(p.s2, id.x, id.y);
    p.back[(uint64_t)id.y * p.ls_kid.x + id.x] = make_float4(b.x, b.w, (float)u.x + (float)u.y, s.y + s.z);
}
'''

ALL["image_kernel"] = IMAGES

# volumes: read a FLOAT2 volume, write a BYTE4 and an INT1 volume, read them back into a buffer
VOLUMES = r'''
#include "lc_hip_device.h"
using namespace lchip;
struct Params {
    alignas(16) dev::Volume src;  // FLOAT2
    alignas(16) dev::Volume b4;   // BYTE4
    alignas(16) dev::Volume i1;   // INT1
    alignas(16) dev::Buffer<float4> back;
    alignas(16) uint4 ls_kid;
};
extern "C" __global__ void __launch_bounds__(64, 1) volume_kernel(const Params p) {
    if (!dev::in_dispatch(p.ls_kid)) return;
    const uint3 id = dev::dispatch_id();
    const float4 v = dev::volume_read_float(p.src, id.x, id.y, id.z);
    dev::volume_write_float(p.b4, id.x, id.y, id.z, make_float4(v.x, v.y, 0.5f, 1.0f));
    dev::volume_write_int(p.i1, id.x, id.y, id.z, make_int4((int)(id.x + 10u * id.y + 100u * id.z) - 500, 0, 0, 0));
    const float4 b = dev::volume_read_float(p.b4, id.x, id.y, id.z);
    const int4 i = dev::volume_read_int(p.i1, id.x, id.y, id.z);
    p.back[((uint64_t)id.z * p.ls_kid.y + id.y) * p.ls_kid.x + id.x] = make_float4(b.x, v.w, (float)i.x, (float)p.src.depth);
}
'''

ALL["volume_kernel"] = VOLUMES

# indirect dispatch: a producer fills the indirect buffer, a consumer records its kernel id and
# dispatch id (the pattern of src/tests/test_indirect.cpp)
INDIRECT_PRODUCER = r'''
#include "lc_hip_device.h"
using namespace lchip;
struct Params {
    alignas(16) dev::Buffer<uint32_t> indirect;
    alignas(16) uint32_t count;
    alignas(16) uint4 ls_kid;
};
extern "C" __global__ void __launch_bounds__(64) producer(const Params p) {
    if (!dev::in_dispatch(p.ls_kid)) return;
    const uint32_t i = dev::dispatch_id().x;
    if (i == 0) dev::indirect_set_dispatch_count(p.indirect, p.count);
    dev::indirect_set_dispatch_kernel(p.indirect, i, make_uint3(64, 1, 1), make_uint3(3 * i + 1, 1, 1), 10 + i);
}
'''
INDIRECT_CONSUMER = r'''
#include "lc_hip_device.h"
using namespace lchip;
struct Params {
    alignas(16) dev::Buffer<uint32_t> out;
    alignas(16) uint4 ls_kid;
};
extern "C" __global__ void __launch_bounds__(64) consumer(const Params p) {
    if (!dev::in_dispatch(p.ls_kid)) return;
    const uint32_t i = dev::dispatch_id().x;
    const uint32_t k = dev::kernel_id(p.ls_kid);
    p.out[64 * (k % 16) + i] = 1000u * k + i + 1u;
}
'''

ALL["producer"] = INDIRECT_PRODUCER
ALL["consumer"] = INDIRECT_CONSUMER
