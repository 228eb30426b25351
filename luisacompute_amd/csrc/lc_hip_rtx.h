// lc_hip_rtx.h -- the DSL's ray-tracing helpers on float triples (src/dsl/rtx/ray.cpp:7-24,
// hit.cpp:15-17), shared by the library's path tracer and the device library of run-time shaders.
// Every operation is spelled out (no contraction, fixed order), so results are reproducible bit
// for bit on the CPU.
#pragma once

#ifndef __HIPCC_RTC__
#include <hip/hip_runtime.h>
#endif

namespace lchip {
namespace rtx {

// ---- float3 arithmetic, every operation spelled out (no contraction, fixed order) ----
struct V3 {
    float x, y, z;
};
__device__ __forceinline__ V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 sub(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ V3 normalize(V3 a) {
    const float l = sqrtf(dot(a, a));
    return v3(a.x / l, a.y / l, a.z / l);
}

// offset_ray_origin (src/dsl/rtx/ray.cpp:16-24): integer ulp offset of 256 n, or 2^-16 n near 0.
__device__ __forceinline__ float offset1(float p, float n) {
    const int of = (int)(256.0f * n);
    const float pi = __int_as_float(__float_as_int(p) + (p < 0.0f ? -of : of));
    return fabsf(p) < (1.0f / 32.0f) ? p + (1.0f / 65536.0f) * n : pi;
}
__device__ __forceinline__ V3 offset_ray_origin(V3 p, V3 n) {
    return v3(offset1(p.x, n.x), offset1(p.y, n.y), offset1(p.z, n.z));
}
// offset_ray_origin(p, n, w) (ray.cpp:26-28): n flipped to the side of w first
__device__ __forceinline__ V3 offset_ray_origin(V3 p, V3 n, V3 w) {
    return offset_ray_origin(p, dot(n, w) < 0.0f ? v3(-n.x, -n.y, -n.z) : n);
}

// triangle_interpolate (src/dsl/rtx/hit.cpp:15-17): (1 - u - v) a + u b + v c
__device__ __forceinline__ V3 triangle_interpolate(float u, float v, V3 a, V3 b, V3 c) {
    const float w = 1.0f - u - v;
    return v3(w * a.x + u * b.x + v * c.x, w * a.y + u * b.y + v * c.y, w * a.z + u * b.z + v * c.z);
}

} // namespace rtx
} // namespace lchip
