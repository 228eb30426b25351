// query.hip -- batched ray queries (RayQueryAll / RayQueryAny: accel.traverse / traverse_any,
// include/luisa/dsl/rtx/ray_query.h:11-243) with built-in candidate callbacks, the form the C ABI
// exposes; kernels of their own include ray_query.h with their callbacks instead.
//
// Built-in callbacks (the reference's own, src/tests/test_procedural.cpp:86-120):
//   surface (non-opaque instances): commit all, ignore all, or the rounded cutout: with
//     (w, u, v) = (1 - u - v, u, v), commit iff length(w, u), length(u, v) and length(v, w) are
//     all below `cutout`;
//   procedural: ignore, or the sphere inscribed in the primitive's AABB (centre = box centre,
//     radius = half the smallest extent) met by the object-space ray: with L = c - o,
//     tc = dot(L, d) / dot(d, d), the centre must lie ahead (tc > 0), the squared distance
//     dot(L, L) - tc^2 dot(d, d) of the ray from the centre at most r^2, and the committed
//     distance is tc - sqrt((r^2 - that) / dot(d, d)) (the test's formula for unit directions,
//     in the ray parameter for any direction). Division and sqrtf are correctly rounded (hipcc's
//     default -fhip-fp32-correctly-rounded-divide-sqrt; __fsqrt_rn is the native approximation).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "lc_hip_internal.h"
#include "ray_query.h"

namespace lchip {

namespace {

struct BuiltinPolicy {
    uint32_t surface_mode;
    float cutout;
    uint32_t procedural_mode;

    __device__ __forceinline__ static float len2(float a, float b) { return sqrtf(a * a + b * b); }

    __device__ __forceinline__ bool surface(uint32_t, uint32_t, float u, float v, float) const {
        if (surface_mode == kQuerySurfaceCommit) return true;
        if (surface_mode != kQuerySurfaceCutout) return false;
        const float w = (1.0f - u) - v;
        return len2(w, u) < cutout && len2(u, v) < cutout && len2(v, w) < cutout;
    }

    __device__ __forceinline__ bool procedural(uint32_t, uint32_t, const float o[3], const float d[3], const float *,
                                               const float *, float4 lo, float4 hi, float, float &t) const {
        if (procedural_mode != kQueryProceduralSphere) return false;
        const float c[3] = {(lo.x + hi.x) * 0.5f, (lo.y + hi.y) * 0.5f, (lo.z + hi.z) * 0.5f};
        const float r = 0.5f * fminf(fminf(hi.x - lo.x, hi.y - lo.y), hi.z - lo.z);
        const float L[3] = {c[0] - o[0], c[1] - o[1], c[2] - o[2]};
        const float dd = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2];
        const float ld = (L[0] * d[0] + L[1] * d[1]) + L[2] * d[2];
        const float ll = (L[0] * L[0] + L[1] * L[1]) + L[2] * L[2];
        const float tc = ld / dd;
        if (!(tc > 0.0f)) return false;
        const float q = ll - (tc * tc) * dd;
        const float rr = r * r;
        if (!(q <= rr)) return false;
        t = tc - sqrtf((rr - q) / dd);
        return true;
    }
};

constexpr int kQueryBlock = 128;
constexpr int kQuerySlds = 8;
constexpr int kQuerySteps = 16;     // walk steps between refills of finished lanes
constexpr uint32_t kQueryChunk = 256; // queue slots a wave claims per atomic

__global__ void k_query_zero(uint32_t *counter) {
    if (threadIdx.x == 0) *counter = 0u;
}

// Persistent launch: a wave claims kQueryChunk queue slots per atomic and hands them to its idle
// lanes; each lane runs kQuerySteps steps of its query between refills, so lanes whose query
// ended take the next ray instead of idling until the wave's longest query ends. `order` (optional)
// is the queue in traversal order (rays sorted by origin cell and direction octant).
__global__ __launch_bounds__(kQueryBlock) void k_query(const QueryParams P) {
    __shared__ uint2 stk[kQuerySlds][kQueryBlock];
    const uint32_t lane = threadIdx.x & 63u;
    const uint64_t glane = (uint64_t)blockIdx.x * kQueryBlock + threadIdx.x;
    rq::QueryWalker<BuiltinPolicy, kQueryBlock, kQuerySlds> W;
    W.S = rq::LaneStack<kQueryBlock, kQuerySlds>{stk, threadIdx.x, P.spill + glane * (uint64_t)P.spill_depth,
                                                 P.spill_depth, P.error, 0};
    const BuiltinPolicy pol{P.surface_mode, P.cutout, P.procedural_mode};
    const uint32_t count = P.count;
    uint32_t pool = 0, pool_end = 0; // wave-uniform: claimed, unassigned slots [pool, pool_end)
    bool drained = false;            // wave-uniform: the queue is used up
    bool active = false, exhausted = false;
    uint32_t ray = 0;
    for (;;) {
        const uint64_t need = __ballot(!active && !exhausted);
        if (need) {
            const uint32_t n_need = (uint32_t)__popcll(need);
            const uint32_t rank = (uint32_t)__popcll(need & ((1ull << lane) - 1ull));
            const uint32_t avail = pool_end - pool;
            uint32_t my = rank < avail ? pool + rank : 0xFFFFFFFFu;
            if (avail >= n_need) {
                pool += n_need;
            } else if (drained) {
                pool = pool_end;
            } else {
                const uint32_t leader = (uint32_t)__ffsll((long long)need) - 1u;
                uint32_t b = 0;
                if (lane == leader) b = atomicAdd(P.counter, kQueryChunk);
                b = __builtin_amdgcn_readlane(b, (int)leader);
                const uint32_t nb = min(b, count), ne = (uint32_t)min<uint64_t>((uint64_t)b + kQueryChunk, count);
                drained = (uint64_t)b + kQueryChunk >= count;
                if (rank >= avail && rank - avail < ne - nb) my = nb + (rank - avail);
                pool = nb + min(n_need - avail, ne - nb);
                pool_end = ne;
            }
            if (!active && !exhausted) {
                if (my < count) {
                    ray = P.order ? P.order[my] : my;
                    W.start(P.geo, P.rays[2 * (uint64_t)ray], P.rays[2 * (uint64_t)ray + 1], P.mask, P.first != 0u);
                    active = true;
                } else if (drained) {
                    exhausted = true;
                }
            }
        }
        if (!__any(active)) break;
        if (active && W.run(P.geo, pol, kQuerySteps)) {
            const rq::CommittedHit &h = W.hit;
            uint2 *o = reinterpret_cast<uint2 *>(static_cast<char *>(P.hits) + 24ull * ray);
            o[0] = make_uint2(h.inst, h.prim);
            o[1] = make_uint2(__float_as_uint(h.u), __float_as_uint(h.v));
            o[2] = make_uint2(h.type, __float_as_uint(h.t));
            active = false;
        }
    }
}

} // namespace

int query_block() { return kQueryBlock; }
int query_lds_entries() { return kQuerySlds; }

int query_blocks_per_cu() {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_query, kQueryBlock, 0) != hipSuccess || n < 1) n = 1;
    return n;
}

hipError_t query_launch(const QueryParams &P, hipStream_t stream, int grid) {
    if (P.count == 0) return hipSuccess;
    hipLaunchKernelGGL(k_query_zero, dim3(1), dim3(64), 0, stream, P.counter);
    hipLaunchKernelGGL(k_query, dim3(grid), dim3(kQueryBlock), 0, stream, P);
    return hipGetLastError();
}

} // namespace lchip
