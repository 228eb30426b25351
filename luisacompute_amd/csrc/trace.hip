// trace.hip -- software trace_closest / trace_any on gfx950 (no ray-tracing hardware on CDNA4).
//
// Replaces the reference's lc_accel_trace_closest / lc_accel_trace_any
// (src/backends/cuda/cuda_builtin/cuda_device_resource.h:1783-1829, OptiX hardware traversal)
// with the observable semantics of the CPU backend's AccelImpl::trace_closest / trace_any
// (src/rust/luisa_compute_backend_impl/src/cpu/accel.rs:449-535):
//   hit  -> {inst, prim, bary = (u, v) weighting vertices 1 and 2, t}
//   miss -> {inst = prim = 0xFFFFFFFF, bary = (0, 0), t = ray.tmax}
//   any  -> 1 if some triangle is hit with t in [tmin, tmax]
//   an instance takes part iff (instance mask & ray mask & 0xff) != 0
//   (cuda_device_resource.h:1776).
//
// Design (wave64):
//   * persistent grid, rays fetched per wave with one atomic for all idle lanes;
//   * 8-wide compressed nodes (lc_hip_bvh.h); a lane keeps a "node group" (child base + hit
//     mask in visit order + internal mask) and a "triangle group" (base + 32-bit mask), the
//     traversal scheme of Ylitie et al. HPG 2017, so a node's hit children cost one stack entry;
//   * stack entries in LDS ([depth][block] layout: lanes of a wave hit consecutive banks), with
//     a per-lane spill area in global memory for depths beyond the LDS part;
//   * box test in the node-local form t = q * (s / d) + (p - o) / d with a per-axis error
//     slack, so the conservative quantised boxes are never culled by rounding;
//   * watertight ray/triangle test (Woop, Benthin, Wald, JCGT 2013) in IEEE single precision
//     with no contraction (compiled with -ffp-contract=off), in the same operation order as
//     the CPU oracle, so t and barycentrics agree bit for bit on the same triangle.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>

#include "lc_hip_bvh.h"
#include "lc_hip_internal.h"

namespace lchip {

namespace {

struct RaySetup {
    float o[3], d[3];
    float inv[3];
    int kx, ky, kz;
    float Sx, Sy, Sz;
    uint32_t oct;
};

__device__ __forceinline__ float sel3(const float v[3], int k) { return k == 0 ? v[0] : (k == 1 ? v[1] : v[2]); }

__device__ __forceinline__ void ray_setup(RaySetup &r) {
    float ax = fabsf(r.d[0]), ay = fabsf(r.d[1]), az = fabsf(r.d[2]);
    int kz = (ax > ay) ? ((ax > az) ? 0 : 2) : ((ay > az) ? 1 : 2);
    int kx = kz + 1 == 3 ? 0 : kz + 1;
    int ky = kx + 1 == 3 ? 0 : kx + 1;
    float dkz = sel3(r.d, kz);
    if (dkz < 0.0f) { int t = kx; kx = ky; ky = t; }
    r.kx = kx; r.ky = ky; r.kz = kz;
    r.Sx = sel3(r.d, kx) / dkz;
    r.Sy = sel3(r.d, ky) / dkz;
    r.Sz = 1.0f / dkz;
    for (int a = 0; a < 3; a++) {
        float dd = r.d[a];
        if (fabsf(dd) < 1e-20f) dd = copysignf(1e-20f, dd);
        r.inv[a] = 1.0f / dd;
    }
    r.oct = (r.d[0] < 0.0f ? 1u : 0u) | (r.d[1] < 0.0f ? 2u : 0u) | (r.d[2] < 0.0f ? 4u : 0u);
}

// Watertight test; identical operation order to the oracle's wt_intersect.
__device__ __forceinline__ bool wt_intersect(const RaySetup &r, const float4 &a, const float4 &b, const float4 &c,
                                             float tmin, float tmax, float &t_out, float &u_out, float &v_out) {
    const float v0[3] = {a.x, a.y, a.z}, v1[3] = {b.x, b.y, b.z}, v2[3] = {c.x, c.y, c.z};
    const int kx = r.kx, ky = r.ky, kz = r.kz;
    const float okx = sel3(r.o, kx), oky = sel3(r.o, ky), okz = sel3(r.o, kz);
    const float A_x = sel3(v0, kx) - okx, A_y = sel3(v0, ky) - oky, A_z = sel3(v0, kz) - okz;
    const float B_x = sel3(v1, kx) - okx, B_y = sel3(v1, ky) - oky, B_z = sel3(v1, kz) - okz;
    const float C_x = sel3(v2, kx) - okx, C_y = sel3(v2, ky) - oky, C_z = sel3(v2, kz) - okz;
    const float Ax = A_x - r.Sx * A_z, Ay = A_y - r.Sy * A_z;
    const float Bx = B_x - r.Sx * B_z, By = B_y - r.Sy * B_z;
    const float Cx = C_x - r.Sx * C_z, Cy = C_y - r.Sy * C_z;
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        double CxBy = (double)Cx * (double)By, CyBx = (double)Cy * (double)Bx;
        double AxCy = (double)Ax * (double)Cy, AyCx = (double)Ay * (double)Cx;
        double BxAy = (double)Bx * (double)Ay, ByAx = (double)By * (double)Ax;
        U = (float)(CxBy - CyBx);
        V = (float)(AxCy - AyCx);
        W = (float)(BxAy - ByAx);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    const float det = U + V + W;
    if (det == 0.0f) return false;
    const float Az = r.Sz * A_z, Bz = r.Sz * B_z, Cz = r.Sz * C_z;
    const float T = U * Az + V * Bz + W * Cz;
    const float rcp = 1.0f / det;
    const float t = T * rcp;
    if (!(t >= tmin && t <= tmax)) return false;
    t_out = t;
    u_out = V * rcp;
    v_out = W * rcp;
    return true;
}

// Bit i of the result = bit (i ^ f) of m: slot order -> visit order for octant f.
__device__ __forceinline__ uint32_t permute8(uint32_t m, uint32_t f) {
    if (f & 1u) m = ((m & 0x55u) << 1) | ((m & 0xAAu) >> 1);
    if (f & 2u) m = ((m & 0x33u) << 2) | ((m & 0xCCu) >> 2);
    if (f & 4u) m = ((m & 0x0Fu) << 4) | ((m & 0xF0u) >> 4);
    return m;
}

__device__ __forceinline__ float ubyte(uint32_t w, int k) { return (float)((w >> (8 * k)) & 0xFFu); }

// Tests the 8 children of a node; returns the hit mask in slot order (occupied slots only).
__device__ __forceinline__ uint32_t intersect_node(const RaySetup &r, const uint4 w0, const uint4 w1, const uint4 w2,
                                                   const uint4 w3, const uint4 w4, float tmin, float tcull) {
    const float p[3] = {__uint_as_float(w0.x), __uint_as_float(w0.y), __uint_as_float(w0.z)};
    float a[3], bn[3], bf[3];
    for (int k = 0; k < 3; k++) {
        float s = __uint_as_float(((w0.w >> (8 * k)) & 0xFFu) << 23);
        a[k] = s * r.inv[k];
        float b = (p[k] - r.o[k]) * r.inv[k];
        // |error| of fma(q, a, b) against the exact slab distance is below 6 ulp of
        // 255|a| + |b|; 2^-21 (16 ulp) keeps the test conservative.
        float e = 0x1p-21f * (255.0f * fabsf(a[k]) + fabsf(b));
        bn[k] = b - e;
        bf[k] = b + e;
    }
    // near/far plane words per axis (x: lo = w2.xy hi = w3.zw; y: lo = w2.zw hi = w4.xy;
    // z: lo = w3.xy hi = w4.zw); a negative direction swaps them.
    const bool nx = r.inv[0] < 0.0f, ny = r.inv[1] < 0.0f, nz = r.inv[2] < 0.0f;
    const uint32_t nx0 = nx ? w3.z : w2.x, nx1 = nx ? w3.w : w2.y, fx0 = nx ? w2.x : w3.z, fx1 = nx ? w2.y : w3.w;
    const uint32_t ny0 = ny ? w4.x : w2.z, ny1 = ny ? w4.y : w2.w, fy0 = ny ? w2.z : w4.x, fy1 = ny ? w2.w : w4.y;
    const uint32_t nz0 = nz ? w4.z : w3.x, nz1 = nz ? w4.w : w3.y, fz0 = nz ? w3.x : w4.z, fz1 = nz ? w3.y : w4.w;
    uint32_t hits = 0;
#pragma unroll
    for (int s = 0; s < 8; s++) {
        const int k = s & 3;
        const uint32_t wnx = s < 4 ? nx0 : nx1, wfx = s < 4 ? fx0 : fx1;
        const uint32_t wny = s < 4 ? ny0 : ny1, wfy = s < 4 ? fy0 : fy1;
        const uint32_t wnz = s < 4 ? nz0 : nz1, wfz = s < 4 ? fz0 : fz1;
        const uint32_t wm = s < 4 ? w1.z : w1.w;
        float tnx = __builtin_fmaf(ubyte(wnx, k), a[0], bn[0]);
        float tny = __builtin_fmaf(ubyte(wny, k), a[1], bn[1]);
        float tnz = __builtin_fmaf(ubyte(wnz, k), a[2], bn[2]);
        float tfx = __builtin_fmaf(ubyte(wfx, k), a[0], bf[0]);
        float tfy = __builtin_fmaf(ubyte(wfy, k), a[1], bf[1]);
        float tfz = __builtin_fmaf(ubyte(wfz, k), a[2], bf[2]);
        float tn = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, tmin));
        float tf = fminf(fminf(tfx, tfy), fminf(tfz, tcull));
        bool occ = ((wm >> (8 * k)) & 0xFFu) != 0u;
        if (occ && tn <= tf) hits |= 1u << s;
    }
    return hits;
}

__device__ __forceinline__ void xform_point(const float m[12], const float p[3], float out[3]) {
    for (int r = 0; r < 3; r++) out[r] = m[4 * r + 0] * p[0] + m[4 * r + 1] * p[1] + m[4 * r + 2] * p[2] + m[4 * r + 3];
}
__device__ __forceinline__ void xform_vector(const float m[12], const float p[3], float out[3]) {
    for (int r = 0; r < 3; r++) out[r] = m[4 * r + 0] * p[0] + m[4 * r + 1] * p[1] + m[4 * r + 2] * p[2];
}

__device__ __forceinline__ void load_node(const WideNode *nodes, uint32_t idx, uint4 &w0, uint4 &w1, uint4 &w2,
                                          uint4 &w3, uint4 &w4) {
    const uint4 *p = reinterpret_cast<const uint4 *>(nodes + idx);
    w0 = p[0]; w1 = p[1]; w2 = p[2]; w3 = p[3]; w4 = p[4];
}

constexpr uint32_t kInstGroup = 0x80000000u; // stack entry tag: remaining instances of a TLAS leaf

} // namespace

// Template parameters: ANY (trace_any), TWO (TLAS + BLAS; otherwise one instance's BLAS),
// COUNT (count node visits / triangle tests), BLOCK threads, SLDS stack entries in LDS.
template<bool ANY, bool TWO, bool COUNT, int BLOCK, int SLDS>
__global__ __launch_bounds__(BLOCK) void k_trace(const TraceParams P) {
    __shared__ uint2 stk[SLDS > 0 ? SLDS : 1][BLOCK];
    const uint32_t tid = threadIdx.x;
    const uint32_t lane = tid & 63u;
    const uint64_t glane = (uint64_t)blockIdx.x * BLOCK + tid;
    uint2 *spill = P.spill + glane * (uint64_t)P.spill_depth;
    int sp = 0;
    // The stack never holds more than (wide depth - 1) node groups plus the TLAS leaf entries;
    // the host sizes spill_depth from the built depths. The bound checks only guard memory.
    auto push = [&](uint2 e) {
        if (sp < SLDS) stk[sp][tid] = e;
        else if ((uint32_t)(sp - SLDS) < P.spill_depth) spill[sp - SLDS] = e;
        else atomicOr(P.error, 1u);
        sp++;
    };
    auto pop = [&]() -> uint2 {
        sp--;
        if (sp < SLDS) return stk[sp][tid];
        if ((uint32_t)(sp - SLDS) < P.spill_depth) return spill[sp - SLDS];
        return make_uint2(0u, 0u);
    };

    bool active = false, exhausted = false;
    uint32_t ray_id = 0;    // index within the batch
    uint32_t batch = 0;     // batch of the current ray
    uint32_t ray_mask = 0;  // its visibility mask
    RaySetup R;           // current (object space when inside an instance)
    float wo[3], wd[3];   // world-space origin/direction (two-level)
    float tmin = 0.0f, tmax = 0.0f;
    float t_best = 0.0f, bu = 0.0f, bv = 0.0f;
    uint32_t b_prim = 0xFFFFFFFFu, b_inst = 0xFFFFFFFFu;
    bool any_hit = false;
    uint32_t g_base = 0, g_mask = 0, t_base = 0, t_mask = 0;
    // two-level state
    bool in_blas = !TWO;
    int sp_inst = 0;
    uint32_t cur_inst = 0;
    const WideNode *nodes = TWO ? P.tlas_nodes : P.blas_nodes;
    const TriRecord *tris = P.blas_tris;
    uint64_t n_nodes = 0, n_tris = 0;

    for (;;) {
        const uint64_t need = __ballot(!active && !exhausted);
        if (need) {
            const uint32_t leader = (uint32_t)__ffsll((long long)need) - 1u;
            uint32_t base = 0;
            if (lane == leader) base = atomicAdd(P.ray_counter, (uint32_t)__popcll(need));
            base = __shfl(base, (int)leader, 64);
            if (!active && !exhausted) {
                const uint32_t my = base + (uint32_t)__popcll(need & ((1ull << lane) - 1ull));
                if (my < P.count) {
                    const uint32_t q = P.order ? P.order[my] : my; // sorted queue: fetch in key order
                    batch = find_batch(P.batches, q);
                    ray_id = q - P.batches.b[batch].start;
                    ray_mask = P.batches.b[batch].mask;
                    const float4 *rays = P.batches.b[batch].rays;
                    const float4 r0 = rays[2 * (uint64_t)ray_id], r1 = rays[2 * (uint64_t)ray_id + 1];
                    wo[0] = r0.x; wo[1] = r0.y; wo[2] = r0.z; tmin = r0.w;
                    wd[0] = r1.x; wd[1] = r1.y; wd[2] = r1.z; tmax = r1.w;
                    t_best = FLT_MAX * 2.0f; // +inf
                    b_prim = 0xFFFFFFFFu; b_inst = 0xFFFFFFFFu; bu = 0.0f; bv = 0.0f;
                    any_hit = false;
                    sp = 0;
                    t_mask = 0;
                    if (TWO) {
                        for (int k = 0; k < 3; k++) { R.o[k] = wo[k]; R.d[k] = wd[k]; }
                        in_blas = false;
                        nodes = P.tlas_nodes;
                        ray_setup(R);
                        g_base = 0;
                        g_mask = P.empty ? 0u : ((1u << R.oct) | (1u << 8));
                    } else {
                        xform_point(P.inst_inv, wo, R.o);
                        xform_vector(P.inst_inv, wd, R.d);
                        ray_setup(R);
                        g_base = 0;
                        g_mask = (P.empty || (P.inst_mask & ray_mask & 0xFFu) == 0u) ? 0u : ((1u << R.oct) | (1u << 8));
                    }
                    active = true;
                } else {
                    exhausted = true;
                }
            }
        }
        if (!__any(active)) break;
        if (active) {
            bool done = false;
            for (int it = 0; it < P.steps; it++) {
                const float tcull = ANY ? tmax : fminf(tmax, t_best);
                if (t_mask != 0u) {
                    const uint32_t b = (uint32_t)__builtin_ctz(t_mask);
                    t_mask &= t_mask - 1u;
                    if (!TWO || in_blas) {
                        const float4 *tr = reinterpret_cast<const float4 *>(tris + t_base + b);
                        const float4 a = tr[0], bb = tr[1], c = tr[2];
                        if (COUNT) n_tris++;
                        float t, u, v;
                        if (wt_intersect(R, a, bb, c, tmin, tmax, t, u, v)) {
                            if (ANY) { any_hit = true; done = true; break; }
                            if (t < t_best) {
                                t_best = t; bu = u; bv = v;
                                b_prim = __float_as_uint(a.w);
                                b_inst = TWO ? cur_inst : 0u;
                            }
                        }
                    } else {
                        // TLAS leaf entry: enter instance t_base + b.
                        const uint32_t id = P.tlas_prims[t_base + b];
                        const InstanceRecord *ir = P.instances + id;
                        const uint32_t imask = ir->vis_mask;
                        const uint64_t bn = ir->blas_nodes;
                        if (bn != 0u && (imask & ray_mask & 0xFFu) != 0u) {
                            if (t_mask != 0u) {
                                const uint2 e = make_uint2(t_base, t_mask | kInstGroup);
                                push(e);
                            }
                            t_mask = 0u;
                            if ((g_mask & 0xFFu) != 0u) {
                                const uint2 e = make_uint2(g_base, g_mask);
                                push(e);
                            }
                            float inv[12];
                            const float4 *m = reinterpret_cast<const float4 *>(ir->inv);
                            const float4 m0 = m[0], m1 = m[1], m2 = m[2];
                            inv[0] = m0.x; inv[1] = m0.y; inv[2] = m0.z; inv[3] = m0.w;
                            inv[4] = m1.x; inv[5] = m1.y; inv[6] = m1.z; inv[7] = m1.w;
                            inv[8] = m2.x; inv[9] = m2.y; inv[10] = m2.z; inv[11] = m2.w;
                            xform_point(inv, wo, R.o);
                            xform_vector(inv, wd, R.d);
                            ray_setup(R);
                            nodes = reinterpret_cast<const WideNode *>(bn);
                            tris = reinterpret_cast<const TriRecord *>(ir->blas_tris);
                            cur_inst = id;
                            in_blas = true;
                            sp_inst = sp;
                            g_base = 0;
                            g_mask = (1u << R.oct) | (1u << 8);
                        }
                    }
                    continue;
                }
                if ((g_mask & 0xFFu) == 0u) {
                    if (TWO && in_blas && sp == sp_inst) {
                        // leave the instance: back to the world-space ray and the TLAS
                        for (int k = 0; k < 3; k++) { R.o[k] = wo[k]; R.d[k] = wd[k]; }
                        ray_setup(R);
                        nodes = P.tlas_nodes;
                        in_blas = false;
                    }
                    if (sp == 0) { done = true; break; }
                    const uint2 e = pop();
                    if (TWO && (e.y & kInstGroup)) {
                        t_base = e.x;
                        t_mask = e.y & ~kInstGroup;
                        continue;
                    }
                    g_base = e.x;
                    g_mask = e.y;
                }
                // visit the next node of the current group
                uint32_t hm = g_mask & 0xFFu;
                const uint32_t im = (g_mask >> 8) & 0xFFu;
                const uint32_t i = (uint32_t)__builtin_ctz(hm);
                hm &= hm - 1u;
                const uint32_t slot = i ^ R.oct;
                const uint32_t node = g_base + (uint32_t)__popc(im & ((1u << slot) - 1u));
                if (hm != 0u) {
                    const uint2 e = make_uint2(g_base, hm | (im << 8));
                    push(e);
                }
                g_mask = 0u;
                uint4 w0, w1, w2, w3, w4;
                load_node(nodes, node, w0, w1, w2, w3, w4);
                if (COUNT) n_nodes++;
                const uint32_t hits = intersect_node(R, w0, w1, w2, w3, w4, tmin, tcull);
                const uint32_t nim = w0.w >> 24;
                g_base = w1.x;
                g_mask = permute8(hits & nim, R.oct) | (nim << 8);
                uint32_t leaf = hits & ~nim;
                t_base = w1.y;
                t_mask = 0u;
                while (leaf) {
                    const uint32_t s = (uint32_t)__builtin_ctz(leaf);
                    leaf &= leaf - 1u;
                    const uint32_t m = ((s < 4 ? w1.z : w1.w) >> (8 * (s & 3))) & 0xFFu;
                    t_mask |= ((1u << (m >> 5)) - 1u) << (m & 31u);
                }
            }
            if (done) {
                // the counting form has no output buffers (out_hits/out_any are null)
                if (COUNT) {
                } else if (ANY) {
                    static_cast<uint8_t *>(P.batches.b[batch].out)[ray_id] = any_hit ? 1u : 0u;
                } else {
                    uint2 *h = reinterpret_cast<uint2 *>(static_cast<char *>(P.batches.b[batch].out) + 24ull * ray_id);
                    if (b_prim != 0xFFFFFFFFu) {
                        h[0] = make_uint2(b_inst, b_prim);
                        h[1] = make_uint2(__float_as_uint(bu), __float_as_uint(bv));
                        h[2] = make_uint2(__float_as_uint(t_best), 0u);
                    } else {
                        h[0] = make_uint2(0xFFFFFFFFu, 0xFFFFFFFFu);
                        h[1] = make_uint2(0u, 0u);
                        h[2] = make_uint2(__float_as_uint(tmax), 0u);
                    }
                }
                active = false;
            }
        }
    }
    if (COUNT) {
        for (int o = 32; o > 0; o >>= 1) {
            n_nodes += __shfl_xor(n_nodes, o, 64);
            n_tris += __shfl_xor(n_tris, o, 64);
        }
        if (lane == 0) {
            atomicAdd(reinterpret_cast<unsigned long long *>(P.stats), (unsigned long long)n_nodes);
            atomicAdd(reinterpret_cast<unsigned long long *>(P.stats) + 1, (unsigned long long)n_tris);
        }
    }
}

__global__ void k_zero_counter(uint32_t *c) { *c = 0u; }

namespace {

template<bool ANY, bool TWO, bool COUNT, int BLOCK, int SLDS>
hipError_t launch_one(const TraceParams &P, hipStream_t stream, int grid) {
    hipLaunchKernelGGL((k_trace<ANY, TWO, COUNT, BLOCK, SLDS>), dim3(grid), dim3(BLOCK), 0, stream, P);
    return hipGetLastError();
}

template<bool ANY, bool TWO, bool COUNT>
hipError_t launch_cfg(const TraceParams &P, hipStream_t stream, int grid, int block, int slds) {
#define LC_TRACE_CFG(B, S) \
    if (block == B && slds == S) return launch_one<ANY, TWO, COUNT, B, S>(P, stream, grid);
    LC_TRACE_CFG(64, 8) LC_TRACE_CFG(64, 16) LC_TRACE_CFG(64, 24)
    LC_TRACE_CFG(128, 8) LC_TRACE_CFG(128, 16) LC_TRACE_CFG(128, 24)
    LC_TRACE_CFG(256, 8) LC_TRACE_CFG(256, 16) LC_TRACE_CFG(256, 24)
#undef LC_TRACE_CFG
    return hipErrorInvalidValue;
}

template<bool ANY, bool TWO, bool COUNT, int BLOCK, int SLDS>
int occupancy_one() {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, (const void *)k_trace<ANY, TWO, COUNT, BLOCK, SLDS>, BLOCK, 0) != hipSuccess)
        return 1;
    return nb > 0 ? nb : 1;
}

template<bool ANY, bool TWO, bool COUNT>
int occupancy_cfg(int block, int slds) {
#define LC_OCC_CFG(B, S) \
    if (block == B && slds == S) return occupancy_one<ANY, TWO, COUNT, B, S>();
    LC_OCC_CFG(64, 8) LC_OCC_CFG(64, 16) LC_OCC_CFG(64, 24)
    LC_OCC_CFG(128, 8) LC_OCC_CFG(128, 16) LC_OCC_CFG(128, 24)
    LC_OCC_CFG(256, 8) LC_OCC_CFG(256, 16) LC_OCC_CFG(256, 24)
#undef LC_OCC_CFG
    return 1;
}

} // namespace

bool trace_config_supported(int block, int slds) {
    return (block == 64 || block == 128 || block == 256) && (slds == 8 || slds == 16 || slds == 24);
}

int trace_blocks_per_cu(bool any, bool two, bool count, int block, int slds) {
    if (any) {
        if (two) return count ? occupancy_cfg<true, true, true>(block, slds) : occupancy_cfg<true, true, false>(block, slds);
        return count ? occupancy_cfg<true, false, true>(block, slds) : occupancy_cfg<true, false, false>(block, slds);
    }
    if (two) return count ? occupancy_cfg<false, true, true>(block, slds) : occupancy_cfg<false, true, false>(block, slds);
    return count ? occupancy_cfg<false, false, true>(block, slds) : occupancy_cfg<false, false, false>(block, slds);
}

hipError_t trace_launch(const TraceParams &P, hipStream_t stream, bool any, bool two, bool count, int grid, int block,
                        int slds) {
    hipLaunchKernelGGL(k_zero_counter, dim3(1), dim3(1), 0, stream, P.ray_counter);
    if (any) {
        if (two) return count ? launch_cfg<true, true, true>(P, stream, grid, block, slds) : launch_cfg<true, true, false>(P, stream, grid, block, slds);
        return count ? launch_cfg<true, false, true>(P, stream, grid, block, slds) : launch_cfg<true, false, false>(P, stream, grid, block, slds);
    }
    if (two) return count ? launch_cfg<false, true, true>(P, stream, grid, block, slds) : launch_cfg<false, true, false>(P, stream, grid, block, slds);
    return count ? launch_cfg<false, false, true>(P, stream, grid, block, slds) : launch_cfg<false, false, false>(P, stream, grid, block, slds);
}

} // namespace lchip
