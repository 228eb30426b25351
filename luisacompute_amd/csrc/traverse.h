// traverse.h -- per-ray BVH8 traversal state machine shared by the persistent trace kernel
// (trace.hip) and the path-tracing kernels (pathtracer.hip).
//
// Observable semantics follow the CPU backend's AccelImpl::trace_closest / trace_any
// (src/rust/luisa_compute_backend_impl/src/cpu/accel.rs:449-535):
//   hit  -> {inst, prim, bary = (u, v) weighting vertices 1 and 2, t}
//   miss -> {inst = prim = 0xFFFFFFFF, bary = (0, 0), t = ray.tmax}
//   any  -> true if some triangle is hit with t in [tmin, tmax]
//   an instance takes part iff (instance mask & ray mask & 0xff) != 0
//   (src/backends/cuda/cuda_builtin/cuda_device_resource.h:1776).
// Among hits at the same t the smaller (instance, primitive) pair wins, so the result does not
// depend on the BVH or on the order in which leaves are visited.
//
// Scheme (Ylitie, Karras, Laine, HPG 2017): a lane keeps a "node group" (child base + hit mask in
// visit order + internal mask) and a "triangle group" (base + 32-bit mask), so the hit children
// of a node cost one stack entry. Stack entries live in LDS ([depth][block] layout, lanes of a
// wave on consecutive banks) with a per-lane spill area in global memory beyond that.
// Box test in node-local form t = q * (s / d) + (p - o) / d with a per-axis error slack, so the
// conservatively quantised boxes are never culled by rounding. Triangle test: watertight
// (Woop, Benthin, Wald, JCGT 2013) in IEEE single precision without contraction
// (-ffp-contract=off), in the same operation order as the CPU oracle.
#pragma once

#ifndef __HIPCC_RTC__ // (hiprtc provides the HIP device environment itself)
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>
#endif
#ifndef FLT_MAX
#define FLT_MAX 3.402823466e+38f
#endif

#include "lc_hip_bvh.h"

namespace lchip {
namespace trav {

struct RaySetup {
    float o[3], d[3];
    float inv[3];
    uint32_t kxyz; // watertight axis permutation: kx | ky << 2 | kz << 4
    float Sx, Sy, Sz;
    uint32_t oct;
};

__device__ __forceinline__ float sel3(const float v[3], int k) { return k == 0 ? v[0] : (k == 1 ? v[1] : v[2]); }

// (x, y, z)[k] for a lane-varying k in 0..2 as two flat selects on scalars: no array for the
// compiler to index (which would move it to scratch memory) and no branches (a nested conditional
// compiles to exec-masked branches, about 25 instructions per selection)
__device__ __forceinline__ float pick3(float x, float y, float z, uint32_t k) {
    float r = z;
    r = k == 1u ? y : r;
    r = k == 0u ? x : r;
    return r;
}

__device__ __forceinline__ void ray_setup(RaySetup &r) {
    float ax = fabsf(r.d[0]), ay = fabsf(r.d[1]), az = fabsf(r.d[2]);
    int kz = (ax > ay) ? ((ax > az) ? 0 : 2) : ((ay > az) ? 1 : 2);
    int kx = kz + 1 == 3 ? 0 : kz + 1;
    int ky = kx + 1 == 3 ? 0 : kx + 1;
    float dkz = sel3(r.d, kz);
    if (dkz < 0.0f) { int t = kx; kx = ky; ky = t; }
    r.kxyz = (uint32_t)kx | ((uint32_t)ky << 2) | ((uint32_t)kz << 4);
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
    // vertex - origin per axis, then the components in (kx, ky, kz) order (selecting component k
    // of the difference equals the difference of the k components)
    const uint32_t kx = r.kxyz & 3u, ky = (r.kxyz >> 2) & 3u, kz = r.kxyz >> 4;
    const float ax = a.x - r.o[0], ay = a.y - r.o[1], az = a.z - r.o[2];
    const float A_x = pick3(ax, ay, az, kx), A_y = pick3(ax, ay, az, ky), A_z = pick3(ax, ay, az, kz);
    const float bx = b.x - r.o[0], by = b.y - r.o[1], bz = b.z - r.o[2];
    const float B_x = pick3(bx, by, bz, kx), B_y = pick3(bx, by, bz, ky), B_z = pick3(bx, by, bz, kz);
    const float cx = c.x - r.o[0], cy = c.y - r.o[1], cz = c.z - r.o[2];
    const float C_x = pick3(cx, cy, cz, kx), C_y = pick3(cx, cy, cz, ky), C_z = pick3(cx, cy, cz, kz);
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

// Quantised planes as IEEE halves: byte q of a plane word becomes the low byte of a half whose high
// byte is 0x64, i.e. the half 1024 + q (exact). One byte permute makes the halves of two planes,
// and a fused multiply-add takes a half operand as it is (v_fma_mix_f32), so a plane costs half a
// permute and one FMA instead of a byte conversion and an FMA.
__device__ __forceinline__ uint32_t plane_halves(uint32_t w, int pair) {
    return __builtin_amdgcn_perm(0x64646464u, w, pair ? 0x07030702u : 0x07010700u);
}
__device__ __forceinline__ float plane_1024(uint32_t halves, int hi) {
    const uint16_t h = hi ? (uint16_t)(halves >> 16) : (uint16_t)(halves & 0xFFFFu);
    return (float)__builtin_bit_cast(_Float16, h);
}

// Tests the 8 children of a node; returns the hit mask in slot order. Empty slots hold the inverted
// box qlo = 255, qhi = 0 (and meta 0), so the box test itself rejects them; a spurious hit on one
// (possible only when every axis of the node has a vanishing scale) is a leaf slot of zero records
// and changes nothing.
__device__ __forceinline__ uint32_t intersect_node(const RaySetup &r, const uint4 w0, const uint4 w1, const uint4 w2,
                                                   const uint4 w3, const uint4 w4, float tmin, float tcull) {
    const float p[3] = {__uint_as_float(w0.x), __uint_as_float(w0.y), __uint_as_float(w0.z)};
    float a[3], bn[3], bf[3];
    for (int k = 0; k < 3; k++) {
        float s = __uint_as_float(((w0.w >> (8 * k)) & 0xFFu) << 23);
        a[k] = s * r.inv[k];
        float b = (p[k] - r.o[k]) * r.inv[k];
        // A plane's distance q a + b is computed as (1024 + q) a + (b - 1024 a): 1024 a is exact
        // (a power-of-two scale times inv), the subtraction and the FMA each round once. The
        // errors of b, of b - e - 1024 a and of the FMA stay below 8 units of 2^-24 (1024|a| +
        // |b|); the slack e = 2^-20 (1024|a| + |b|) (16 such units) keeps the test conservative.
        float e = 0x1p-20f * (1024.0f * fabsf(a[k]) + fabsf(b));
        const float shift = 1024.0f * a[k];
        bn[k] = (b - e) - shift;
        bf[k] = (b + e) - shift;
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
        const int pair = k >> 1, hi = k & 1;
        float tnx = __builtin_fmaf(plane_1024(plane_halves(wnx, pair), hi), a[0], bn[0]);
        float tny = __builtin_fmaf(plane_1024(plane_halves(wny, pair), hi), a[1], bn[1]);
        float tnz = __builtin_fmaf(plane_1024(plane_halves(wnz, pair), hi), a[2], bn[2]);
        float tfx = __builtin_fmaf(plane_1024(plane_halves(wfx, pair), hi), a[0], bf[0]);
        float tfy = __builtin_fmaf(plane_1024(plane_halves(wfy, pair), hi), a[1], bf[1]);
        float tfz = __builtin_fmaf(plane_1024(plane_halves(wfz, pair), hi), a[2], bf[2]);
        float tn = fmaxf(fmaxf(tnx, tny), fmaxf(tnz, tmin));
        float tf = fminf(fminf(tfx, tfy), fminf(tfz, tcull));
        if (tn <= tf) hits |= 1u << s;
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
constexpr uint32_t kTriGroup = 0x40000000u;  // stack entry tag: postponed triangles of a BLAS node

// Wave-uniform step choice (speculative traversal, Aila and Laine, "Understanding the Efficiency
// of Ray Traversal on GPUs", HPG 2009): a step tests triangles only when at least
// kTriQuarters/4 of the wave's live lanes hold triangles; otherwise it visits nodes, and lanes
// that hold triangles but still have a node group visit a node too (their triangles wait, and
// are pushed as a tagged stack entry if the new node brings leaves of its own). Each step then
// runs one of the two code paths instead of both.
constexpr bool kSpeculate = true;      // closest hit
constexpr bool kSpeculateAny = false;  // any hit (a lane tests its triangles first: a hit ends it)
constexpr bool kSpeculateNodes = true; // triangle-holding lanes visit nodes in node steps
constexpr int kTriQuarters = 2;
constexpr int kTriQuartersAny = 2;

// (t, inst, prim) is nearer than the current best, or as near with a smaller (inst, prim) key.
__device__ __forceinline__ bool nearer(float t, uint32_t inst, uint32_t prim, float t_best, uint32_t b_inst,
                                       uint32_t b_prim) {
    if (t != t_best) return t < t_best;
    return inst != b_inst ? inst < b_inst : prim < b_prim;
}

// One ray's traversal. ANY: stop at the first hit; TWO: TLAS + BLAS (otherwise the single
// instance of G); COUNT: count node visits and triangle tests. The stack's LDS part is the
// column `tid` of an [SLDS][BLOCK] array; entries beyond SLDS go to `spill`.
template<bool ANY, bool TWO, bool COUNT, int BLOCK, int SLDS>
struct Walker {
    uint2 (*stk)[BLOCK];
    uint32_t tid;
    uint2 *spill;
    uint32_t spill_depth;
    uint32_t *error;
    int sp;
    RaySetup R;           // current (object space when inside an instance)
    float wo[3], wd[3];   // world-space origin/direction (two-level)
    float tmin, tmax;
    uint32_t ray_mask;
    float t_best, bu, bv;
    uint32_t b_prim, b_inst;
    uint32_t b_rec;       // leaf record (TriRecord index in the hit instance's BLAS) of the best hit
    bool any_hit;
    uint32_t g_base, g_mask, t_base, t_mask;
    bool in_blas;
    int sp_inst;
    uint32_t cur_inst;
    const WideNode *nodes;
    const TriRecord *tris;
    // counting form only: node visits, triangle tests, loop iterations this lane spent live in
    // run(), and wave iterations (counted by the wave's first live lane)
    uint64_t n_nodes, n_tris, n_live, n_wave;

    __device__ __forceinline__ void init(const Geometry &G, uint2 (*s)[BLOCK], uint32_t t, uint2 *spill_lane,
                                         uint32_t depth, uint32_t *err) {
        stk = s; tid = t; spill = spill_lane; spill_depth = depth; error = err;
        sp = 0; n_nodes = 0; n_tris = 0; n_live = 0; n_wave = 0;
        // single-instance form: the BLAS pointers never change (kept uniform)
        in_blas = !TWO;
        nodes = TWO ? G.tlas_nodes : G.blas_nodes;
        tris = G.blas_tris;
        sp_inst = 0;
        cur_inst = 0;
    }

    // The stack never holds more than (wide depth - 1) node groups plus the TLAS leaf entries;
    // the host sizes spill_depth from the built depths. The bound checks only guard memory.
    __device__ __forceinline__ void push(uint2 e) {
        if (sp < SLDS) stk[sp][tid] = e;
        else if ((uint32_t)(sp - SLDS) < spill_depth) spill[sp - SLDS] = e;
        else atomicOr(error, 1u);
        sp++;
    }
    __device__ __forceinline__ uint2 pop() {
        sp--;
        if (sp < SLDS) return stk[sp][tid];
        if ((uint32_t)(sp - SLDS) < spill_depth) return spill[sp - SLDS];
        return make_uint2(0u, 0u);
    }

    __device__ __forceinline__ void start(const Geometry &G, float ox, float oy, float oz, float tmin_, float dx,
                                          float dy, float dz, float tmax_, uint32_t mask) {
        wo[0] = ox; wo[1] = oy; wo[2] = oz; tmin = tmin_;
        wd[0] = dx; wd[1] = dy; wd[2] = dz; tmax = tmax_;
        ray_mask = mask;
        t_best = FLT_MAX * 2.0f; // +inf
        b_prim = 0xFFFFFFFFu; b_inst = 0xFFFFFFFFu; bu = 0.0f; bv = 0.0f;
        any_hit = false;
        sp = 0;
        t_mask = 0;
        g_base = 0;
        if (TWO) {
            for (int k = 0; k < 3; k++) { R.o[k] = wo[k]; R.d[k] = wd[k]; }
            in_blas = false;
            nodes = G.tlas_nodes;
            ray_setup(R);
            g_mask = G.empty ? 0u : ((1u << R.oct) | (1u << 8));
            sp_inst = 0;
            cur_inst = 0;
        } else {
            xform_point(G.inst_inv, wo, R.o);
            xform_vector(G.inst_inv, wd, R.d);
            ray_setup(R);
            g_mask = (G.empty || (G.inst_mask & ray_mask & 0xFFu) == 0u) ? 0u : ((1u << R.oct) | (1u << 8));
        }
    }

    // Runs up to `steps` traversal steps; true when the ray is finished. Closest hit takes a
    // triangle step when 4 x (lanes holding triangles) >= tri_quarters x (live lanes) (0: the
    // default kTriQuarters); with
    // refill_below > 0 the wave stops early (false for every live lane) once fewer lanes than that
    // hold a ray, so that the caller can refill the idle ones (can_refill: slots are left).
    __device__ __forceinline__ bool run(const Geometry &G, int steps, uint32_t tri_quarters = kTriQuarters,
                                        uint32_t refill_below = 0u, bool can_refill = false) {
        constexpr bool SPEC = ANY ? kSpeculateAny : kSpeculate;
        for (int it = 0; it < steps; it++) {
            if (COUNT) {
                n_live++;
                if (__lane_id() == (uint32_t)__ffsll((long long)__ballot(1)) - 1u) n_wave++;
            }
            if (refill_below != 0u && can_refill && it > 0 && (uint32_t)__popcll(__ballot(1)) < refill_below)
                return false;
            const float tcull = ANY ? tmax : fminf(tmax, t_best);
            bool tri_step = t_mask != 0u;
            if (SPEC) {
                if (t_mask == 0u && (g_mask & 0xFFu) == 0u && sp == 0) return true;
                const uint32_t n_live = (uint32_t)__popcll(__ballot(1));
                const uint32_t n_tri = (uint32_t)__popcll(__ballot(t_mask != 0u));
                const uint32_t tq = ANY ? (uint32_t)kTriQuartersAny : (tri_quarters ? tri_quarters : (uint32_t)kTriQuarters);
                tri_step = 4u * n_tri >= tq * n_live;
                if (tri_step && t_mask == 0u) continue;
                if (!tri_step && !kSpeculateNodes && t_mask != 0u) continue;
            }
            if (tri_step) {
                const uint32_t b = (uint32_t)__builtin_ctz(t_mask);
                t_mask &= t_mask - 1u;
                if (!TWO || in_blas) {
                    const float4 *tr = reinterpret_cast<const float4 *>(tris + t_base + b);
                    const float4 a = tr[0], bb = tr[1], c = tr[2];
                    if (COUNT) n_tris++;
                    float t, u, v;
                    if (wt_intersect(R, a, bb, c, tmin, tmax, t, u, v)) {
                        if (ANY) { any_hit = true; return true; }
                        const uint32_t prim = __float_as_uint(a.w);
                        const uint32_t inst = TWO ? cur_inst : 0u;
                        if (nearer(t, inst, prim, t_best, b_inst, b_prim)) {
                            t_best = t; bu = u; bv = v;
                            b_prim = prim;
                            b_inst = inst;
                            b_rec = t_base + b;
                        }
                    }
                } else {
                    // TLAS leaf entry: enter instance t_base + b.
                    const uint32_t id = G.tlas_prims[t_base + b];
                    const InstanceRecord *ir = G.instances + id;
                    const uint32_t imask = ir->vis_mask;
                    const uint64_t bn = ir->blas_nodes;
                    // procedural primitives report only to ray queries (plain traces run without
                    // intersection functions, cpu/accel.rs:476-484)
                    if (bn != 0u && (imask & ray_mask & 0xFFu) != 0u && !(ir->flags & kInstProcedural)) {
                        if (t_mask != 0u) push(make_uint2(t_base, t_mask | kInstGroup));
                        t_mask = 0u;
                        if ((g_mask & 0xFFu) != 0u) push(make_uint2(g_base, g_mask));
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
                if (SPEC && t_mask != 0u) continue; // holds triangles and has no node to visit
                if (TWO && in_blas && sp == sp_inst) {
                    // leave the instance: back to the world-space ray and the TLAS
                    for (int k = 0; k < 3; k++) { R.o[k] = wo[k]; R.d[k] = wd[k]; }
                    ray_setup(R);
                    nodes = G.tlas_nodes;
                    in_blas = false;
                }
                if (sp == 0) return true;
                const uint2 e = pop();
                if (TWO && (e.y & kInstGroup)) {
                    t_base = e.x;
                    t_mask = e.y & ~kInstGroup;
                    continue;
                }
                if (SPEC && (e.y & kTriGroup)) {
                    t_base = e.x;
                    t_mask = e.y & ~kTriGroup;
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
            if (hm != 0u) push(make_uint2(g_base, hm | (im << 8)));
            g_mask = 0u;
            uint4 w0, w1, w2, w3, w4;
            load_node(nodes, node, w0, w1, w2, w3, w4);
            if (COUNT) n_nodes++;
            const uint32_t hits = intersect_node(R, w0, w1, w2, w3, w4, tmin, tcull);
            const uint32_t nim = w0.w >> 24;
            g_base = w1.x;
            g_mask = permute8(hits & nim, R.oct) | (nim << 8);
            uint32_t leaf = hits & ~nim;
            uint32_t nt = 0u;
            while (leaf) {
                const uint32_t s = (uint32_t)__builtin_ctz(leaf);
                leaf &= leaf - 1u;
                const uint32_t m = ((s < 4 ? w1.z : w1.w) >> (8 * (s & 3))) & 0xFFu;
                nt |= ((1u << (m >> 5)) - 1u) << (m & 31u);
            }
            if (SPEC) {
                if (nt != 0u) {
                    // the postponed group goes to the stack (instances in the TLAS, triangles in a BLAS)
                    if (t_mask != 0u) push(make_uint2(t_base, t_mask | ((TWO && !in_blas) ? kInstGroup : kTriGroup)));
                    t_base = w1.y;
                    t_mask = nt;
                }
            } else {
                t_base = w1.y;
                t_mask = nt;
            }
        }
        return false;
    }

    // Closest-hit record (24-B SurfaceHit as three uint2). The padding word after t carries the
    // leaf record index when `record` is set (queues internal to the path tracer), else 0.
    __device__ __forceinline__ void write_hit(uint2 *h, bool record) const {
        const bool hit = b_prim != 0xFFFFFFFFu;
        const uint32_t t_bits = hit ? __float_as_uint(t_best) : __float_as_uint(tmax);
        h[0] = make_uint2(hit ? b_inst : 0xFFFFFFFFu, b_prim);
        h[1] = make_uint2(hit ? __float_as_uint(bu) : 0u, hit ? __float_as_uint(bv) : 0u);
        h[2] = make_uint2(t_bits, (hit && record) ? b_rec : 0u);
    }
};

} // namespace trav
} // namespace lchip
