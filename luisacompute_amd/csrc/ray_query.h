// ray_query.h -- ray queries over the BVH8 layout: traverse / traverse_any with surface and
// procedural candidate callbacks (include/luisa/dsl/rtx/ray_query.h:11-243), for HIP kernels that
// include this header and pass their callbacks as a policy object.
//
// Semantics follow the reference CPU backend's AccelImpl::ray_query
// (src/rust/luisa_compute_backend_impl/src/cpu/accel.rs:582-800):
//   - every triangle the ray hits with t in [tmin, tmax] in an instance whose mask passes is a
//     surface candidate; a candidate of an opaque instance is committed without the callback
//     (the callback is disabled for opaque instances, cpu/accel.rs:378-393), otherwise the
//     callback decides;
//   - every procedural primitive whose AABB the ray reaches is a procedural candidate; the
//     callback returns the distance it commits, which counts only inside [tmin, tmax);
//   - traverse: the result is the nearest committed candidate; traverse_any (terminate on first):
//     the first committed candidate ends the query;
//   - result: CommittedHit {inst, prim, bary, hit_type, committed_ray_t}
//     (include/luisa/runtime/rtx/hit.h:18-26); a miss reports inst = prim = 0xFFFFFFFF,
//     bary = (0, 0), hit_type 0, t = ray.tmax; a procedural hit reports bary = (0, 0).
// Callbacks that are functions of the candidate alone make the nearest-commit result independent
// of the visiting order; ties at equal t go to the smaller (instance, primitive) pair, as in
// trace_closest.
//
// The walk is one ray per lane. Nearest-commit queries take wave-uniform steps (leaf records or
// nodes, as the closest-hit walker of traverse.h); first-commit queries let each lane finish a
// node's leaf records before its next node (a commit ends them). The walk is resumable
// (QueryWalker::run advances a bounded number of steps), so persistent kernels refill the lanes
// whose query finished, as the trace kernel does. Procedural BLASes hold one
// ProcRecord per leaf slot; the callback receives the object-space ray (and the world-space one),
// so geometry inside the AABB does not depend on how conservatively boxes are culled.
#pragma once

#include "lc_hip_bvh.h"
#include "traverse.h"

namespace lchip {
namespace rq {

constexpr uint32_t kHitMiss = 0u, kHitSurface = 1u, kHitProcedural = 2u;

struct CommittedHit {
    uint32_t inst, prim;
    float u, v;
    uint32_t type;
    float t;
};

// Per-lane stack: the first SLDS entries in column `tid` of an [SLDS][BLOCK] LDS array, the rest
// in a global spill area of `depth` entries (sized by the host from the built depths; the bound
// check only guards memory and flags an error).
template<int BLOCK, int SLDS>
struct LaneStack {
    uint2 (*lds)[BLOCK];
    uint32_t tid;
    uint2 *spill;
    uint32_t depth;
    uint32_t *error;
    int sp;
    __device__ __forceinline__ void push(uint2 e) {
        if constexpr (SLDS > 0) {
            if (sp < SLDS) {
                lds[sp][tid] = e;
                sp++;
                return;
            }
        }
        if ((uint32_t)(sp - SLDS) < depth) spill[sp - SLDS] = e;
        else atomicOr(error, 1u);
        sp++;
    }
    __device__ __forceinline__ uint2 pop() {
        sp--;
        if constexpr (SLDS > 0) {
            if (sp < SLDS) return lds[sp][tid];
        }
        if ((uint32_t)(sp - SLDS) < depth) return spill[sp - SLDS];
        return make_uint2(0u, 0u);
    }
};

// One query as a resumable walk: start() takes the ray, run() advances it by up to `steps`
// steps and returns true when it is finished, hit holds the answer. Policy provides
//   bool surface(uint32_t inst, uint32_t prim, float u, float v, float t) const
//   bool procedural(uint32_t inst, uint32_t prim, const float obj_o[3], const float obj_d[3],
//                   const float world_o[3], const float world_d[3], float4 lo, float4 hi,
//                   float tmax, float &t) const
// (`tmax`: the current committed distance, as candidate.ray().t_max() reports it).
template<class Policy, int BLOCK, int SLDS>
struct QueryWalker {
    LaneStack<BLOCK, SLDS> S;
    trav::RaySetup R;
    float wo[3], wd[3];
    float tmin, tmax, t_best;
    uint32_t ray_mask;
    bool first, in_blas;
    CommittedHit hit;
    const WideNode *nodes;
    const char *recs;
    uint32_t cur_inst, cur_flags;
    int sp_inst;
    uint32_t g_base, g_mask, t_base, t_mask;

    // enters instance `id` if it has geometry and its mask passes (pushing what the TLAS walk
    // still holds); false otherwise
    __device__ __forceinline__ bool enter(const QueryGeometry &Q, uint32_t id) {
        const InstanceRecord *ir = Q.instances + id;
        const uint64_t bn = ir->blas_nodes;
        if (bn == 0u || (ir->vis_mask & ray_mask & 0xFFu) == 0u) return false;
        if (t_mask != 0u) S.push(make_uint2(t_base, t_mask | trav::kInstGroup));
        t_mask = 0u;
        if ((g_mask & 0xFFu) != 0u) S.push(make_uint2(g_base, g_mask));
        trav::xform_point(ir->inv, wo, R.o);
        trav::xform_vector(ir->inv, wd, R.d);
        trav::ray_setup(R);
        nodes = reinterpret_cast<const WideNode *>(bn);
        recs = reinterpret_cast<const char *>(ir->blas_tris);
        cur_inst = id;
        cur_flags = ir->flags;
        in_blas = true;
        sp_inst = S.sp;
        g_base = 0;
        g_mask = (1u << R.oct) | (1u << 8);
        return true;
    }

    __device__ __forceinline__ void start(const QueryGeometry &Q, const float4 r0, const float4 r1, uint32_t mask,
                                          bool first_) {
        wo[0] = r0.x; wo[1] = r0.y; wo[2] = r0.z; tmin = r0.w;
        wd[0] = r1.x; wd[1] = r1.y; wd[2] = r1.z; tmax = r1.w;
        ray_mask = mask;
        first = first_;
        hit = CommittedHit{0xFFFFFFFFu, 0xFFFFFFFFu, 0.0f, 0.0f, kHitMiss, tmax};
        t_best = FLT_MAX * 2.0f; // +inf
        S.sp = 0;
        in_blas = false;
        sp_inst = 0;
        g_base = 0; g_mask = 0; t_base = 0; t_mask = 0;
        nodes = Q.tlas_nodes;
        recs = nullptr;
        cur_inst = 0; cur_flags = 0;
        if (Q.empty) return; // g_mask = t_mask = 0, sp = 0: finished at the first step
        if (Q.two) {
            for (int k = 0; k < 3; k++) { R.o[k] = wo[k]; R.d[k] = wd[k]; }
            trav::ray_setup(R);
            g_mask = (1u << R.oct) | (1u << 8);
        } else {
            enter(Q, 0u);
        }
    }

    __device__ __forceinline__ bool run(const QueryGeometry &Q, const Policy &pol, int steps) {
        for (int it = 0; it < steps; it++) {
            bool leaf_step = t_mask != 0u;
            if (!first) {
                // nearest-commit queries take wave-uniform steps like the closest-hit walker
                // (traverse.h): leaf records when at least half of the live lanes hold some,
                // otherwise nodes (lanes holding records visit a node too and push them)
                if (t_mask == 0u && (g_mask & 0xFFu) == 0u && S.sp == 0) return true;
                const uint32_t n_live = (uint32_t)__popcll(__ballot(1));
                const uint32_t n_leaf = (uint32_t)__popcll(__ballot(t_mask != 0u));
                leaf_step = 2u * n_leaf >= n_live;
                if (leaf_step && t_mask == 0u) continue;
            }
            if (leaf_step) {
                const uint32_t b = (uint32_t)__builtin_ctz(t_mask);
                t_mask &= t_mask - 1u;
                if (!in_blas) {
                    enter(Q, Q.tlas_prims[t_base + b]);
                    continue;
                }
                const uint32_t rec = t_base + b;
                if (cur_flags & kInstProcedural) {
                    const float4 *pr = reinterpret_cast<const float4 *>(recs + 32ull * rec);
                    const float4 lo = pr[0], hi = pr[1];
                    const uint32_t prim = __float_as_uint(lo.w);
                    float t = 0.0f;
                    const float t_far = fminf(tmax, t_best);
                    if (pol.procedural(cur_inst, prim, R.o, R.d, wo, wd, lo, hi, t_far, t) && t >= tmin && t < tmax &&
                        trav::nearer(t, cur_inst, prim, t_best, hit.inst, hit.prim)) {
                        t_best = t;
                        hit = CommittedHit{cur_inst, prim, 0.0f, 0.0f, kHitProcedural, t};
                        if (first) return true;
                    }
                } else {
                    const float4 *tr = reinterpret_cast<const float4 *>(recs + 48ull * rec);
                    const float4 a = tr[0], bb = tr[1], c = tr[2];
                    float t, u, v;
                    if (trav::wt_intersect(R, a, bb, c, tmin, tmax, t, u, v)) {
                        const uint32_t prim = __float_as_uint(a.w);
                        if (trav::nearer(t, cur_inst, prim, t_best, hit.inst, hit.prim) &&
                            ((cur_flags & kInstOpaque) || pol.surface(cur_inst, prim, u, v, t))) {
                            t_best = t;
                            hit = CommittedHit{cur_inst, prim, u, v, kHitSurface, t};
                            if (first) return true;
                        }
                    }
                }
                continue;
            }
            if ((g_mask & 0xFFu) == 0u) {
                if (t_mask != 0u) continue; // holds leaf records and has no node to visit
                if (in_blas && S.sp == sp_inst) {
                    if (!Q.two) return true;
                    // leave the instance: back to the world-space ray and the TLAS
                    for (int k = 0; k < 3; k++) { R.o[k] = wo[k]; R.d[k] = wd[k]; }
                    trav::ray_setup(R);
                    nodes = Q.tlas_nodes;
                    in_blas = false;
                }
                if (S.sp == 0) return true;
                const uint2 e = S.pop();
                if (e.y & trav::kInstGroup) {
                    t_base = e.x;
                    t_mask = e.y & ~trav::kInstGroup;
                    continue;
                }
                if (e.y & trav::kTriGroup) {
                    t_base = e.x;
                    t_mask = e.y & ~trav::kTriGroup;
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
            if (hm != 0u) S.push(make_uint2(g_base, hm | (im << 8)));
            uint4 w0, w1, w2, w3, w4;
            trav::load_node(nodes, node, w0, w1, w2, w3, w4);
            const float tcull = first ? tmax : fminf(tmax, t_best);
            const uint32_t hits = trav::intersect_node(R, w0, w1, w2, w3, w4, tmin, tcull);
            const uint32_t nim = w0.w >> 24;
            g_base = w1.x;
            g_mask = trav::permute8(hits & nim, R.oct) | (nim << 8);
            uint32_t leaf = hits & ~nim;
            uint32_t nt = 0u;
            while (leaf) {
                const uint32_t s = (uint32_t)__builtin_ctz(leaf);
                leaf &= leaf - 1u;
                const uint32_t m = ((s < 4 ? w1.z : w1.w) >> (8 * (s & 3))) & 0xFFu;
                nt |= ((1u << (m >> 5)) - 1u) << (m & 31u);
            }
            if (nt != 0u) {
                // records still held go to the stack (instances in the TLAS, leaf records in a BLAS)
                if (t_mask != 0u) S.push(make_uint2(t_base, t_mask | (in_blas ? trav::kTriGroup : trav::kInstGroup)));
                t_base = w1.y;
                t_mask = nt;
            }
        }
        return false;
    }
};

// The whole query in one call (for kernels that trace one query per lane).
template<class Policy, int BLOCK, int SLDS>
__device__ CommittedHit ray_query(const QueryGeometry &Q, LaneStack<BLOCK, SLDS> &S, const float4 r0, const float4 r1,
                                  uint32_t ray_mask, bool first, const Policy &pol) {
    QueryWalker<Policy, BLOCK, SLDS> W;
    W.S = S;
    W.start(Q, r0, r1, ray_mask, first);
    while (!W.run(Q, pol, 1 << 30)) {
    }
    return W.hit;
}

} // namespace rq
} // namespace lchip
