// bvh_build.hip -- GPU BVH construction for gfx950: Morton LBVH -> greedy 8-wide collapse with
// conservatively quantised child boxes.
//
// Replaces the reference's OptiX BLAS/IAS build (src/backends/cuda/cuda_primitive.cpp:20-113,
// cuda_accel.cpp:26-126) and Embree's rtcCommitScene (src/rust/luisa_compute_backend_impl/src/
// cpu/accel.rs:205-260). Neither reference path has a software builder; the algorithm here is:
//   1. centroid bounds (wave64 shuffle reduction + one atomic per block),
//   2. 63-bit Morton codes, radix sort (rocPRIM, keys + primitive ids),
//   3. binary radix tree and its boxes in one bottom-up pass (Apetrei, "Fast and Simple
//      Agglomerative LBVH Construction", CGVC 2014) with acquire/release handoff between the two
//      children of each node,
//   4. top-down collapse to 8-wide nodes, one launch per wide level with device-side work
//      counts (no host round trip between levels); each node greedily opens its widest child
//      until it has 8 children, assigns children to octant slots, quantises child boxes to
//      8 bits per plane and copies its leaf triangles into a contiguous block.
// The same builder serves the TLAS (primitives = instance world boxes).
#include <hip/hip_runtime.h>
#include <cstring>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_select.hpp>

#include <cfloat>
#include <cstdint>

#include "lc_hip_bvh.h"
#include "lc_hip_internal.h"

namespace lchip {

namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ int f2ord(float f) {
    int i = __float_as_int(f);
    return i >= 0 ? i : i ^ 0x7FFFFFFF;
}
__device__ __forceinline__ float ord2f(int i) { return __int_as_float(i >= 0 ? i : i ^ 0x7FFFFFFF); }

struct PrimSource {
    // Triangles: vertex buffer (16-B aligned stride) + index triples.
    const char *vbase;
    uint64_t vstride;
    uint64_t nverts;
    const uint32_t *tri; // 3 per triangle
    // Boxes (TLAS): 2 float4 per primitive {lo.xyz, -}, {hi.xyz, -}. Null for triangles.
    const float4 *boxes;
    uint32_t *error;     // bit 1: a triangle index was out of range (clamped)
};

__device__ __forceinline__ float4 load_vertex(const PrimSource &s, uint32_t i) {
    if (i >= s.nverts) {
        atomicOr(s.error, 2u);
        i = (uint32_t)(s.nverts - 1);
    }
    return *reinterpret_cast<const float4 *>(s.vbase + (uint64_t)i * s.vstride);
}

__device__ __forceinline__ void prim_box(const PrimSource &s, uint32_t p, float3 &lo, float3 &hi) {
    if (s.boxes) {
        float4 a = s.boxes[2 * (uint64_t)p], b = s.boxes[2 * (uint64_t)p + 1];
        lo = make_float3(a.x, a.y, a.z);
        hi = make_float3(b.x, b.y, b.z);
        return;
    }
    uint32_t i0 = s.tri[3 * (uint64_t)p], i1 = s.tri[3 * (uint64_t)p + 1], i2 = s.tri[3 * (uint64_t)p + 2];
    float4 a = load_vertex(s, i0), b = load_vertex(s, i1), c = load_vertex(s, i2);
    lo = make_float3(fminf(a.x, fminf(b.x, c.x)), fminf(a.y, fminf(b.y, c.y)), fminf(a.z, fminf(b.z, c.z)));
    hi = make_float3(fmaxf(a.x, fmaxf(b.x, c.x)), fmaxf(a.y, fmaxf(b.y, c.y)), fmaxf(a.z, fmaxf(b.z, c.z)));
}

__global__ void k_init(BuildCounters *ctr, uint32_t *slots, uint32_t nslots) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) {
        ctr->node_count = 1;
        ctr->tri_count = 0;
        ctr->depth = 0;
        ctr->root_ref = 0;
        for (int a = 0; a < 3; a++) {
            ctr->cbounds[a] = f2ord(FLT_MAX);
            ctr->cbounds[3 + a] = f2ord(-FLT_MAX);
            ctr->bounds[a] = 0.0f;
            ctr->bounds[3 + a] = 0.0f;
        }
    }
    if (i < kMaxCollapseLevels) ctr->level_count[i] = 0;
    for (uint32_t k = i; k < nslots; k += gridDim.x * blockDim.x) slots[k] = 0u;
}

__device__ __forceinline__ float wave_min(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ float wave_max(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}

__global__ __launch_bounds__(kBlock) void k_centroid_bounds(PrimSource src, uint32_t n, BuildCounters *ctr) {
    float c[6] = {FLT_MAX, FLT_MAX, FLT_MAX, -FLT_MAX, -FLT_MAX, -FLT_MAX};
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) {
        float3 lo, hi;
        prim_box(src, p, lo, hi);
        float cx = 0.5f * (lo.x + hi.x), cy = 0.5f * (lo.y + hi.y), cz = 0.5f * (lo.z + hi.z);
        c[0] = fminf(c[0], cx); c[1] = fminf(c[1], cy); c[2] = fminf(c[2], cz);
        c[3] = fmaxf(c[3], cx); c[4] = fmaxf(c[4], cy); c[5] = fmaxf(c[5], cz);
    }
    __shared__ float part[kBlock / 64][6];
    int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int a = 0; a < 3; a++) {
        float mn = wave_min(c[a]), mx = wave_max(c[3 + a]);
        if (lane == 0) { part[wave][a] = mn; part[wave][3 + a] = mx; }
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        int a = threadIdx.x;
        float v = part[0][a];
        for (int w = 1; w < kBlock / 64; w++) v = a < 3 ? fminf(v, part[w][a]) : fmaxf(v, part[w][a]);
        if (a < 3) atomicMin(&ctr->cbounds[a], f2ord(v));
        else atomicMax(&ctr->cbounds[a], f2ord(v));
    }
}

__device__ __forceinline__ uint64_t expand21(uint32_t v) {
    uint64_t x = v & 0x1fffffu;
    x = (x | x << 32) & 0x1f00000000ffffull;
    x = (x | x << 16) & 0x1f0000ff0000ffull;
    x = (x | x << 8) & 0x100f00f00f00f00full;
    x = (x | x << 4) & 0x10c30c30c30c30c3ull;
    x = (x | x << 2) & 0x1249249249249249ull;
    return x;
}

__global__ __launch_bounds__(kBlock) void k_morton(PrimSource src, uint32_t n, const BuildCounters *ctr,
                                                   uint64_t *keys, uint32_t *ids) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    float lo[3], sc[3];
    for (int a = 0; a < 3; a++) {
        float l = ord2f(ctr->cbounds[a]), h = ord2f(ctr->cbounds[3 + a]);
        float ext = h - l;
        lo[a] = l;
        sc[a] = ext > 0.0f ? 2097152.0f / ext : 0.0f;
    }
    float3 blo, bhi;
    prim_box(src, p, blo, bhi);
    float c[3] = {0.5f * (blo.x + bhi.x), 0.5f * (blo.y + bhi.y), 0.5f * (blo.z + bhi.z)};
    uint32_t q[3];
    for (int a = 0; a < 3; a++) {
        float f = (c[a] - lo[a]) * sc[a];
        f = fminf(fmaxf(f, 0.0f), 2097151.0f);
        q[a] = (uint32_t)f;
    }
    keys[p] = expand21(q[0]) << 2 | expand21(q[1]) << 1 | expand21(q[2]);
    ids[p] = p;
}

// Common-prefix length of sorted keys i and j (-1 outside [0, n)); equal keys continue with the
// common prefix of the indices, so every pair of positions is ordered (Karras 2012, sec. 4).
__device__ __forceinline__ int prefix_len(const uint64_t *keys, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    uint64_t a = keys[i], b = keys[j];
    if (a != b) return __clzll((long long)(a ^ b));
    return 64 + __clz(i ^ j);
}

struct BinaryTree {
    float4 *cbox;          // 4 per internal node: left lo, left hi, right lo, right hi
    uint2 *child;          // per internal node: left ref, right ref
    uint2 *range;          // per internal node: first, last sorted position
    uint32_t *slots;       // handoff word per internal node (0 = no child arrived yet)
    uint32_t *parent_int;  // parent of internal node i
    uint32_t *parent_leaf; // parent of leaf (sorted position) i
};

// Binary radix tree in one parallel pass: internal node i finds its key range by exponential and
// binary search over prefix lengths and splits it at the highest differing bit (Karras,
// "Maximizing Parallelism in the Construction of BVHs, Octrees, and k-d Trees", HPG 2012).
// No synchronisation: every node depends on the sorted keys only. Root = internal node 0.
__global__ __launch_bounds__(kBlock) void k_karras(uint32_t n, const uint64_t *keys, BinaryTree bt) {
    const int N = (int)n;
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= N - 1) return;
    const int d = prefix_len(keys, N, i, i + 1) > prefix_len(keys, N, i, i - 1) ? 1 : -1;
    const int dmin = prefix_len(keys, N, i, i - d);
    int lmax = 2;
    while (prefix_len(keys, N, i, i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (prefix_len(keys, N, i, i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = prefix_len(keys, N, i, j);
    int s = 0, t = l;
    do {
        t = (t + 1) >> 1;
        if (prefix_len(keys, N, i, i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    const int gamma = i + s * d + min(d, 0);
    const int lo = min(i, j), hi = max(i, j);
    const uint32_t left = lo == gamma ? ((uint32_t)gamma | kLeafBit) : (uint32_t)gamma;
    const uint32_t right = hi == gamma + 1 ? ((uint32_t)(gamma + 1) | kLeafBit) : (uint32_t)(gamma + 1);
    bt.child[i] = make_uint2(left, right);
    bt.range[i] = make_uint2((uint32_t)lo, (uint32_t)hi);
    if (left & kLeafBit) bt.parent_leaf[gamma] = (uint32_t)i; else bt.parent_int[gamma] = (uint32_t)i;
    if (right & kLeafBit) bt.parent_leaf[gamma + 1] = (uint32_t)i; else bt.parent_int[gamma + 1] = (uint32_t)i;
}

__device__ __forceinline__ uint32_t ref_first(const BinaryTree &bt, uint32_t r) {
    return (r & kLeafBit) ? (r & ~kLeafBit) : bt.range[r].x;
}
__device__ __forceinline__ uint32_t ref_last(const BinaryTree &bt, uint32_t r) {
    return (r & kLeafBit) ? (r & ~kLeafBit) : bt.range[r].y;
}

__device__ __forceinline__ void publish_root(BuildCounters *ctr, uint2 *queue0, uint32_t root, float3 lo, float3 hi) {
    ctr->root_ref = root;
    ctr->bounds[0] = lo.x; ctr->bounds[1] = lo.y; ctr->bounds[2] = lo.z;
    ctr->bounds[3] = hi.x; ctr->bounds[4] = hi.y; ctr->bounds[5] = hi.z;
    queue0[0] = make_uint2(root, 0u);
    ctr->level_count[0] = 1;
}

constexpr uint32_t kNoEntry = 0xFFFFFFFFu;

// Bottom-up boxes, one phase. The workgroup owns entries [e0, e1) -- complete subtrees ordered by
// first leaf -- and so the leaf range [L0, L1]. A thread climbs from its entry while the parent's
// key range lies inside [L0, L1]; then both children of that parent are climbed by threads of the
// same workgroup, and the handoff (the first child to arrive stops, the second merges the boxes
// and goes on) needs only workgroup-scope ordering. A parent whose range leaves the workgroup is
// deferred: the child is written to pend[first leaf] and becomes an entry of the next phase.
// whole != 0: the workgroup owns all leaves (last phase, one workgroup) or, with agent scope, the
// fallback for inputs whose phases stop shrinking.
template<bool AGENT>
__global__ void k_bounds_phase(PrimSource src, uint32_t n, const uint32_t *ids, BinaryTree bt,
                               const uint32_t *entries, uint32_t m, uint32_t *pend, BuildCounters *ctr,
                               uint2 *queue0, uint32_t whole) {
    const uint32_t e0 = blockIdx.x * blockDim.x;
    if (e0 >= m) return;
    const uint32_t e1 = min(e0 + blockDim.x, m);
    uint32_t L0 = 0, L1 = n - 1;
    if (!whole) {
        L0 = ref_first(bt, entries ? entries[e0] : (e0 | kLeafBit));
        L1 = ref_last(bt, entries ? entries[e1 - 1] : ((e1 - 1) | kLeafBit));
    }
    const uint32_t e = e0 + threadIdx.x;
    if (e >= e1) return;
    uint32_t cur = entries ? entries[e] : (e | kLeafBit);
    float3 lo, hi;
    if (cur & kLeafBit) {
        prim_box(src, ids[cur & ~kLeafBit], lo, hi);
    } else {
        const float4 *cb = bt.cbox + 4 * (uint64_t)cur;
        float4 a = cb[0], b = cb[1], c = cb[2], d = cb[3];
        lo = make_float3(fminf(a.x, c.x), fminf(a.y, c.y), fminf(a.z, c.z));
        hi = make_float3(fmaxf(b.x, d.x), fmaxf(b.y, d.y), fmaxf(b.z, d.z));
    }
    for (;;) {
        if (cur == 0u) { // internal node 0 spans every leaf: the root
            publish_root(ctr, queue0, 0u, lo, hi);
            return;
        }
        const uint32_t p = (cur & kLeafBit) ? bt.parent_leaf[cur & ~kLeafBit] : bt.parent_int[cur];
        const uint2 pr = bt.range[p];
        if (pr.x < L0 || pr.y > L1) {
            pend[ref_first(bt, cur)] = cur;
            return;
        }
        const uint32_t side = bt.child[p].x == cur ? 0u : 1u;
        bt.cbox[4 * (uint64_t)p + 2 * side] = make_float4(lo.x, lo.y, lo.z, 0.0f);
        bt.cbox[4 * (uint64_t)p + 2 * side + 1] = make_float4(hi.x, hi.y, hi.z, 0.0f);
        const uint32_t first = AGENT
            ? __hip_atomic_exchange(&bt.slots[p], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT)
            : __hip_atomic_exchange(&bt.slots[p], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (first == 0u) return; // the sibling's thread completes p
        const float4 slo = bt.cbox[4 * (uint64_t)p + 2 * (1u - side)];
        const float4 shi = bt.cbox[4 * (uint64_t)p + 2 * (1u - side) + 1];
        lo = make_float3(fminf(lo.x, slo.x), fminf(lo.y, slo.y), fminf(lo.z, slo.z));
        hi = make_float3(fmaxf(hi.x, shi.x), fmaxf(hi.y, shi.y), fmaxf(hi.z, shi.z));
        cur = p;
    }
}

__global__ void k_single_leaf_root(PrimSource src, const uint32_t *ids, BuildCounters *ctr, uint2 *queue0) {
    float3 lo, hi;
    prim_box(src, ids[0], lo, hi);
    publish_root(ctr, queue0, kLeafBit, lo, hi);
}

__global__ void k_fill_u32(uint32_t *p, uint64_t n, uint32_t v) {
    for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < n; i += (uint64_t)gridDim.x * blockDim.x) p[i] = v;
}

// Clears the pend slots of the entries just compacted out of it.
__global__ void k_clear_pend(const uint32_t *entries, const uint32_t *count, uint32_t *pend, BinaryTree bt) {
    const uint32_t m = *count;
    for (uint32_t e = blockIdx.x * blockDim.x + threadIdx.x; e < m; e += gridDim.x * blockDim.x)
        pend[ref_first(bt, entries[e])] = kNoEntry;
}

struct IsEntry {
    __device__ bool operator()(uint32_t v) const { return v != kNoEntry; }
};

struct WideOut {
    WideNode *nodes;
    TriRecord *tris;     // triangle BLAS: leaf triangle records
    uint32_t *prim_out;  // box BVH (TLAS): leaf primitive ids instead of triangle records
};

__device__ __forceinline__ float box_area(float3 lo, float3 hi) {
    float dx = hi.x - lo.x, dy = hi.y - lo.y, dz = hi.z - lo.z;
    return dx * dy + dy * dz + dz * dx;
}

// Smallest power-of-two scale (as a biased exponent) with p + 255 * 2^(e-127) >= hi, checked in
// double so the decoded range is exact.
__device__ __forceinline__ uint32_t pick_exponent(float lo, float hi) {
    float ext = hi - lo;
    uint32_t e = 1;
    if (ext > 0.0f) {
        int ex;
        frexpf(ext / 255.0f, &ex); // ext/255 = m * 2^ex, m in [0.5, 1)
        int b = ex + 127;
        e = (uint32_t)max(1, min(254, b));
    }
    while (e < 254 && (double)lo + 255.0 * (double)__int_as_float((int)(e << 23)) < (double)hi) e++;
    return e;
}

__device__ __forceinline__ uint32_t quant_lo(float v, float p, float s) {
    float q = floorf((v - p) / s);
    q = fminf(fmaxf(q, 0.0f), 255.0f);
    uint32_t qi = (uint32_t)q;
    while (qi > 0 && (double)p + (double)qi * (double)s > (double)v) qi--;
    return qi;
}

__device__ __forceinline__ uint32_t quant_hi(float v, float p, float s) {
    float q = ceilf((v - p) / s);
    q = fminf(fmaxf(q, 0.0f), 255.0f);
    uint32_t qi = (uint32_t)q;
    while (qi < 255 && (double)p + (double)qi * (double)s < (double)v) qi++;
    return qi;
}

// One wide level: each work item (binary ref, wide slot) becomes one WideNode.
__global__ __launch_bounds__(kBlock) void k_collapse(PrimSource src, const uint32_t *ids, BinaryTree bt,
                                                     BuildCounters *ctr, uint32_t level, const uint2 *qin,
                                                     uint2 *qout, WideOut out, uint32_t split_small) {
    const uint32_t count = ctr->level_count[level];
    if (count == 0) return;
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicMax(&ctr->depth, level + 1);
    for (uint32_t w = blockIdx.x * blockDim.x + threadIdx.x; w < count; w += gridDim.x * blockDim.x) {
        uint2 item = qin[w];
        uint32_t ref[8], cnt[8];
        float3 clo[8], chi[8];
        int nc;
        auto ref_count = [&](uint32_t r) -> uint32_t {
            if (r & kLeafBit) return 1u;
            uint2 rg = bt.range[r];
            return rg.y - rg.x + 1u;
        };
        if (item.x & kLeafBit) {
            nc = 1;
            ref[0] = item.x;
            cnt[0] = 1;
            prim_box(src, ids[item.x & ~kLeafBit], clo[0], chi[0]);
        } else {
            nc = 2;
            uint2 ch = bt.child[item.x];
            const float4 *cb = bt.cbox + 4 * (uint64_t)item.x;
            float4 a = cb[0], b = cb[1], c = cb[2], d = cb[3];
            ref[0] = ch.x; clo[0] = make_float3(a.x, a.y, a.z); chi[0] = make_float3(b.x, b.y, b.z);
            ref[1] = ch.y; clo[1] = make_float3(c.x, c.y, c.z); chi[1] = make_float3(d.x, d.y, d.z);
            cnt[0] = ref_count(ch.x);
            cnt[1] = ref_count(ch.y);
        }
        // Greedy opening: widest child that must become internal, then (optionally) widest
        // multi-triangle leaf candidate while slots remain.
        while (nc < 8) {
            int best = -1;
            float best_area = -1.0f;
            for (int j = 0; j < nc; j++) {
                if (cnt[j] > kLeafMax) {
                    float a = box_area(clo[j], chi[j]);
                    if (a > best_area) { best_area = a; best = j; }
                }
            }
            if (best < 0 && split_small) {
                for (int j = 0; j < nc; j++) {
                    if (!(ref[j] & kLeafBit) && cnt[j] > 1) {
                        float a = box_area(clo[j], chi[j]);
                        if (a > best_area) { best_area = a; best = j; }
                    }
                }
            }
            if (best < 0) break;
            uint32_t r = ref[best];
            uint2 ch = bt.child[r];
            const float4 *cb = bt.cbox + 4 * (uint64_t)r;
            float4 a = cb[0], b = cb[1], c = cb[2], d = cb[3];
            ref[best] = ch.x; clo[best] = make_float3(a.x, a.y, a.z); chi[best] = make_float3(b.x, b.y, b.z);
            cnt[best] = ref_count(ch.x);
            ref[nc] = ch.y; clo[nc] = make_float3(c.x, c.y, c.z); chi[nc] = make_float3(d.x, d.y, d.z);
            cnt[nc] = ref_count(ch.y);
            nc++;
        }
        // Node box, origin and scale.
        float3 nlo = clo[0], nhi = chi[0];
        for (int j = 1; j < nc; j++) {
            nlo = make_float3(fminf(nlo.x, clo[j].x), fminf(nlo.y, clo[j].y), fminf(nlo.z, clo[j].z));
            nhi = make_float3(fmaxf(nhi.x, chi[j].x), fmaxf(nhi.y, chi[j].y), fmaxf(nhi.z, chi[j].z));
        }
        uint32_t ex = pick_exponent(nlo.x, nhi.x), ey = pick_exponent(nlo.y, nhi.y), ez = pick_exponent(nlo.z, nhi.z);
        float sx = __int_as_float((int)(ex << 23)), sy = __int_as_float((int)(ey << 23)), sz = __int_as_float((int)(ez << 23));
        // Octant slot assignment: greedy max of dot(child centre - node centre, octant direction).
        float3 pc = make_float3(0.5f * (nlo.x + nhi.x), 0.5f * (nlo.y + nhi.y), 0.5f * (nlo.z + nhi.z));
        int slot_of[8];
        uint32_t used_slots = 0, used_children = 0;
        for (int k = 0; k < nc; k++) {
            float best = -FLT_MAX;
            int bc = 0, bs = 0;
            for (int j = 0; j < nc; j++) {
                if (used_children & (1u << j)) continue;
                float dx = 0.5f * (clo[j].x + chi[j].x) - pc.x;
                float dy = 0.5f * (clo[j].y + chi[j].y) - pc.y;
                float dz = 0.5f * (clo[j].z + chi[j].z) - pc.z;
                for (int s = 0; s < 8; s++) {
                    if (used_slots & (1u << s)) continue;
                    float v = ((s & 1) ? dx : -dx) + ((s & 2) ? dy : -dy) + ((s & 4) ? dz : -dz);
                    if (v > best) { best = v; bc = j; bs = s; }
                }
            }
            used_children |= 1u << bc;
            used_slots |= 1u << bs;
            slot_of[bc] = bs;
        }
        int child_at[8];
        for (int s = 0; s < 8; s++) child_at[s] = -1;
        for (int j = 0; j < nc; j++) child_at[slot_of[j]] = j;
        // Allocation of internal children and leaf triangles.
        uint32_t imask = 0, ninternal = 0, nleaf_tris = 0;
        for (int s = 0; s < 8; s++) {
            int j = child_at[s];
            if (j < 0) continue;
            if (cnt[j] > kLeafMax) { imask |= 1u << s; ninternal++; }
            else nleaf_tris += cnt[j];
        }
        uint32_t child_base = ninternal ? atomicAdd(&ctr->node_count, ninternal) : 0u;
        uint32_t tri_base = nleaf_tris ? atomicAdd(&ctr->tri_count, nleaf_tris) : 0u;
        uint32_t qbase = ninternal ? atomicAdd(&ctr->level_count[level + 1], ninternal) : 0u;
        uint32_t meta[2] = {0, 0};
        uint32_t qlo[3][2] = {{0, 0}, {0, 0}, {0, 0}}, qhi[3][2] = {{0, 0}, {0, 0}, {0, 0}};
        uint32_t rank = 0, toff = 0;
        for (int s = 0; s < 8; s++) {
            int j = child_at[s];
            uint32_t sh = 8u * (uint32_t)(s & 3), wi = (uint32_t)(s >> 2);
            uint32_t lx, ly, lz, hx, hy, hz, m;
            if (j < 0) {
                lx = ly = lz = 255u;
                hx = hy = hz = 0u;
                m = 0u;
            } else {
                lx = quant_lo(clo[j].x, nlo.x, sx); hx = quant_hi(chi[j].x, nlo.x, sx);
                ly = quant_lo(clo[j].y, nlo.y, sy); hy = quant_hi(chi[j].y, nlo.y, sy);
                lz = quant_lo(clo[j].z, nlo.z, sz); hz = quant_hi(chi[j].z, nlo.z, sz);
                if (imask & (1u << s)) {
                    m = 0xFFu;
                    qout[qbase + rank] = make_uint2(ref[j], child_base + rank);
                    rank++;
                } else {
                    uint32_t first = (ref[j] & kLeafBit) ? (ref[j] & ~kLeafBit) : bt.range[ref[j]].x;
                    m = (cnt[j] << 5) | toff;
                    for (uint32_t k = 0; k < cnt[j]; k++) {
                        uint32_t prim = ids[first + k];
                        uint32_t dst = tri_base + toff + k;
                        if (out.tris) {
                            uint32_t i0 = src.tri[3 * (uint64_t)prim], i1 = src.tri[3 * (uint64_t)prim + 1], i2 = src.tri[3 * (uint64_t)prim + 2];
                            float4 a = load_vertex(src, i0), b = load_vertex(src, i1), c = load_vertex(src, i2);
                            float4 *tr = reinterpret_cast<float4 *>(out.tris + dst);
                            tr[0] = make_float4(a.x, a.y, a.z, __uint_as_float(prim));
                            tr[1] = make_float4(b.x, b.y, b.z, 0.0f);
                            tr[2] = make_float4(c.x, c.y, c.z, 0.0f);
                        } else {
                            out.prim_out[dst] = prim;
                        }
                    }
                    toff += cnt[j];
                }
            }
            meta[wi] |= m << sh;
            qlo[0][wi] |= lx << sh; qlo[1][wi] |= ly << sh; qlo[2][wi] |= lz << sh;
            qhi[0][wi] |= hx << sh; qhi[1][wi] |= hy << sh; qhi[2][wi] |= hz << sh;
        }
        uint4 *dst = reinterpret_cast<uint4 *>(out.nodes + item.y);
        dst[0] = make_uint4(__float_as_uint(nlo.x), __float_as_uint(nlo.y), __float_as_uint(nlo.z),
                            ex | (ey << 8) | (ez << 16) | (imask << 24));
        dst[1] = make_uint4(child_base, tri_base, meta[0], meta[1]);
        dst[2] = make_uint4(qlo[0][0], qlo[0][1], qlo[1][0], qlo[1][1]);
        dst[3] = make_uint4(qlo[2][0], qlo[2][1], qhi[0][0], qhi[0][1]);
        dst[4] = make_uint4(qhi[1][0], qhi[1][1], qhi[2][0], qhi[2][1]);
    }
}

template<typename T>
T *carve(char *&p, size_t count) {
    uintptr_t a = (reinterpret_cast<uintptr_t>(p) + 255) & ~uintptr_t(255);
    T *r = reinterpret_cast<T *>(a);
    p = reinterpret_cast<char *>(a + count * sizeof(T));
    return r;
}

} // namespace

// Scratch bytes for building over n primitives (everything but the output arrays).
static size_t sort_temp_bytes(uint32_t n) {
    size_t bytes = 0;
    rocprim::double_buffer<uint64_t> k(nullptr, nullptr);
    rocprim::double_buffer<uint32_t> v(nullptr, nullptr);
    rocprim::radix_sort_pairs(nullptr, bytes, k, v, n, 0, 63, (hipStream_t)0);
    return bytes;
}

static size_t select_temp_bytes(uint32_t n) {
    size_t bytes = 0;
    rocprim::select(nullptr, bytes, (const uint32_t *)nullptr, (uint32_t *)nullptr, (uint32_t *)nullptr, n, IsEntry{},
                    (hipStream_t)0);
    return bytes;
}

size_t bvh_scratch_bytes(uint32_t n) {
    size_t m = std::max<uint32_t>(n, 2u);
    size_t b = 0;
    b += 2 * m * sizeof(uint64_t) + 256 * 2;        // keys
    b += 2 * m * sizeof(uint32_t) + 256 * 2;        // ids
    b += 4 * m * sizeof(float4) + 256;              // cbox
    b += m * sizeof(uint2) + 256;                   // child
    b += m * sizeof(uint2) + 256;                   // range
    b += m * sizeof(uint32_t) + 256;                // slots
    b += 2 * m * sizeof(uint32_t) + 256 * 2;        // parents
    b += 3 * m * sizeof(uint32_t) + 256 * 3;        // pend + two entry lists
    b += 2 * m * sizeof(uint2) + 256 * 2;           // collapse queues
    b += sizeof(BuildCounters) + 256;               // counters
    b += sizeof(uint32_t) + 256;                    // selected count
    b += std::max(sort_temp_bytes((uint32_t)m), select_temp_bytes((uint32_t)m)) + 256; // rocPRIM temp
    return b;
}

size_t bvh_node_capacity(uint32_t n) { return std::max<uint32_t>(n, 2u); }

// Builds a wide BVH over n primitives (triangles if boxes == nullptr, else boxes).
// Output arrays: nodes (bvh_node_capacity(n) entries), tris (n entries) or prim_out (n entries).
// The counters land in `counters_out` (device) for the caller to read back. `pinned_word` is a
// host-pinned word used to read the per-phase entry counts (the stream is synchronized once per
// bottom-up phase; about log_64(n) phases).
hipError_t bvh_build(hipStream_t stream, void *scratch, size_t scratch_bytes, const void *vbase, uint64_t vstride,
                     uint64_t nverts, const uint32_t *tri, const float4 *boxes, uint32_t n, WideNode *nodes,
                     TriRecord *tris, uint32_t *prim_out, BuildCounters **counters_out, uint32_t split_small,
                     int grid_collapse, uint32_t *error_flag, uint32_t *pinned_word) {
    if (scratch_bytes < bvh_scratch_bytes(n)) return hipErrorInvalidValue;
    size_t m = std::max<uint32_t>(n, 2u);
    char *p = static_cast<char *>(scratch);
    uint64_t *k0 = carve<uint64_t>(p, m), *k1 = carve<uint64_t>(p, m);
    uint32_t *i0 = carve<uint32_t>(p, m), *i1 = carve<uint32_t>(p, m);
    BinaryTree bt;
    bt.cbox = carve<float4>(p, 4 * m);
    bt.child = carve<uint2>(p, m);
    bt.range = carve<uint2>(p, m);
    bt.slots = carve<uint32_t>(p, m);
    bt.parent_int = carve<uint32_t>(p, m);
    bt.parent_leaf = carve<uint32_t>(p, m);
    uint32_t *pend = carve<uint32_t>(p, m);
    uint32_t *ent[2] = {carve<uint32_t>(p, m), carve<uint32_t>(p, m)};
    uint2 *q0 = carve<uint2>(p, m), *q1 = carve<uint2>(p, m);
    BuildCounters *ctr = carve<BuildCounters>(p, 1);
    uint32_t *dcount = carve<uint32_t>(p, 1);
    size_t temp_bytes = std::max(sort_temp_bytes((uint32_t)m), select_temp_bytes((uint32_t)m));
    void *temp = carve<char>(p, temp_bytes);
    *counters_out = ctr;

    if (!boxes && n > 0 && nverts == 0) return hipErrorInvalidValue;
    PrimSource src{static_cast<const char *>(vbase), vstride, nverts, tri, boxes, error_flag};
    const uint32_t nslots = n > 1 ? n - 1 : 0;
    hipLaunchKernelGGL(k_init, dim3(std::max<uint32_t>(1, std::min<uint32_t>((nslots + kBlock - 1) / kBlock, 4096))),
                       dim3(kBlock), 0, stream, ctr, bt.slots, nslots);
    if (n == 0) return hipGetLastError();
    uint32_t grid = (n + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_centroid_bounds, dim3(std::min<uint32_t>(grid, 2048)), dim3(kBlock), 0, stream, src, n, ctr);
    hipLaunchKernelGGL(k_morton, dim3(grid), dim3(kBlock), 0, stream, src, n, ctr, k0, i0);
    rocprim::double_buffer<uint64_t> kb(k0, k1);
    rocprim::double_buffer<uint32_t> vb(i0, i1);
    size_t tb = temp_bytes;
    hipError_t e = rocprim::radix_sort_pairs(temp, tb, kb, vb, n, 0, 63, stream);
    if (e != hipSuccess) return e;
    const uint64_t *keys = kb.current();
    const uint32_t *ids = vb.current();
    if (n == 1) {
        hipLaunchKernelGGL(k_single_leaf_root, dim3(1), dim3(1), 0, stream, src, ids, ctr, q0);
    } else {
        hipLaunchKernelGGL(k_karras, dim3((n - 1 + kBlock - 1) / kBlock), dim3(kBlock), 0, stream, n, keys, bt);
        hipLaunchKernelGGL(k_fill_u32, dim3(std::min<uint32_t>(grid, 4096)), dim3(kBlock), 0, stream, pend, (uint64_t)n,
                           kNoEntry);
        // Bottom-up phases. Entries of phase 0 are the leaves in sorted order (entries == null).
        const uint32_t kPhaseBlock = 256, kLastBlock = 1024;
        const uint32_t *entries = nullptr;
        uint32_t count = n;
        int which = 0;
        for (;;) {
            if (count <= kLastBlock) {
                uint32_t blk = (count + 63) / 64 * 64;
                hipLaunchKernelGGL(k_bounds_phase<false>, dim3(1), dim3(blk), 0, stream, src, n, ids, bt, entries, count,
                                   pend, ctr, q0, 1u);
                break;
            }
            hipLaunchKernelGGL(k_bounds_phase<false>, dim3((count + kPhaseBlock - 1) / kPhaseBlock), dim3(kPhaseBlock), 0,
                               stream, src, n, ids, bt, entries, count, pend, ctr, q0, 0u);
            tb = temp_bytes;
            e = rocprim::select(temp, tb, (const uint32_t *)pend, ent[which], dcount, n, IsEntry{}, stream);
            if (e != hipSuccess) return e;
            hipLaunchKernelGGL(k_clear_pend, dim3(256), dim3(kBlock), 0, stream, ent[which], dcount, pend, bt);
            e = hipMemcpyAsync(pinned_word, dcount, sizeof(uint32_t), hipMemcpyDeviceToHost, stream);
            if (e != hipSuccess) return e;
            e = hipStreamSynchronize(stream);
            if (e != hipSuccess) return e;
            const uint32_t next = *pinned_word;
            entries = ent[which];
            which ^= 1;
            if (next == 0) break; // the root was reached in this phase
            if (next > count - count / 8) {
                // Phases stopped shrinking (sibling subtrees keep straddling workgroup
                // boundaries): finish with agent-scope handoffs over all remaining entries.
                hipLaunchKernelGGL(k_bounds_phase<true>, dim3((next + kPhaseBlock - 1) / kPhaseBlock), dim3(kPhaseBlock),
                                   0, stream, src, n, ids, bt, entries, next, pend, ctr, q0, 1u);
                break;
            }
            count = next;
        }
    }
    WideOut out{nodes, tris, prim_out};
    // A root-to-leaf path of the radix tree crosses each of the 63 key bits and 32 index bits
    // at most once, and every wide level descends at least one binary level, so the wide depth
    // is at most 96: launching kCollapseLaunches levels always completes the tree (levels
    // without work return at once). The host checks level_count[kCollapseLaunches] == 0.
    for (uint32_t level = 0; level < kCollapseLaunches; level++) {
        const uint2 *qin = (level & 1) ? q1 : q0;
        uint2 *qout = (level & 1) ? q0 : q1;
        hipLaunchKernelGGL(k_collapse, dim3(grid_collapse), dim3(kBlock), 0, stream, src, ids, bt, ctr, level, qin,
                           qout, out, split_small);
    }
    return hipGetLastError();
}

} // namespace lchip
