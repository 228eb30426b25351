// bvh_build.hip -- GPU BVH construction for gfx950: Morton LBVH -> 8-wide collapse with
// conservatively quantised child boxes.
//
// Replaces the reference's OptiX BLAS/IAS build (src/backends/cuda/cuda_primitive.cpp:20-113,
// cuda_accel.cpp:26-126) and Embree's rtcCommitScene (src/rust/luisa_compute_backend_impl/src/
// cpu/accel.rs:205-260). Neither reference path has a software builder; the steps here are:
//   1. centroid bounds (wave64 shuffle reduction + one atomic per block); primitive boxes are
//      stored once, in input order, by the Morton kernel;
//   2. 30-bit Morton codes (10 bits per axis), radix sort of (code, primitive id) (rocPRIM);
//   3. binary radix tree from the sorted codes alone (Karras 2012), equal codes ordered by index;
//   4. node boxes bottom-up in phases whose child handoffs stay inside one workgroup (so they
//      only need workgroup-scope ordering -- an agent-scope release/acquire writes back and
//      invalidates the XCD-local L2 and made a single-pass climb 9x slower than the whole build);
//      the first phase, over all leaves, hands off through LDS;
//   5. top-down collapse to 8-wide nodes, one launch per wide level, 8 lanes per node: each lane
//      holds one child, the group opens its widest child until it has 8, assigns children to
//      octant slots, quantises each child box to 8 bits per plane and lists leaf primitives;
//   6. triangle records in leaf order, one thread per record.
// The same builder serves the TLAS (primitives = instance world boxes).
#include <hip/hip_runtime.h>
#include <cstring>

#include <cfloat>
#include <cstdint>
#include <cstdlib>

#include "lc_hip_bvh.h"
#include "lc_hip_internal.h"

namespace lchip {

namespace {

constexpr int kBlock = 256;
// Collapse workgroup: groups of 8 lanes, one queue/triangle reservation per workgroup iteration.
// collapse workgroup (8 lanes per wide node; 512 measured best: 4.39 ms builds vs 4.90 at 256 and
// 4.62 at 1024, the workgroup's one reservation atomic per pass against occupancy) and the refit's
constexpr int kCollapseBlock = 512;
constexpr int kRefitBlock = 256;
constexpr uint32_t kNoEntry = 0xFFFFFFFFu;
// bottom-up phases: entries per workgroup, the single-workgroup last phase, leaves per run of
// the first phase
constexpr uint32_t kPhaseBlock = 256, kLastBlock = 1024, kLeafRun = 512;
static_assert(kLastBlock >= kLeafRun && kLeafRun >= kPhaseBlock, "phase sizes");

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

// Input box of primitive p (box builds: the TLAS's instance boxes).
__device__ __forceinline__ void stored_box(const float4 *pbox, uint32_t p, float3 &lo, float3 &hi) {
    float4 a = pbox[2 * (uint64_t)p], b = pbox[2 * (uint64_t)p + 1];
    lo = make_float3(a.x, a.y, a.z);
    hi = make_float3(b.x, b.y, b.z);
}

// Leaf boxes by sorted position: triangle builds read the records written in sorted order
// (srec, coalesced with the sort order), box builds the input boxes through the sorted ids.
struct LeafBoxes {
    const float4 *pbox;
    const uint32_t *ids;
    const TriRecord *srec;
};

__device__ __forceinline__ void record_box(const TriRecord *r, float3 &lo, float3 &hi) {
    const float4 *v = reinterpret_cast<const float4 *>(r);
    const float4 a = v[0], b = v[1], c = v[2];
    lo = make_float3(fminf(a.x, fminf(b.x, c.x)), fminf(a.y, fminf(b.y, c.y)), fminf(a.z, fminf(b.z, c.z)));
    hi = make_float3(fmaxf(a.x, fmaxf(b.x, c.x)), fmaxf(a.y, fmaxf(b.y, c.y)), fmaxf(a.z, fmaxf(b.z, c.z)));
}

__device__ __forceinline__ void leaf_box(const LeafBoxes &lb, uint32_t pos, float3 &lo, float3 &hi) {
    if (lb.srec) record_box(lb.srec + pos, lo, hi);
    else stored_box(lb.pbox, lb.ids[pos], lo, hi);
}

// Copies record *src to *dst; returns its box.
__device__ __forceinline__ void copy_record(const TriRecord *src, TriRecord *dst, float3 &lo, float3 &hi) {
    const float4 *v = reinterpret_cast<const float4 *>(src);
    const float4 a = v[0], b = v[1], c = v[2];
    float4 *d = reinterpret_cast<float4 *>(dst);
    d[0] = a;
    d[1] = b;
    d[2] = c;
    lo = make_float3(fminf(a.x, fminf(b.x, c.x)), fminf(a.y, fminf(b.y, c.y)), fminf(a.z, fminf(b.z, c.z)));
    hi = make_float3(fmaxf(a.x, fmaxf(b.x, c.x)), fmaxf(a.y, fmaxf(b.y, c.y)), fmaxf(a.z, fmaxf(b.z, c.z)));
}

// Triangle record of primitive `prim` (prim id in v0.w) at *dst; returns its box.
__device__ __forceinline__ void write_record(const PrimSource &src, uint32_t prim, TriRecord *dst, float3 &lo,
                                             float3 &hi) {
    const uint32_t i0 = src.tri[3 * (uint64_t)prim], i1 = src.tri[3 * (uint64_t)prim + 1],
                   i2 = src.tri[3 * (uint64_t)prim + 2];
    const float4 a = load_vertex(src, i0), b = load_vertex(src, i1), c = load_vertex(src, i2);
    float4 *tr = reinterpret_cast<float4 *>(dst);
    tr[0] = make_float4(a.x, a.y, a.z, __uint_as_float(prim));
    tr[1] = make_float4(b.x, b.y, b.z, 0.0f);
    tr[2] = make_float4(c.x, c.y, c.z, 0.0f);
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
    if (i < kMaxCollapseLevels) ctr->level[i] = 0ull;
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

__device__ __forceinline__ uint32_t expand10(uint32_t v) {
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

// 30-bit Morton code of the box centroid in the centroid bounds; a triangle build also writes
// each triangle's record in primitive order (prec), read back in sorted order by the leaf phase.
__global__ __launch_bounds__(kBlock) void k_morton(PrimSource src, uint32_t n, const BuildCounters *ctr,
                                                   uint32_t *keys, TriRecord *prec) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    float lo[3], sc[3];
    for (int a = 0; a < 3; a++) {
        float l = ord2f(ctr->cbounds[a]), h = ord2f(ctr->cbounds[3 + a]);
        float ext = h - l;
        lo[a] = l;
        sc[a] = ext > 0.0f ? 1024.0f / ext : 0.0f;
    }
    float3 blo, bhi;
    if (prec) write_record(src, p, prec + p, blo, bhi);
    else prim_box(src, p, blo, bhi);
    float c[3] = {0.5f * (blo.x + bhi.x), 0.5f * (blo.y + bhi.y), 0.5f * (blo.z + bhi.z)};
    uint32_t q[3];
    for (int a = 0; a < 3; a++) {
        float f = (c[a] - lo[a]) * sc[a];
        f = fminf(fmaxf(f, 0.0f), 1023.0f);
        q[a] = (uint32_t)f;
    }
    keys[p] = expand10(q[0]) << 2 | expand10(q[1]) << 1 | expand10(q[2]);
}

struct BinaryTree {
    float4 *cbox;          // 4 per internal node: left lo, left hi, right lo, right hi
    uint2 *child;          // per internal node: left ref, right ref
    uint2 *range;          // per internal node: first, last sorted position
    uint32_t *slots;       // handoff word per internal node (0 = no child arrived yet)
    // Parent links of internal node i / leaf (sorted position) i: (parent, e | side << 31) where
    // side is 1 for a right child and e is the parent's range end the child does not share (its
    // last leaf for a left child, its first for a right one). A climbing thread that knows its
    // node's range [f, l] gets the parent's range [f, e] or [e, l] from this one load.
    uint2 *parent_int;
    uint2 *parent_leaf;
};

__device__ __forceinline__ uint2 parent_link(uint32_t parent, uint32_t end, uint32_t side) {
    return make_uint2(parent, end | (side << 31));
}

// Binary radix tree in one parallel pass: internal node i finds its key range by exponential and
// binary search over prefix lengths and splits it at the highest differing bit (Karras,
// "Maximizing Parallelism in the Construction of BVHs, Octrees, and k-d Trees", HPG 2012).
// No synchronisation: every node depends on the sorted keys only. Root = internal node 0.
// Most nodes span a few keys, so the workgroup first copies the keys of its nodes and kBlock more
// on each side into LDS; the searches read keys there and go to memory only beyond that window.
__global__ __launch_bounds__(kBlock) void k_karras(uint32_t n, const uint32_t *keys, BinaryTree bt) {
    __shared__ uint32_t s_keys[3 * kBlock];
    const int N = (int)n;
    const int w0 = (int)(blockIdx.x * blockDim.x) - (int)kBlock; // position of s_keys[0]
    for (int k = (int)threadIdx.x; k < 3 * (int)kBlock; k += (int)blockDim.x) {
        const int g = w0 + k;
        s_keys[k] = (g >= 0 && g < N) ? keys[g] : 0u;
    }
    __syncthreads();
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= N - 1) return;
    const uint32_t ki = s_keys[i - w0];
    // common-prefix length of sorted keys i and j (-1 outside [0, N)); equal keys continue with
    // the common prefix of the positions, so every pair of positions is ordered (Karras 2012,
    // sec. 4). The key of j comes from the window when it lies there.
    const auto plen = [&](int j) -> int {
        if (j < 0 || j >= N) return -1;
        const int k = j - w0;
        const uint32_t kj = (k >= 0 && k < 3 * (int)kBlock) ? s_keys[k] : keys[j];
        if (ki != kj) return __clz((int)(ki ^ kj));
        return 32 + __clz(i ^ j);
    };
    const int d = plen(i + 1) > plen(i - 1) ? 1 : -1;
    const int dmin = plen(i - d);
    int lmax = 2;
    while (plen(i + lmax * d) > dmin) lmax <<= 1;
    int l = 0;
    for (int t = lmax >> 1; t >= 1; t >>= 1)
        if (plen(i + (l + t) * d) > dmin) l += t;
    const int j = i + l * d;
    const int dnode = plen(j);
    int s = 0, t = l;
    do {
        t = (t + 1) >> 1;
        if (plen(i + (s + t) * d) > dnode) s += t;
    } while (t > 1);
    const int gamma = i + s * d + min(d, 0);
    const int lo = min(i, j), hi = max(i, j);
    const uint32_t left = lo == gamma ? ((uint32_t)gamma | kLeafBit) : (uint32_t)gamma;
    const uint32_t right = hi == gamma + 1 ? ((uint32_t)(gamma + 1) | kLeafBit) : (uint32_t)(gamma + 1);
    bt.child[i] = make_uint2(left, right);
    bt.range[i] = make_uint2((uint32_t)lo, (uint32_t)hi);
    const uint2 lp = parent_link((uint32_t)i, (uint32_t)hi, 0u), rp = parent_link((uint32_t)i, (uint32_t)lo, 1u);
    if (left & kLeafBit) bt.parent_leaf[gamma] = lp; else bt.parent_int[gamma] = lp;
    if (right & kLeafBit) bt.parent_leaf[gamma + 1] = rp; else bt.parent_int[gamma + 1] = rp;
}

__device__ __forceinline__ uint32_t ref_first(const BinaryTree &bt, uint32_t r) {
    return (r & kLeafBit) ? (r & ~kLeafBit) : bt.range[r].x;
}
__device__ __forceinline__ uint32_t ref_last(const BinaryTree &bt, uint32_t r) {
    return (r & kLeafBit) ? (r & ~kLeafBit) : bt.range[r].y;
}
__device__ __forceinline__ uint32_t ref_count(const BinaryTree &bt, uint32_t r) {
    if (r & kLeafBit) return 1u;
    const uint2 rg = bt.range[r];
    return rg.y - rg.x + 1u;
}

__device__ __forceinline__ void publish_root(BuildCounters *ctr, uint2 *queue0, uint32_t root, float3 lo, float3 hi) {
    ctr->root_ref = root;
    ctr->bounds[0] = lo.x; ctr->bounds[1] = lo.y; ctr->bounds[2] = lo.z;
    ctr->bounds[3] = hi.x; ctr->bounds[4] = hi.y; ctr->bounds[5] = hi.z;
    queue0[0] = make_uint2(root, 0u);
    ctr->level[0] = 1ull << 32;
}

// Deferred entries of a phase, compacted per workgroup in thread order (which is first-leaf
// order: a thread's final subtree contains its starting entry): out[blockIdx.x * blockDim.x + r],
// count in bcount[blockIdx.x]. Every thread of the workgroup calls it once.
__device__ __forceinline__ void compact_deferred(uint32_t deferred, uint32_t *out, uint32_t *bcount,
                                                 uint32_t chunk = 0xFFFFFFFFu) {
    if (chunk == 0xFFFFFFFFu) chunk = blockIdx.x;
    __shared__ uint32_t s_wave_count[16];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const bool has = deferred != kNoEntry;
    const uint64_t mask = __ballot(has);
    const uint32_t rank = (uint32_t)__popcll(mask & ((1ull << lane) - 1ull));
    if (lane == 0u) s_wave_count[wave] = (uint32_t)__popcll(mask);
    __syncthreads();
    uint32_t base = 0, total = 0;
    for (uint32_t w = 0; w < ((blockDim.x + 63u) >> 6); w++) {
        if (w < wave) base += s_wave_count[w];
        total += s_wave_count[w];
    }
    if (has) out[chunk * blockDim.x + base + rank] = deferred;
    if (threadIdx.x == 0u) bcount[chunk] = total;
}

// One thread's climb from entry `cur` (a leaf ref or an internal node): boxes go up while the
// parent's key range lies inside [L0, L1]; the first child to arrive at a parent stops, the second
// merges the sibling's box and goes on. Returns the entry to defer (kNoEntry: stopped at a
// parent its sibling completes, or published the root).
template<bool AGENT>
__device__ __forceinline__ uint32_t climb(const LeafBoxes &lb, const BinaryTree &bt, uint32_t cur, uint32_t L0,
                                          uint32_t L1, BuildCounters *ctr, uint2 *queue0) {
    float3 lo, hi;
    uint32_t f, l;
    if (cur & kLeafBit) {
        f = l = cur & ~kLeafBit;
        leaf_box(lb, f, lo, hi);
    } else {
        const uint2 rg = bt.range[cur];
        f = rg.x;
        l = rg.y;
        const float4 *cb = bt.cbox + 4 * (uint64_t)cur;
        float4 a = cb[0], b = cb[1], c = cb[2], d = cb[3];
        lo = make_float3(fminf(a.x, c.x), fminf(a.y, c.y), fminf(a.z, c.z));
        hi = make_float3(fmaxf(b.x, d.x), fmaxf(b.y, d.y), fmaxf(b.z, d.z));
    }
    for (;;) {
        if (cur == 0u) { // internal node 0 spans every leaf: the root
            publish_root(ctr, queue0, 0u, lo, hi);
            return kNoEntry;
        }
        const uint2 link = (cur & kLeafBit) ? bt.parent_leaf[f] : bt.parent_int[cur];
        const uint32_t p = link.x, side = link.y >> 31, end = link.y & ~kLeafBit;
        const uint32_t pf = side ? end : f, pl = side ? l : end;
        if (pf < L0 || pl > L1) return cur;
        bt.cbox[4 * (uint64_t)p + 2 * side] = make_float4(lo.x, lo.y, lo.z, 0.0f);
        bt.cbox[4 * (uint64_t)p + 2 * side + 1] = make_float4(hi.x, hi.y, hi.z, 0.0f);
        const uint32_t first = AGENT
            ? __hip_atomic_exchange(&bt.slots[p], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT)
            : __hip_atomic_exchange(&bt.slots[p], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (first == 0u) return kNoEntry; // the sibling's thread completes p
        const float4 slo = bt.cbox[4 * (uint64_t)p + 2 * (1u - side)];
        const float4 shi = bt.cbox[4 * (uint64_t)p + 2 * (1u - side) + 1];
        lo = make_float3(fminf(lo.x, slo.x), fminf(lo.y, slo.y), fminf(lo.z, slo.z));
        hi = make_float3(fmaxf(hi.x, shi.x), fmaxf(hi.y, shi.y), fmaxf(hi.z, shi.z));
        cur = p;
        f = pf;
        l = pl;
    }
}

// Bottom-up boxes, one phase. The workgroup owns entries [e0, e1) -- complete subtrees ordered by
// first leaf -- and so the leaf range [L0, L1]. A thread climbs from its entry while the parent's
// key range lies inside [L0, L1]; then both children of that parent are climbed by threads of the
// same workgroup, and the handoff needs only workgroup-scope ordering. A parent whose range leaves
// the workgroup is deferred: the child becomes an entry of the next pass (compact_deferred into
// `pend`). whole != 0: the workgroup owns all leaves (a build of at most kLastBlock primitives,
// one workgroup); nothing is deferred then.
__global__ void k_bounds_phase(LeafBoxes lb, uint32_t n, BinaryTree bt, const uint32_t *entries, uint32_t m,
                               uint32_t *pend, uint32_t *bcount, BuildCounters *ctr, uint2 *queue0, uint32_t whole) {
    const uint32_t e0 = blockIdx.x * blockDim.x;
    if (e0 >= m) return; // whole workgroups only
    const uint32_t e1 = min(e0 + blockDim.x, m);
    uint32_t L0 = 0, L1 = n - 1;
    if (!whole) {
        L0 = ref_first(bt, entries ? entries[e0] : (e0 | kLeafBit));
        L1 = ref_last(bt, entries ? entries[e1 - 1] : ((e1 - 1) | kLeafBit));
    }
    const uint32_t e = e0 + threadIdx.x;
    uint32_t deferred = kNoEntry;
    if (e < e1) deferred = climb<false>(lb, bt, entries ? entries[e] : (e | kLeafBit), L0, L1, ctr, queue0);
    if (!whole) compact_deferred(deferred, pend, bcount);
}

// The bottom-up passes after the first run over entry lists whose length a kernel left in device
// memory (`count`), so the host never reads it. k_bounds_phase_dev: a phase as k_bounds_phase,
// workgroup w taking the chunks of kPhaseBlock entries w, w + gridDim.x, ... (chunk c's deferred
// entries go to pend[c * kPhaseBlock ...], their number to bcount[c]). k_bounds_last: one
// workgroup climbs from every remaining entry to the root (threads stride over the entries; the
// handoffs stay in the workgroup). k_gather_dev: the ordered compaction of a phase's deferred
// entries into the next list and its count.
__global__ __launch_bounds__(kPhaseBlock) void k_bounds_phase_dev(LeafBoxes lb, uint32_t n, BinaryTree bt,
                                                                  const uint32_t *entries, const uint32_t *count,
                                                                  uint32_t *pend, uint32_t *bcount, BuildCounters *ctr,
                                                                  uint2 *queue0) {
    const uint32_t m = *count;
    const uint32_t chunks = (m + kPhaseBlock - 1u) / kPhaseBlock;
    for (uint32_t c = blockIdx.x; c < chunks; c += gridDim.x) {
        const uint32_t e0 = c * kPhaseBlock, e1 = min(e0 + kPhaseBlock, m);
        const uint32_t L0 = ref_first(bt, entries[e0]), L1 = ref_last(bt, entries[e1 - 1u]);
        const uint32_t e = e0 + threadIdx.x;
        uint32_t deferred = kNoEntry;
        if (e < e1) deferred = climb<false>(lb, bt, entries[e], L0, L1, ctr, queue0);
        compact_deferred(deferred, pend, bcount, c);
        __syncthreads(); // compact_deferred's LDS counts are rewritten by the next chunk
    }
}

__global__ __launch_bounds__(kLastBlock) void k_bounds_last(LeafBoxes lb, uint32_t n, BinaryTree bt,
                                                            const uint32_t *entries, const uint32_t *count,
                                                            BuildCounters *ctr, uint2 *queue0) {
    const uint32_t m = *count;
    for (uint32_t e = threadIdx.x; e < m; e += blockDim.x) climb<false>(lb, bt, entries[e], 0u, n - 1u, ctr, queue0);
}

__global__ __launch_bounds__(kPhaseBlock) void k_gather_dev(const uint32_t *src, const uint32_t *bcount,
                                                            const uint32_t *boff, const uint32_t *count,
                                                            uint32_t *dst, uint32_t *total) {
    const uint32_t chunks = (*count + kPhaseBlock - 1u) / kPhaseBlock;
    if (chunks == 0u && blockIdx.x == 0u && threadIdx.x == 0u) *total = 0u;
    for (uint32_t c = blockIdx.x; c < chunks; c += gridDim.x) {
        const uint32_t k = bcount[c], o = boff[c];
        for (uint32_t t = threadIdx.x; t < k; t += blockDim.x) dst[o + t] = src[(uint64_t)c * kPhaseBlock + t];
        if (c == chunks - 1u && threadIdx.x == 0u) *total = o + k;
    }
}

// First bottom-up phase: workgroup b owns leaves [b B, b B + B) and climbs from every leaf. A node
// whose range lies inside the run has its index inside the run too (a radix-tree node's index is
// one end of its range), so both children of such a node hand off through LDS: each writes its
// box to its side's LDS slot, then an LDS atomic picks the second arrival, which merges and goes
// on. Boxes are also stored to cbox for the collapse. Parents reaching outside the run are
// deferred as in k_bounds_phase. Requires n > B (the root never lies inside a run). A triangle
// build (srec != null) also writes each leaf's triangle record at its sorted position.
template<int B>
__global__ __launch_bounds__(B) void k_bounds_leaves(LeafBoxes lb, const TriRecord *prec, TriRecord *srec, uint32_t n,
                                                     BinaryTree bt, uint32_t *pend, uint32_t *bcount) {
    __shared__ float s_box[B][2][6];
    __shared__ uint32_t s_arrived[B];
    // parent links of the run's leaves and of its internal nodes (a node completed inside the run
    // has its range, and so its index, inside the run), read once with coalesced loads: the climb
    // then takes its links from LDS instead of one dependent global load per step
    __shared__ uint2 s_leaf_link[B], s_int_link[B];
    const uint32_t L0 = blockIdx.x * B, L1 = min(L0 + (uint32_t)B - 1u, n - 1u);
    s_arrived[threadIdx.x] = 0u;
    {
        const uint32_t i = L0 + threadIdx.x;
        if (i <= L1) s_leaf_link[threadIdx.x] = bt.parent_leaf[i];
        if (i <= L1 && i + 1u < n) s_int_link[threadIdx.x] = bt.parent_int[i]; // internal nodes 0..n-2
    }
    __syncthreads();
    const uint32_t leaf = L0 + threadIdx.x;
    uint32_t deferred = kNoEntry;
    if (leaf <= L1) {
        uint32_t cur = leaf | kLeafBit, f = leaf, l = leaf;
        float3 lo, hi;
        if (srec) copy_record(prec + lb.ids[leaf], srec + leaf, lo, hi); // triangle records in sorted order
        else stored_box(lb.pbox, lb.ids[leaf], lo, hi);
        for (;;) {
            const uint2 link = (cur & kLeafBit) ? s_leaf_link[f - L0] : s_int_link[cur - L0];
            const uint32_t p = link.x, side = link.y >> 31, end = link.y & ~kLeafBit;
            const uint32_t pf = side ? end : f, pl = side ? l : end;
            if (pf < L0 || pl > L1) {
                deferred = cur;
                break;
            }
            bt.cbox[4 * (uint64_t)p + 2 * side] = make_float4(lo.x, lo.y, lo.z, 0.0f);
            bt.cbox[4 * (uint64_t)p + 2 * side + 1] = make_float4(hi.x, hi.y, hi.z, 0.0f);
            const uint32_t k = p - L0;
            float *mine = s_box[k][side];
            mine[0] = lo.x; mine[1] = lo.y; mine[2] = lo.z;
            mine[3] = hi.x; mine[4] = hi.y; mine[5] = hi.z;
            const uint32_t before = __hip_atomic_fetch_add(&s_arrived[k], 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (before == 0u) break; // the sibling's thread completes p
            const float *other = s_box[k][1u - side];
            lo = make_float3(fminf(lo.x, other[0]), fminf(lo.y, other[1]), fminf(lo.z, other[2]));
            hi = make_float3(fmaxf(hi.x, other[3]), fmaxf(hi.y, other[4]), fmaxf(hi.z, other[5]));
            cur = p;
            f = pf;
            l = pl;
        }
    }
    compact_deferred(deferred, pend, bcount);
}

// Next phase's entries: workgroup b of the phase left bcount[b] entries at src[b * run]; they go
// to dst[boff[b] ...] (boff = exclusive scan of bcount). The last workgroup writes the total.
__global__ void k_gather_deferred(const uint32_t *src, const uint32_t *bcount, const uint32_t *boff, uint32_t run,
                                  uint32_t nblocks, uint32_t *dst, uint32_t *total) {
    const uint32_t b = blockIdx.x;
    const uint32_t c = bcount[b], o = boff[b];
    for (uint32_t t = threadIdx.x; t < c; t += blockDim.x) dst[o + t] = src[(uint64_t)b * run + t];
    if (b == nblocks - 1u && threadIdx.x == 0u) *total = o + c;
}

__global__ void k_single_leaf_root(LeafBoxes lb, BuildCounters *ctr, uint2 *queue0) {
    float3 lo, hi;
    leaf_box(lb, 0u, lo, hi);
    publish_root(ctr, queue0, kLeafBit, lo, hi);
}

// Triangle records in sorted order for builds too small for the leaf phase.
__global__ void k_sorted_records(const TriRecord *prec, const uint32_t *ids, TriRecord *srec, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float3 lo, hi;
        copy_record(prec + ids[i], srec + i, lo, hi);
    }
}

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
    // 2^(b-127) > ext / 255 up to the rounding of ext and of the division, so one step up at most
    // restores lo + 255 s >= hi (checked exactly in double precision)
    if (e < 254 && (double)lo + 255.0 * (double)__int_as_float((int)(e << 23)) < (double)hi) e++;
    return e;
}

// 1 / s for the power of two s = 2^(e - 127): exact, so (v - p) * inv equals (v - p) / s.
__device__ __forceinline__ float inv_scale(uint32_t e) {
    return e <= 253u ? __int_as_float((int)((254u - e) << 23)) : 1.0f / __int_as_float((int)(e << 23));
}

// Conservative quantised bounds of v in node frame (p, s): the largest q with p + q s <= v and
// the smallest with p + q s >= v, within [0, 255]. The float quotient is off by at most one unit
// (the rounding of v - p is far below s), so one exact double-precision check corrects it.
__device__ __forceinline__ uint32_t quant_lo(float v, float p, float s, float inv) {
    float q = floorf((v - p) * inv);
    q = fminf(fmaxf(q, 0.0f), 255.0f);
    uint32_t qi = (uint32_t)q;
    if (qi > 0 && (double)p + (double)qi * (double)s > (double)v) qi--;
    return qi;
}

__device__ __forceinline__ uint32_t quant_hi(float v, float p, float s, float inv) {
    float q = ceilf((v - p) * inv);
    q = fminf(fmaxf(q, 0.0f), 255.0f);
    uint32_t qi = (uint32_t)q;
    if (qi < 255 && (double)p + (double)qi * (double)s < (double)v) qi++;
    return qi;
}

// Reductions over the 8 lanes of a collapse group through DPP lane moves (no LDS round trips):
// pairs and quads by quad permutes, then the two quads of the group by the half-row mirror (lane
// i and 7 - i lie in different quads). The 8 lanes of a group are active together.
template<int CTRL>
__device__ __forceinline__ uint32_t g8_move(uint32_t v) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)v, (int)v, CTRL, 0xF, 0xF, false);
}
__device__ __forceinline__ float g8_min(float v) {
    v = fminf(v, __uint_as_float(g8_move<0xB1>(__float_as_uint(v))));
    v = fminf(v, __uint_as_float(g8_move<0x4E>(__float_as_uint(v))));
    return fminf(v, __uint_as_float(g8_move<0x141>(__float_as_uint(v))));
}
__device__ __forceinline__ float g8_max(float v) {
    v = fmaxf(v, __uint_as_float(g8_move<0xB1>(__float_as_uint(v))));
    v = fmaxf(v, __uint_as_float(g8_move<0x4E>(__float_as_uint(v))));
    return fmaxf(v, __uint_as_float(g8_move<0x141>(__float_as_uint(v))));
}
__device__ __forceinline__ uint32_t g8_or(uint32_t v) {
    v |= g8_move<0xB1>(v);
    v |= g8_move<0x4E>(v);
    return v | g8_move<0x141>(v);
}
template<int CTRL>
__device__ __forceinline__ uint64_t g8_max_step(uint64_t v) {
    const uint64_t w = ((uint64_t)g8_move<CTRL>((uint32_t)(v >> 32)) << 32) | g8_move<CTRL>((uint32_t)v);
    return w > v ? w : v;
}
__device__ __forceinline__ uint64_t g8_max_u64(uint64_t v) {
    return g8_max_step<0x141>(g8_max_step<0x4E>(g8_max_step<0xB1>(v)));
}
__device__ __forceinline__ uint32_t ord_u32(float f) {
    uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

// ---- SAH refinement (FAST_TRACE builds). The radix tree is cut into clusters: the highest
// nodes with at most kSahCluster primitives (their parent has more). Each cluster's subtree is
// rebuilt top-down by binned SAH over its primitives (one wave per cluster), and its primitives
// are reordered into the new tree's leaf order, so every node still covers a contiguous run of
// positions and the rest of the build (bottom-up boxes from the cluster roots, collapse) is
// unchanged. The rebuilt subtree keeps the radix tree's index rule: a node split after position
// g has its left child at index g and its right child at g + 1 (leaf or internal), so its
// internal nodes take the indices the old subtree's internal nodes held, f + 1 .. l - 1, and the
// cluster root keeps its own index (its parent's link is untouched).
constexpr uint32_t kSahCluster = 64; // primitives per cluster (one per lane of a wave)
static_assert(kSahCluster == 64, "k_sah_clusters runs one wave64 per workgroup");
constexpr int kSahBins = 16;         // bins per axis: 15 split candidates per axis, 45 lanes

// Cluster roots in position order: the thread at position `pos` climbs from leaf pos while the
// parent starts at pos and holds at most kSahCluster primitives; it reports the node it stops at
// when that node's parent is larger (or the node is the root). Every cluster is reported by the
// thread at its first position, so the list comes out ordered by first leaf.
__global__ __launch_bounds__(kPhaseBlock) void k_sah_cut(uint32_t n, BinaryTree bt, uint32_t *pend, uint32_t *bcount) {
    const uint32_t pos = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t found = kNoEntry;
    if (pos < n) {
        uint32_t ref = pos | kLeafBit, f = pos, l = pos;
        for (;;) {
            if (ref == 0u) { // the root: the whole input is one cluster
                found = ref;
                break;
            }
            const uint2 link = (ref & kLeafBit) ? bt.parent_leaf[pos] : bt.parent_int[ref];
            const uint32_t p = link.x, side = link.y >> 31, end = link.y & ~kLeafBit;
            const uint32_t pf = side ? end : f, pl = side ? l : end;
            if (pl - pf + 1u > kSahCluster) {
                found = ref;
                break;
            }
            if (pf != pos) break; // inside a cluster that starts before pos
            ref = p;
            f = pf;
            l = pl;
        }
    }
    compact_deferred(found, pend, bcount);
}

// Wave64 reductions through DPP row operations (VALU lane moves, no LDS round trips): pairs,
// quads, half rows, rows, then rows 1 / 3 take lane 15 / 47 and rows 2 / 3 lane 31; lane 63 ends
// with the whole wave's value, broadcast with readlane. Every lane of the wave takes part.
template<bool MAX, int CTRL, int ROW_MASK>
__device__ __forceinline__ float dpp_step(float v) {
    const int o = __builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), CTRL, ROW_MASK, 0xF, false);
    return MAX ? fmaxf(v, __int_as_float(o)) : fminf(v, __int_as_float(o));
}
template<bool MAX>
__device__ __forceinline__ float wave_reduce(float v) {
    v = dpp_step<MAX, 0xB1, 0xF>(v);  // quad_perm [1,0,3,2]
    v = dpp_step<MAX, 0x4E, 0xF>(v);  // quad_perm [2,3,0,1]
    v = dpp_step<MAX, 0x141, 0xF>(v); // row_half_mirror
    v = dpp_step<MAX, 0x140, 0xF>(v); // row_mirror
    v = dpp_step<MAX, 0x142, 0xA>(v); // row_bcast:15 into rows 1 and 3
    v = dpp_step<MAX, 0x143, 0xC>(v); // row_bcast:31 into rows 2 and 3
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}
__device__ __forceinline__ float wave_sum_min(float v, bool on) { return wave_reduce<false>(on ? v : FLT_MAX); }
__device__ __forceinline__ float wave_sum_max(float v, bool on) { return wave_reduce<true>(on ? v : -FLT_MAX); }
__device__ __forceinline__ float lane_value(float v, uint32_t lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), (int)lane));
}

// Bin of a doubled centroid (NaN and values below the first boundary go to bin 0).
__device__ __forceinline__ int sah_bin(float cv, float cmin, float scale) {
    const float x = (cv - cmin) * scale;
    return x >= 1.0f ? min(kSahBins - 1, (int)fminf(x, (float)kSahBins)) : 0;
}

// Orders the LDS accesses of one wave (LDS operations of a wave complete in order; this keeps the
// compiler from moving them across the point). Unlike __syncthreads it does not wait for the
// wave's global stores, so a workgroup of exactly one wave uses it in place of a barrier.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// One wave (one workgroup of 64) per cluster, clusters taken in a grid-stride loop. The cluster's
// primitives sit in LDS in position order (box, doubled centroid, local index) and move when a
// task is partitioned, so the candidate scans read consecutive LDS words. A task (positions
// [a, b], node index) is split by the lanes: lane L < 45 scores candidate (axis L / 15, boundary
// L % 15 + 1) over the task's primitives, the cheapest area(left) * n_left + area(right) * n_right
// wins (lowest lane on ties), and a stable partition reorders the positions. Tasks of two
// primitives need no search. Centroids that coincide on every axis are split at the middle.
__global__ __launch_bounds__(64) void k_sah_clusters(const uint32_t *entries, const uint32_t *count_ptr, BinaryTree bt,
                                                     uint32_t *ids, const TriRecord *prec, TriRecord *srec,
                                                     const float4 *pbox) {
    __shared__ float s_box[6][kSahCluster];
    __shared__ float s_cen[3][kSahCluster];
    __shared__ uint32_t s_loc[kSahCluster];
    __shared__ uint32_t s_pos[kSahCluster];
    __shared__ uint2 s_task[kSahCluster];
    const uint32_t lane = threadIdx.x;
    const uint32_t count = *count_ptr;
    for (uint32_t c = blockIdx.x; c < count; c += gridDim.x) {
        const uint32_t ref = entries[c];
        if (ref & kLeafBit) { // a single primitive: its record goes to its position unchanged
            const uint32_t pos = ref & ~kLeafBit;
            if (srec && lane < 3u) {
                const float4 *src = reinterpret_cast<const float4 *>(prec + ids[pos]);
                reinterpret_cast<float4 *>(srec + pos)[lane] = src[lane];
            }
            continue;
        }
        const uint2 rg = bt.range[ref];
        const uint32_t f = rg.x, m = rg.y - rg.x + 1u; // 2 <= m <= kSahCluster
        const bool mine = lane < m;
        uint32_t id = 0;
        float4 r0 = make_float4(0.f, 0.f, 0.f, 0.f), r1 = r0, r2 = r0;
        if (mine) {
            float3 lo, hi;
            id = ids[f + lane];
            if (srec) {
                const float4 *src = reinterpret_cast<const float4 *>(prec + id);
                r0 = src[0];
                r1 = src[1];
                r2 = src[2];
                lo = make_float3(fminf(r0.x, fminf(r1.x, r2.x)), fminf(r0.y, fminf(r1.y, r2.y)), fminf(r0.z, fminf(r1.z, r2.z)));
                hi = make_float3(fmaxf(r0.x, fmaxf(r1.x, r2.x)), fmaxf(r0.y, fmaxf(r1.y, r2.y)), fmaxf(r0.z, fmaxf(r1.z, r2.z)));
            } else {
                stored_box(pbox, id, lo, hi);
            }
            s_box[0][lane] = lo.x; s_box[1][lane] = lo.y; s_box[2][lane] = lo.z;
            s_box[3][lane] = hi.x; s_box[4][lane] = hi.y; s_box[5][lane] = hi.z;
            s_cen[0][lane] = lo.x + hi.x; s_cen[1][lane] = lo.y + hi.y; s_cen[2][lane] = lo.z + hi.z;
            s_loc[lane] = lane;
        }
        if (lane == 0u) s_task[0] = make_uint2(m - 1u, ref); // positions [0, m - 1] packed as b, root index
        wave_sync();
        int sp = 1;
        while (sp > 0) { // uniform: every lane pops the same task
            sp--;
            const uint2 task = s_task[sp];
            const uint32_t a = task.x >> 16, b = task.x & 0xFFFFu, idx = task.y;
            const uint32_t mt = b - a + 1u;
            // lane j holds the primitive at position a + j
            const bool on = lane < mt;
            const uint32_t pj = on ? a + lane : a;
            const float3 blo = make_float3(s_box[0][pj], s_box[1][pj], s_box[2][pj]);
            const float3 bhi = make_float3(s_box[3][pj], s_box[4][pj], s_box[5][pj]);
            const float bc[3] = {s_cen[0][pj], s_cen[1][pj], s_cen[2][pj]};
            const uint32_t bloc = s_loc[pj];
            uint32_t nl = 1u;
            float3 llo, lhi, rlo, rhi;
            bool by_bins = false;
            uint32_t best_axis = 0, best_k = 0;
            float cmin[3] = {0.f, 0.f, 0.f}, cscale[3] = {0.f, 0.f, 0.f};
            if (mt > 2u) {
                for (int ax = 0; ax < 3; ax++) {
                    const float mn = wave_sum_min(bc[ax], on), mx = wave_sum_max(bc[ax], on);
                    cmin[ax] = mn;
                    cscale[ax] = mx > mn ? fminf((float)kSahBins / (mx - mn), FLT_MAX) : 0.0f;
                }
                // candidate of this lane: a scan of the task's positions
                float cost = FLT_MAX;
                float3 clo = make_float3(FLT_MAX, FLT_MAX, FLT_MAX), chi = make_float3(-FLT_MAX, -FLT_MAX, -FLT_MAX);
                float3 dlo = clo, dhi = chi;
                const uint32_t ax = lane / (kSahBins - 1), k = lane % (kSahBins - 1) + 1u;
                if (lane < 3u * (kSahBins - 1) && cscale[ax] > 0.0f) {
                    const float *cen = s_cen[ax];
                    uint32_t cl = 0;
#pragma unroll 4
                    for (uint32_t q = a; q <= b; q++) {
                        const float3 qlo = make_float3(s_box[0][q], s_box[1][q], s_box[2][q]);
                        const float3 qhi = make_float3(s_box[3][q], s_box[4][q], s_box[5][q]);
                        if ((uint32_t)sah_bin(cen[q], cmin[ax], cscale[ax]) < k) {
                            cl++;
                            clo = make_float3(fminf(clo.x, qlo.x), fminf(clo.y, qlo.y), fminf(clo.z, qlo.z));
                            chi = make_float3(fmaxf(chi.x, qhi.x), fmaxf(chi.y, qhi.y), fmaxf(chi.z, qhi.z));
                        } else {
                            dlo = make_float3(fminf(dlo.x, qlo.x), fminf(dlo.y, qlo.y), fminf(dlo.z, qlo.z));
                            dhi = make_float3(fmaxf(dhi.x, qhi.x), fmaxf(dhi.y, qhi.y), fmaxf(dhi.z, qhi.z));
                        }
                    }
                    if (cl > 0u && cl < mt)
                        cost = box_area(clo, chi) * (float)cl + box_area(dlo, dhi) * (float)(mt - cl);
                }
                // cheapest candidate, lowest lane on ties (NaN costs never win)
                const float best = wave_sum_min(cost == cost ? cost : FLT_MAX, true);
                const uint64_t at = __ballot(cost == best);
                const uint32_t win = at ? (uint32_t)__ffsll((long long)at) - 1u : 0u;
                if (best < FLT_MAX && at != 0u && win < 3u * (kSahBins - 1)) {
                    by_bins = true;
                    best_axis = win / (kSahBins - 1);
                    best_k = win % (kSahBins - 1) + 1u;
                    llo = make_float3(lane_value(clo.x, win), lane_value(clo.y, win), lane_value(clo.z, win));
                    lhi = make_float3(lane_value(chi.x, win), lane_value(chi.y, win), lane_value(chi.z, win));
                    rlo = make_float3(lane_value(dlo.x, win), lane_value(dlo.y, win), lane_value(dlo.z, win));
                    rhi = make_float3(lane_value(dhi.x, win), lane_value(dhi.y, win), lane_value(dhi.z, win));
                }
            }
            // partition the task's positions: left = bins below best_k (or the first half)
            bool left;
            if (by_bins) {
                left = sah_bin(bc[best_axis], cmin[best_axis], cscale[best_axis]) < (int)best_k;
            } else {
                left = lane < mt / 2u;
            }
            const uint64_t lmask = __ballot(on && left), rmask = __ballot(on && !left);
            nl = (uint32_t)__popcll(lmask);
            const uint64_t below = (1ull << lane) - 1ull;
            if (!by_bins && mt == 2u) { // two primitives: one each side
                llo = make_float3(s_box[0][a], s_box[1][a], s_box[2][a]);
                lhi = make_float3(s_box[3][a], s_box[4][a], s_box[5][a]);
                rlo = make_float3(s_box[0][b], s_box[1][b], s_box[2][b]);
                rhi = make_float3(s_box[3][b], s_box[4][b], s_box[5][b]);
            } else if (!by_bins) { // boxes of the halves
                const bool il = on && left, ir = on && !left;
                llo = make_float3(wave_sum_min(blo.x, il), wave_sum_min(blo.y, il), wave_sum_min(blo.z, il));
                lhi = make_float3(wave_sum_max(bhi.x, il), wave_sum_max(bhi.y, il), wave_sum_max(bhi.z, il));
                rlo = make_float3(wave_sum_min(blo.x, ir), wave_sum_min(blo.y, ir), wave_sum_min(blo.z, ir));
                rhi = make_float3(wave_sum_max(bhi.x, ir), wave_sum_max(bhi.y, ir), wave_sum_max(bhi.z, ir));
            }
            wave_sync(); // every lane has read the task's positions
            if (on && mt > 2u) { // move the primitives to their partitioned positions
                const uint32_t r = a + (left ? (uint32_t)__popcll(lmask & below) : nl + (uint32_t)__popcll(rmask & below));
                s_box[0][r] = blo.x; s_box[1][r] = blo.y; s_box[2][r] = blo.z;
                s_box[3][r] = bhi.x; s_box[4][r] = bhi.y; s_box[5][r] = bhi.z;
                s_cen[0][r] = bc[0]; s_cen[1][r] = bc[1]; s_cen[2][r] = bc[2];
                s_loc[r] = bloc;
            }
            // the node: split after position g; children at g (left) and g + 1 (right)
            const uint32_t g = a + nl - 1u;
            const uint32_t lref = nl == 1u ? ((f + a) | kLeafBit) : (f + g);
            const uint32_t rref = mt - nl == 1u ? ((f + b) | kLeafBit) : (f + g + 1u);
            if (lane == 0u) {
                bt.child[idx] = make_uint2(lref, rref);
                bt.range[idx] = make_uint2(f + a, f + b);
                float4 *cb = bt.cbox + 4 * (uint64_t)idx;
                cb[0] = make_float4(llo.x, llo.y, llo.z, 0.0f);
                cb[1] = make_float4(lhi.x, lhi.y, lhi.z, 0.0f);
                cb[2] = make_float4(rlo.x, rlo.y, rlo.z, 0.0f);
                cb[3] = make_float4(rhi.x, rhi.y, rhi.z, 0.0f);
                if (mt - nl >= 2u) s_task[sp] = make_uint2(((a + nl) << 16) | b, f + g + 1u);
                if (nl >= 2u) s_task[sp + (mt - nl >= 2u ? 1 : 0)] = make_uint2((a << 16) | g, f + g);
            }
            sp += (mt - nl >= 2u ? 1 : 0) + (nl >= 2u ? 1 : 0);
            wave_sync(); // positions and the task stack written
        }
        // primitives to their final positions
        if (mine) s_pos[s_loc[lane]] = lane;
        wave_sync();
        if (mine) {
            const uint32_t pos = f + s_pos[lane];
            ids[pos] = id;
            if (srec) {
                float4 *dst = reinterpret_cast<float4 *>(srec + pos);
                dst[0] = r0;
                dst[1] = r1;
                dst[2] = r2;
            }
        }
        wave_sync(); // LDS reused by the next cluster
    }
}

// One wide level. Each group of 8 lanes turns one work item (binary ref, wide node index) into a
// WideNode of at most `width` (2..8) children; lane k of the group holds child candidate k. All shuffles are executed by every lane
// of a group (groups without an item compute on empty data), and each group's control flow is
// uniform across its 8 lanes; the loop over items is uniform across the block.
//
// A candidate carries its sorted range [first, last]: in the radix tree an internal node split at
// gamma has children covering [first, gamma] and [gamma + 1, last] (child.x & ~kLeafBit == gamma
// for leaf and internal children alike), so opening a child needs one round of loads (its child
// pair and boxes) and no range lookups.
//
// Wide nodes of level l + 1 are numbered after those of levels 0..l in queue order, so a child's
// node index is (nodes of levels 0..l) + its queue slot. Queue slots and triangle ranges are
// reserved with one 64-bit atomic per workgroup iteration. Leaf slots list their primitives in
// out.prim_out in triangle-record order: primitive ids for a box build, sorted positions for a
// triangle build, whose records are then copied from the sorted records (k_tri_records).
//
// Level l reads its work items from q[l & 1] and appends the next level's to q[(l + 1) & 1].
// TAIL: one workgroup walks levels `level`, `level` + 1, ... until one is empty: the levels below
// the ones launched over the whole GPU, present only in deep trees, with no host read of a level's
// count. The level words are read with agent-scope atomic loads (the tail's own reservations
// update them while it runs); a level's queue entries are written and read by the same workgroup,
// ordered by the barrier that ends each item round.
template<int BLOCK, bool TAIL>
__global__ __launch_bounds__(BLOCK) void k_collapse(LeafBoxes lb, BinaryTree bt, BuildCounters *ctr, uint32_t level,
                                                     uint2 *q0, uint2 *q1, WideOut out, uint32_t split_small,
                                                     uint32_t leaf_max, uint32_t width) {
    __shared__ __align__(16) uint32_t s_node[BLOCK / 8][20];
    __shared__ uint32_t s_wave[2][BLOCK / 64];
    __shared__ unsigned long long s_base;
    const auto level_word = [&](uint32_t l) {
        return __hip_atomic_load(&ctr->level[l], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    };
  for (uint32_t lv = level; lv < (TAIL ? kCollapseLaunches : level + 1u); lv++) {
    const uint2 *qin = (lv & 1u) ? q1 : q0;
    uint2 *qout = (lv & 1u) ? q0 : q1;
    const uint32_t count = (uint32_t)(level_word(lv) >> 32);
    if (count == 0) return;
    if (blockIdx.x == 0 && threadIdx.x == 0) atomicMax(&ctr->depth, lv + 1);
    uint32_t next_start = 0, tri_start = 0;
    for (uint32_t j = 0; j <= lv; j++) {
        const uint64_t lw = level_word(j);
        next_start += (uint32_t)(lw >> 32);
        tri_start += (uint32_t)lw;
    }
    const uint32_t lane = threadIdx.x & 7u;
    const uint32_t wl = threadIdx.x & 63u;
    const uint32_t groups = blockDim.x >> 3;
    for (uint32_t base = blockIdx.x * groups; base < count; base += gridDim.x * groups) {
        const uint32_t item_idx = base + (threadIdx.x >> 3);
        const bool active = item_idx < count;
        const uint2 item = active ? qin[item_idx] : make_uint2(kLeafBit, 0u);
        // ---- gather children: lanes 0/1 start with the binary node's two children
        bool valid = false;
        uint32_t ref = 0, first = 0, last = 0;
        float3 lo = make_float3(FLT_MAX, FLT_MAX, FLT_MAX), hi = make_float3(-FLT_MAX, -FLT_MAX, -FLT_MAX);
        uint32_t nc;
        if (!active) {
            nc = 8;
        } else if (item.x & kLeafBit) {
            nc = 1;
            if (lane == 0) {
                ref = item.x;
                first = last = item.x & ~kLeafBit;
                valid = true;
                leaf_box(lb, first, lo, hi);
            }
        } else {
            nc = 2;
            if (lane < 2) {
                const uint2 ch = bt.child[item.x], rg = bt.range[item.x];
                const uint32_t gamma = ch.x & ~kLeafBit;
                ref = lane == 0 ? ch.x : ch.y;
                first = lane == 0 ? rg.x : gamma + 1u;
                last = lane == 0 ? gamma : rg.y;
                const float4 a = bt.cbox[4 * (uint64_t)item.x + 2 * lane], b = bt.cbox[4 * (uint64_t)item.x + 2 * lane + 1];
                lo = make_float3(a.x, a.y, a.z);
                hi = make_float3(b.x, b.y, b.z);
                valid = true;
            }
        }
        // ---- open the widest child (must-split before may-split) until 8 children; while two
        // slots are free the runner-up opens in the same round (its loads overlap the winner's:
        // half the dependent load rounds)
        for (int it = 0; it < 7; it++) {
            uint64_t key = 0;
            if (valid) {
                const uint32_t cnt = last - first + 1u;
                const uint32_t prio = cnt > leaf_max ? 2u : ((split_small && !(ref & kLeafBit) && cnt > 1u) ? 1u : 0u);
                if (prio) key = ((uint64_t)prio << 40) | ((uint64_t)__float_as_uint(fmaxf(box_area(lo, hi), 0.0f)) << 3) | (7u - lane);
            }
            const uint64_t best = g8_max_u64(key);
            if (nc >= width || best == 0) break; // uniform within the group
            const uint32_t winner = 7u - (uint32_t)(best & 7u);
            const uint64_t best2 = g8_max_u64(nc + 2u <= width && lane != winner ? key : 0ull);
            const bool two = best2 != 0;
            const uint32_t runner = 7u - (uint32_t)(best2 & 7u);
            const uint32_t wref = __shfl(ref, (int)winner, 8), wlast = __shfl(last, (int)winner, 8);
            const uint32_t rref = __shfl(ref, (int)runner, 8), rlast = __shfl(last, (int)runner, 8);
            // lanes: the winner and lane nc take the winner's children, the runner-up and lane
            // nc + 1 the runner-up's (side 0 stays in the opened child's lane)
            const bool wside = lane == winner || lane == nc;
            const bool rside = two && (lane == runner || lane == nc + 1u);
            if (wside || rside) {
                const uint32_t side = (lane == winner || (two && lane == runner)) ? 0u : 1u;
                const uint32_t pref = wside ? wref : rref, plast = wside ? wlast : rlast;
                const uint2 ch = bt.child[pref];
                const float4 a = bt.cbox[4 * (uint64_t)pref + 2 * side], b = bt.cbox[4 * (uint64_t)pref + 2 * side + 1];
                const uint32_t gamma = ch.x & ~kLeafBit;
                ref = side ? ch.y : ch.x;
                if (side) first = gamma + 1u;
                last = side ? plast : gamma;
                lo = make_float3(a.x, a.y, a.z);
                hi = make_float3(b.x, b.y, b.z);
                valid = true;
            }
            nc += two ? 2u : 1u;
        }
        const uint32_t cnt = valid ? last - first + 1u : 0u;
        // ---- node box, origin and scales
        const float3 nlo = make_float3(g8_min(lo.x), g8_min(lo.y), g8_min(lo.z));
        const float3 nhi = make_float3(g8_max(hi.x), g8_max(hi.y), g8_max(hi.z));
        const uint32_t ex = pick_exponent(nlo.x, nhi.x), ey = pick_exponent(nlo.y, nhi.y), ez = pick_exponent(nlo.z, nhi.z);
        const float sx = __int_as_float((int)(ex << 23)), sy = __int_as_float((int)(ey << 23)), sz = __int_as_float((int)(ez << 23));
        const float ix = inv_scale(ex), iy = inv_scale(ey), iz = inv_scale(ez);
        // ---- octant slots: repeatedly give the best remaining (child, free slot) pair its slot,
        // scoring dot(child centre - node centre, octant direction)
        const float3 pc = make_float3(0.5f * (nlo.x + nhi.x), 0.5f * (nlo.y + nhi.y), 0.5f * (nlo.z + nhi.z));
        const float dx = 0.5f * (lo.x + hi.x) - pc.x, dy = 0.5f * (lo.y + hi.y) - pc.y, dz = 0.5f * (lo.z + hi.z) - pc.z;
        uint32_t slot = 8u, used = 0u;
        const uint32_t nvalid = active ? nc : 0u;
        for (uint32_t k = 0; k < nvalid; k++) {
            uint64_t key = 0;
            if (valid && slot == 8u) {
                float bv = -FLT_MAX;
                uint32_t bs = (uint32_t)__ffs((int)(~used & 0xFFu)) - 1u; // lowest free slot (NaN scores)
                for (uint32_t s = 0; s < 8; s++) {
                    if (used & (1u << s)) continue;
                    const float v = ((s & 1) ? dx : -dx) + ((s & 2) ? dy : -dy) + ((s & 4) ? dz : -dz);
                    if (v > bv) { bv = v; bs = s; }
                }
                key = ((uint64_t)ord_u32(bv) << 8) | (bs << 3) | (7u - lane);
            }
            const uint64_t best = g8_max_u64(key);
            const uint32_t bs = (uint32_t)(best >> 3) & 7u;
            if (lane == 7u - (uint32_t)(best & 7u)) slot = bs;
            used |= 1u << bs;
        }
        // ---- internal / leaf, ranks and triangle offsets in slot order
        const bool internal = valid && cnt > leaf_max;
        const uint32_t imask = g8_or(internal ? (1u << slot) : 0u);
        const uint32_t leafcnt = (valid && !internal) ? cnt : 0u;
        // each lane's leaf count in the nibble of its slot (counts <= kLeafMax < 16, slots
        // distinct), OR-reduced: the group's counts by slot; toff = the nibbles below this slot
        static_assert(kLeafMax < 16, "leaf counts packed in nibbles");
        const uint32_t by_slot = g8_or(leafcnt ? leafcnt << (4u * slot) : 0u);
        const auto nibble_sum = [](uint32_t x) {
            x = (x & 0x0F0F0F0Fu) + ((x >> 4) & 0x0F0F0F0Fu);
            return (x * 0x01010101u) >> 24;
        };
        const uint32_t ttotal = nibble_sum(by_slot);
        const uint32_t toff = slot < 8u ? nibble_sum(by_slot & ((1u << (4u * slot)) - 1u)) : ttotal;
        // ---- reservations: inclusive scan of the 8 groups' counts (uniform within a group) over
        // the wave, wave totals through LDS, one 64-bit atomic per block for the next level's
        // queue slots (upper half) and this level's triangle records (lower half)
        const uint32_t gi = active ? (uint32_t)__popc(imask) : 0u, gt = active ? ttotal : 0u;
        uint32_t si = gi, st = gt;
        for (int o = 8; o < 64; o <<= 1) {
            const uint32_t ui = __shfl_up(si, (unsigned)o, 64), ut = __shfl_up(st, (unsigned)o, 64);
            if (wl >= (uint32_t)o) { si += ui; st += ut; }
        }
        const uint32_t wave = threadIdx.x >> 6;
        if (wl == 63u) {
            s_wave[0][wave] = si;
            s_wave[1][wave] = st;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t bi = 0, btri = 0;
            for (uint32_t w = 0; w < BLOCK / 64; w++) {
                bi += s_wave[0][w];
                btri += s_wave[1][w];
            }
            unsigned long long b = 0;
            if (bi | btri)
                b = atomicAdd(reinterpret_cast<unsigned long long *>(&ctr->level[lv + 1]),
                              ((unsigned long long)bi << 32) | btri);
            s_base = b;
        }
        __syncthreads();
        uint32_t wpi = 0, wpt = 0;
        for (uint32_t w = 0; w < wave; w++) {
            wpi += s_wave[0][w];
            wpt += s_wave[1][w];
        }
        const unsigned long long bbase = s_base;
        const uint32_t qbase = (uint32_t)(bbase >> 32) + wpi + si - gi;
        const uint32_t tri_base = tri_start + (uint32_t)bbase + wpt + st - gt;
        const uint32_t child_base = next_start + qbase;
        const uint32_t rank = valid ? (uint32_t)__popc(imask & ((1u << slot) - 1u)) : 0u;
        uint32_t meta = 0u;
        if (active && valid) {
            if (internal) {
                meta = 0xFFu;
                qout[qbase + rank] = make_uint2(ref, child_base + rank);
            } else {
                meta = (cnt << 5) | toff;
                uint32_t prim[kLeafMax];
                // triangle builds list sorted positions (records are in srec), box builds ids
                for (uint32_t k = 0; k < kLeafMax; k++) prim[k] = k < cnt ? (lb.srec ? first + k : lb.ids[first + k]) : 0u;
                for (uint32_t k = 0; k < kLeafMax; k++)
                    if (k < cnt) out.prim_out[tri_base + toff + k] = prim[k];
            }
        }
        // ---- assemble the node in LDS: header words from lane 0, each child's bytes at its slot;
        // empty slots keep qlo = 255, qhi = 0, meta = 0 (never hit). Lanes 0..4 store the 80 B.
        uint32_t *nw = s_node[threadIdx.x >> 3];
        uint8_t *nb = reinterpret_cast<uint8_t *>(nw);
        for (uint32_t w = lane; w < 20u; w += 8u) nw[w] = (w >= 8u && w < 14u) ? 0xFFFFFFFFu : 0u;
        __syncthreads();
        if (lane == 0u) {
            nw[0] = __float_as_uint(nlo.x);
            nw[1] = __float_as_uint(nlo.y);
            nw[2] = __float_as_uint(nlo.z);
            nw[3] = ex | (ey << 8) | (ez << 16) | (imask << 24);
            nw[4] = child_base;
            nw[5] = tri_base;
        }
        if (valid) {
            nb[24 + slot] = (uint8_t)meta;
            nb[32 + slot] = (uint8_t)quant_lo(lo.x, nlo.x, sx, ix);
            nb[40 + slot] = (uint8_t)quant_lo(lo.y, nlo.y, sy, iy);
            nb[48 + slot] = (uint8_t)quant_lo(lo.z, nlo.z, sz, iz);
            nb[56 + slot] = (uint8_t)quant_hi(hi.x, nlo.x, sx, ix);
            nb[64 + slot] = (uint8_t)quant_hi(hi.y, nlo.y, sy, iy);
            nb[72 + slot] = (uint8_t)quant_hi(hi.z, nlo.z, sz, iz);
        }
        __syncthreads();
        if (active && lane < 5u)
            reinterpret_cast<uint4 *>(out.nodes + item.y)[lane] = reinterpret_cast<const uint4 *>(nw)[lane];
        __syncthreads();
    }
  }
}

// Totals of the per-level words (one thread).
__global__ void k_finish_counts(BuildCounters *ctr) {
    uint32_t nodes = 0, tris = 0;
    for (uint32_t l = 0; l < kMaxCollapseLevels; l++) {
        const uint64_t w = ctr->level[l];
        nodes += (uint32_t)(w >> 32);
        tris += (uint32_t)w;
    }
    ctr->node_count = nodes;
    ctr->tri_count = tris;
}

// Triangle records in leaf order: record i = sorted record order[i]. A node's leaf triangles are
// a few nearby sorted positions, so the reads stay close to sequential; writes are coalesced.
__global__ __launch_bounds__(kBlock) void k_tri_records(const uint32_t *order, const TriRecord *srec, TriRecord *tris,
                                                        uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float4 *src = reinterpret_cast<const float4 *>(srec + order[i]);
        float4 *dst = reinterpret_cast<float4 *>(tris + i);
        const float4 a = src[0], b = src[1], c = src[2];
        dst[0] = a;
        dst[1] = b;
        dst[2] = c;
    }
}

// ---- refit: new vertex positions, same topology (a MeshBuildCommand that updates, see
// src/backends/cuda/cuda_mesh.cpp:30-69). Every record reloads its triangle's vertices; then the
// wide levels, deepest first, recompute each slot box from the child's exact box (internal) or
// the leaf's records, and requantise against the new node box. Slot order, child and triangle
// bases and leaf ranges are kept.

// Records: the primitive id in v0.w names the triangle to reload.
__global__ __launch_bounds__(kBlock) void k_refit_records(PrimSource src, TriRecord *tris, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        float4 *tr = reinterpret_cast<float4 *>(tris + i);
        const uint32_t prim = __float_as_uint(tr[0].w);
        const uint32_t i0 = src.tri[3 * (uint64_t)prim], i1 = src.tri[3 * (uint64_t)prim + 1],
                       i2 = src.tri[3 * (uint64_t)prim + 2];
        const float4 a = load_vertex(src, i0), b = load_vertex(src, i1), c = load_vertex(src, i2);
        tr[0] = make_float4(a.x, a.y, a.z, __uint_as_float(prim));
        tr[1] = make_float4(b.x, b.y, b.z, 0.0f);
        tr[2] = make_float4(c.x, c.y, c.z, 0.0f);
    }
}

// One wide level; 8 lanes per node, lane k = slot k. exact[2 i], exact[2 i + 1]: node i's box.
__global__ __launch_bounds__(kRefitBlock) void k_refit_level(BuildCounters *ctr, uint32_t level, WideNode *nodes,
                                                                const TriRecord *tris, float4 *exact) {
    __shared__ __align__(16) uint32_t s_node[kRefitBlock / 8][20];
    const uint32_t count = (uint32_t)(ctr->level[level] >> 32);
    uint32_t start = 0;
    for (uint32_t j = 0; j < level; j++) start += (uint32_t)(ctr->level[j] >> 32);
    const uint32_t lane = threadIdx.x & 7u;
    const uint32_t groups = blockDim.x >> 3;
    for (uint32_t base = blockIdx.x * groups; base < count; base += gridDim.x * groups) {
        const uint32_t item = base + (threadIdx.x >> 3);
        const bool active = item < count;
        const uint32_t node = start + item;
        uint4 w0 = make_uint4(0u, 0u, 0u, 0u), w1 = make_uint4(0u, 0u, 0u, 0u);
        if (active) {
            const uint4 *src = reinterpret_cast<const uint4 *>(nodes + node);
            w0 = src[0];
            w1 = src[1];
        }
        const uint32_t imask = w0.w >> 24;
        const uint32_t meta = ((lane < 4u ? w1.z : w1.w) >> (8u * (lane & 3u))) & 0xFFu;
        float3 lo = make_float3(FLT_MAX, FLT_MAX, FLT_MAX), hi = make_float3(-FLT_MAX, -FLT_MAX, -FLT_MAX);
        if (meta != 0u) {
            if (imask & (1u << lane)) {
                const uint32_t c = w1.x + (uint32_t)__popc(imask & ((1u << lane) - 1u));
                const float4 a = exact[2 * (uint64_t)c], b = exact[2 * (uint64_t)c + 1];
                lo = make_float3(a.x, a.y, a.z);
                hi = make_float3(b.x, b.y, b.z);
            } else {
                const uint32_t cnt = meta >> 5, off = meta & 31u;
                for (uint32_t k = 0; k < cnt; k++) {
                    const float4 *tr = reinterpret_cast<const float4 *>(tris + w1.y + off + k);
                    for (int v = 0; v < 3; v++) {
                        const float4 p = tr[v];
                        lo = make_float3(fminf(lo.x, p.x), fminf(lo.y, p.y), fminf(lo.z, p.z));
                        hi = make_float3(fmaxf(hi.x, p.x), fmaxf(hi.y, p.y), fmaxf(hi.z, p.z));
                    }
                }
            }
        }
        const float3 nlo = make_float3(g8_min(lo.x), g8_min(lo.y), g8_min(lo.z));
        const float3 nhi = make_float3(g8_max(hi.x), g8_max(hi.y), g8_max(hi.z));
        const uint32_t ex = pick_exponent(nlo.x, nhi.x), ey = pick_exponent(nlo.y, nhi.y), ez = pick_exponent(nlo.z, nhi.z);
        const float sx = __int_as_float((int)(ex << 23)), sy = __int_as_float((int)(ey << 23)), sz = __int_as_float((int)(ez << 23));
        const float ix = inv_scale(ex), iy = inv_scale(ey), iz = inv_scale(ez);
        uint32_t *nw = s_node[threadIdx.x >> 3];
        uint8_t *nb = reinterpret_cast<uint8_t *>(nw);
        for (uint32_t w = 8u + lane; w < 20u; w += 8u) nw[w] = w < 14u ? 0xFFFFFFFFu : 0u;
        if (lane == 0u) {
            nw[0] = __float_as_uint(nlo.x);
            nw[1] = __float_as_uint(nlo.y);
            nw[2] = __float_as_uint(nlo.z);
            nw[3] = ex | (ey << 8) | (ez << 16) | (imask << 24);
            nw[4] = w1.x;
            nw[5] = w1.y;
            nw[6] = w1.z;
            nw[7] = w1.w;
        }
        __syncthreads();
        if (active && meta != 0u) {
            nb[32 + lane] = (uint8_t)quant_lo(lo.x, nlo.x, sx, ix);
            nb[40 + lane] = (uint8_t)quant_lo(lo.y, nlo.y, sy, iy);
            nb[48 + lane] = (uint8_t)quant_lo(lo.z, nlo.z, sz, iz);
            nb[56 + lane] = (uint8_t)quant_hi(hi.x, nlo.x, sx, ix);
            nb[64 + lane] = (uint8_t)quant_hi(hi.y, nlo.y, sy, iy);
            nb[72 + lane] = (uint8_t)quant_hi(hi.z, nlo.z, sz, iz);
        }
        __syncthreads();
        if (active && lane < 5u)
            reinterpret_cast<uint4 *>(nodes + node)[lane] = reinterpret_cast<const uint4 *>(nw)[lane];
        if (active && lane == 0u) {
            exact[2 * (uint64_t)node] = make_float4(nlo.x, nlo.y, nlo.z, 0.0f);
            exact[2 * (uint64_t)node + 1] = make_float4(nhi.x, nhi.y, nhi.z, 0.0f);
            if (node == 0u) {
                ctr->bounds[0] = nlo.x; ctr->bounds[1] = nlo.y; ctr->bounds[2] = nlo.z;
                ctr->bounds[3] = nhi.x; ctr->bounds[4] = nhi.y; ctr->bounds[5] = nhi.z;
            }
        }
        __syncthreads();
    }
}

// ---- procedural primitives: AABBs (packed float3 min, float3 max: 24 B, include/luisa/runtime/
// rtx/aabb.h:7-10) -> the builder's box layout, and leaf records in leaf order after the build.
__global__ void k_aabbs_to_boxes(const float *aabb, uint32_t n, float4 *boxes) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float *a = aabb + 6 * (uint64_t)i;
        boxes[2 * (uint64_t)i] = make_float4(a[0], a[1], a[2], 0.0f);
        boxes[2 * (uint64_t)i + 1] = make_float4(a[3], a[4], a[5], 0.0f);
    }
}

__global__ void k_proc_records(const uint32_t *ids, const float4 *boxes, uint32_t n, ProcRecord *out) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t p = ids[i];
        const float4 lo = boxes[2 * (uint64_t)p], hi = boxes[2 * (uint64_t)p + 1];
        float4 *o = reinterpret_cast<float4 *>(out + i);
        o[0] = make_float4(lo.x, lo.y, lo.z, __uint_as_float(p));
        o[1] = make_float4(hi.x, hi.y, hi.z, 0.0f);
    }
}

// ---- validation (tests and debugging): structural and containment invariants of a built tree.
// err[0] child index out of range      err[1] triangle range out of range
// err[2] leaf triangle outside its slot box
// err[3] child node extent outside its parent's slot box
// err[4] a triangle record referenced != 1 times   err[5] a primitive id covered != 1 times
// err[6] an empty slot with an occupied box (qlo <= qhi on every axis)
struct DecodedSlot {
    double lo[3], hi[3];
};

__device__ __forceinline__ uint32_t node_byte(const uint4 *w, int word, int slot) {
    const uint32_t *u = reinterpret_cast<const uint32_t *>(w);
    return (u[word + (slot >> 2)] >> (8 * (slot & 3))) & 0xFFu;
}

__device__ __forceinline__ DecodedSlot decode_slot(const uint4 *w, int s) {
    DecodedSlot d;
    const double p[3] = {__uint_as_float(w[0].x), __uint_as_float(w[0].y), __uint_as_float(w[0].z)};
    // word indices (uint32 units): qlo x/y/z at 8/10/12, qhi x/y/z at 14/16/18
    for (int a = 0; a < 3; a++) {
        const double sc = __int_as_float((int)(((w[0].w >> (8 * a)) & 0xFFu) << 23));
        d.lo[a] = p[a] + sc * (double)node_byte(w, 8 + 2 * a, s);
        d.hi[a] = p[a] + sc * (double)node_byte(w, 14 + 2 * a, s);
    }
    return d;
}

__global__ void k_validate_nodes(const WideNode *nodes, uint32_t node_count, const TriRecord *tris, uint32_t tri_count,
                                 uint32_t nprims, uint32_t *tri_marks, uint32_t *prim_marks, uint32_t *err) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= node_count) return;
    uint4 w[5];
    for (int k = 0; k < 5; k++) w[k] = reinterpret_cast<const uint4 *>(nodes + i)[k];
    const uint32_t imask = w[0].w >> 24, child_base = w[1].x, tri_base = w[1].y;
    for (int s = 0; s < 8; s++) {
        const uint32_t meta = node_byte(w, 6, s);
        const DecodedSlot b = decode_slot(w, s);
        if (meta == 0u) {
            if (node_byte(w, 8, s) <= node_byte(w, 14, s) && node_byte(w, 10, s) <= node_byte(w, 16, s) &&
                node_byte(w, 12, s) <= node_byte(w, 18, s))
                atomicAdd(&err[6], 1u);
            continue;
        }
        if (imask & (1u << s)) {
            const uint32_t c = child_base + (uint32_t)__popc(imask & ((1u << s) - 1u));
            if (c >= node_count || c == 0u) { atomicAdd(&err[0], 1u); continue; }
            uint4 cw[5];
            for (int k = 0; k < 5; k++) cw[k] = reinterpret_cast<const uint4 *>(nodes + c)[k];
            // The child's origin is its exact lower corner; its decoded slot tops exceed the exact
            // upper corner by less than one child quantum.
            const double cp[3] = {__uint_as_float(cw[0].x), __uint_as_float(cw[0].y), __uint_as_float(cw[0].z)};
            bool bad = false;
            for (int a = 0; a < 3; a++) bad |= cp[a] < b.lo[a];
            for (int cs = 0; cs < 8; cs++) {
                if (node_byte(cw, 6, cs) == 0u) continue;
                const DecodedSlot cb = decode_slot(cw, cs);
                for (int a = 0; a < 3; a++) {
                    const double q = __int_as_float((int)(((cw[0].w >> (8 * a)) & 0xFFu) << 23));
                    bad |= cb.hi[a] - q > b.hi[a];
                }
            }
            if (bad) atomicAdd(&err[3], 1u);
        } else {
            const uint32_t off = meta & 31u, cnt = meta >> 5;
            if (cnt == 0u || (uint64_t)tri_base + off + cnt > tri_count) { atomicAdd(&err[1], 1u); continue; }
            for (uint32_t k = 0; k < cnt; k++) {
                const uint32_t t = tri_base + off + k;
                atomicAdd(&tri_marks[t], 1u);
                if (tris) {
                    const float *v = reinterpret_cast<const float *>(tris + t);
                    const uint32_t prim = __float_as_uint(v[3]);
                    if (prim < nprims) atomicAdd(&prim_marks[prim], 1u); else atomicAdd(&err[5], 1u);
                    for (int q = 0; q < 3; q++)
                        for (int a = 0; a < 3; a++) {
                            const double x = v[4 * q + a];
                            if (x < b.lo[a] || x > b.hi[a]) atomicAdd(&err[2], 1u);
                        }
                }
            }
        }
    }
}

__global__ void k_validate_marks(const uint32_t *tri_marks, uint32_t tri_count, const uint32_t *prim_marks,
                                 uint32_t nprims, uint32_t *err) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < tri_count && tri_marks[i] != 1u) atomicAdd(&err[4], 1u);
    if (prim_marks && i < nprims && prim_marks[i] != 1u) atomicAdd(&err[5], 1u);
}

template<typename T>
T *carve(char *&p, size_t count) {
    uintptr_t a = (reinterpret_cast<uintptr_t>(p) + 255) & ~uintptr_t(255);
    T *r = reinterpret_cast<T *>(a);
    p = reinterpret_cast<char *>(a + count * sizeof(T));
    return r;
}

// Workgroups of a bottom-up phase (runs of >= kPhaseBlock entries or leaves).
size_t phase_blocks(size_t m) { return m / kPhaseBlock + 1; }

// Workgroup size of the collapse (LUISA_HIP_COLLAPSE_BLOCK: 256, 512 or 1024 threads, 8 per wide
// node) and its resident workgroups per CU. A workgroup reserves its nodes' next-level queue slots
// and triangle ranges with one atomic per pass over its items.
struct CollapseShape {
    int block;
    int blocks_per_cu;
};

template<int B>
int collapse_occupancy() {
    int b = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_collapse<B, false>, B, 0) != hipSuccess || b < 1) b = 1;
    return b;
}

CollapseShape collapse_shape() {
    static const CollapseShape shape = [] {
        int block = kCollapseBlock;
        if (const char *v = std::getenv("LUISA_HIP_COLLAPSE_BLOCK")) {
            const int b = std::atoi(v);
            if (b == 256 || b == 512 || b == 1024) block = b;
        }
        const int bpc = block == 1024 ? collapse_occupancy<1024>() : (block == 512 ? collapse_occupancy<512>()
                                                                                    : collapse_occupancy<256>());
        return CollapseShape{block, bpc};
    }();
    return shape;
}

// The scratch arrays of a build over n primitives (bvh_scratch_bytes(n) bytes). The counters
// stay valid after the build; a refit reuses them and the cbox array.
struct ScratchLayout {
    uint32_t *k0, *k1, *i0, *i1;
    TriRecord *srec, *prec; // triangle builds: records in sorted and in primitive order
    BinaryTree bt;
    uint32_t *pend, *ent[2];
    uint32_t *bcount, *boff; // per-workgroup deferred-entry counts of a phase and their scan
    uint2 *q0, *q1;
    BuildCounters *ctr;
    uint32_t *dcount;
    size_t temp_bytes;
    void *temp;
    ScratchLayout(void *scratch, uint32_t n) {
        const size_t m = std::max<uint32_t>(n, 2u);
        char *p = static_cast<char *>(scratch);
        k0 = carve<uint32_t>(p, m); k1 = carve<uint32_t>(p, m);
        i0 = carve<uint32_t>(p, m); i1 = carve<uint32_t>(p, m);
        srec = carve<TriRecord>(p, m);
        prec = carve<TriRecord>(p, m);
        bt.cbox = carve<float4>(p, 4 * m);
        bt.child = carve<uint2>(p, m);
        bt.range = carve<uint2>(p, m);
        bt.slots = carve<uint32_t>(p, m);
        bt.parent_int = carve<uint2>(p, m);
        bt.parent_leaf = carve<uint2>(p, m);
        pend = carve<uint32_t>(p, m);
        ent[0] = carve<uint32_t>(p, m);
        ent[1] = carve<uint32_t>(p, m);
        q0 = carve<uint2>(p, m); q1 = carve<uint2>(p, m);
        bcount = carve<uint32_t>(p, phase_blocks(m));
        boff = carve<uint32_t>(p, phase_blocks(m));
        ctr = carve<BuildCounters>(p, 1);
        dcount = carve<uint32_t>(p, 2);
        temp_bytes = radix_sort_count_bytes((uint32_t)m);
        temp = carve<char>(p, temp_bytes);
    }
};

} // namespace

// Scratch bytes for building over n primitives (everything but the output arrays).
size_t bvh_scratch_bytes(uint32_t n) {
    size_t m = std::max<uint32_t>(n, 2u);
    size_t b = 0;
    b += 2 * m * sizeof(uint32_t) + 256 * 2;        // keys
    b += 2 * m * sizeof(uint32_t) + 256 * 2;        // ids
    b += 2 * m * sizeof(TriRecord) + 256 * 2;       // triangle records, sorted and primitive order
    b += 4 * m * sizeof(float4) + 256;              // cbox
    b += m * sizeof(uint2) + 256;                   // child
    b += m * sizeof(uint2) + 256;                   // range
    b += m * sizeof(uint32_t) + 256;                // slots
    b += 2 * m * sizeof(uint2) + 256 * 2;           // parent links
    b += 3 * m * sizeof(uint32_t) + 256 * 3;        // pend + two entry lists
    b += 2 * phase_blocks(m) * sizeof(uint32_t) + 256 * 2; // phase workgroup counts + scan
    b += 2 * m * sizeof(uint2) + 256 * 2;           // collapse queues
    b += sizeof(BuildCounters) + 256;               // counters
    b += 2 * sizeof(uint32_t) + 256;                // entry counts of the two lists
    b += radix_sort_count_bytes((uint32_t)m) + 256; // digit counts of the key sort
    return b;
}

size_t bvh_node_capacity(uint32_t n) { return std::max<uint32_t>(n, 2u); }

// Checks a built tree; err7 (device, 7 words) receives the violation counts listed above.
hipError_t bvh_validate(hipStream_t stream, const WideNode *nodes, uint32_t node_count, const TriRecord *tris,
                        uint32_t tri_count, uint32_t nprims, uint32_t *marks, uint32_t *err7) {
    hipError_t e = hipMemsetAsync(marks, 0, sizeof(uint32_t) * ((size_t)tri_count + nprims), stream);
    if (e != hipSuccess) return e;
    e = hipMemsetAsync(err7, 0, 7 * sizeof(uint32_t), stream);
    if (e != hipSuccess) return e;
    uint32_t *prim_marks = tris ? marks + tri_count : nullptr;
    if (node_count)
        hipLaunchKernelGGL(k_validate_nodes, dim3((node_count + kBlock - 1) / kBlock), dim3(kBlock), 0, stream, nodes,
                           node_count, tris, tri_count, nprims, marks, prim_marks, err7);
    const uint32_t nmax = std::max(tri_count, nprims);
    if (nmax)
        hipLaunchKernelGGL(k_validate_marks, dim3((nmax + kBlock - 1) / kBlock), dim3(kBlock), 0, stream, marks, tri_count,
                           prim_marks, nprims, err7);
    return hipGetLastError();
}

// Builds a wide BVH over n primitives (triangles if boxes == nullptr, else boxes).
// Output arrays: nodes (bvh_node_capacity(n) entries), tris (n entries) or prim_out (n entries).
// The counters land in `counters_out` (device) for the caller to read back. The build only
// enqueues work: every count a step depends on is read on the device, so the host never waits.
hipError_t bvh_build(hipStream_t stream, void *scratch, size_t scratch_bytes, const void *vbase, uint64_t vstride,
                     uint64_t nverts, const uint32_t *tri, const float4 *boxes, uint32_t n, WideNode *nodes,
                     TriRecord *tris, uint32_t *prim_out, BuildCounters **counters_out, uint32_t split_small,
                     uint32_t leaf_max, uint32_t sah_refine, uint32_t width, int cu_count, uint32_t *error_flag,
                     uint32_t collapse_head) {
    if (leaf_max < 1 || leaf_max > kLeafMax) return hipErrorInvalidValue;
    if (width < 2 || width > 8) return hipErrorInvalidValue;
    if (scratch_bytes < bvh_scratch_bytes(n)) return hipErrorInvalidValue;
    // the collapse grid fills the CUs once (its loop strides over the level's items)
    const CollapseShape cs = collapse_shape();
    const int grid_collapse = std::max(1, cu_count) * cs.blocks_per_cu;
    ScratchLayout L(scratch, n);
    uint32_t *k0 = L.k0, *k1 = L.k1, *i0 = L.i0, *i1 = L.i1;
    BinaryTree bt = L.bt;
    uint32_t *pend = L.pend;
    uint32_t *ent[2] = {L.ent[0], L.ent[1]};
    uint2 *q0 = L.q0, *q1 = L.q1;
    BuildCounters *ctr = L.ctr;
    uint32_t *dcount = L.dcount;
    const size_t temp_bytes = L.temp_bytes;
    void *temp = L.temp;
    *counters_out = ctr;

    if (!boxes && n > 0 && nverts == 0) return hipErrorInvalidValue;
    PrimSource src{static_cast<const char *>(vbase), vstride, nverts, tri, boxes, error_flag};
    const uint32_t nslots = n > 1 ? n - 1 : 0;
    hipLaunchKernelGGL(k_init, dim3(std::max<uint32_t>(1, std::min<uint32_t>((nslots + kBlock - 1) / kBlock, 4096))),
                       dim3(kBlock), 0, stream, ctr, bt.slots, nslots);
    if (n == 0) return hipGetLastError();
    const uint32_t grid = (n + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_centroid_bounds, dim3(std::min<uint32_t>(grid, 2048)), dim3(kBlock), 0, stream, src, n, ctr);
    TriRecord *prec = boxes ? nullptr : L.prec;
    hipLaunchKernelGGL(k_morton, dim3(grid), dim3(kBlock), 0, stream, src, n, ctr, k0, prec);
    // stable sort of the primitive indices by key (k0 -> k1, ids -> i1); the sort's ping-pong
    // buffers are i0 and the first entry list, which the bottom-up phases fill later
    hipError_t e = radix_sort_pairs_with(stream, temp, temp_bytes, ent[0], i0, k0, k1, nullptr, i1, n, nullptr, 0u, 30u);
    if (e != hipSuccess) return e;
    const uint32_t *keys = k1;
    uint32_t *ids = i1;
    // SAH refinement of the radix tree's clusters (below); needs at least one split to choose
    const bool sah = sah_refine != 0u && n > 2u;
    // leaf boxes: triangle builds keep their records in sorted order (written by the leaf phase,
    // by the SAH cluster pass, or here when the build is too small for the leaf phase); box
    // builds read the input boxes
    TriRecord *srec = boxes ? nullptr : L.srec;
    const LeafBoxes lb{boxes, ids, srec};
    if (srec && n <= kLastBlock && !sah)
        hipLaunchKernelGGL(k_sorted_records, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, stream,
                           (const TriRecord *)prec, ids, srec, n);
    if (n == 1) {
        hipLaunchKernelGGL(k_single_leaf_root, dim3(1), dim3(1), 0, stream, lb, ctr, q0);
    } else {
        hipLaunchKernelGGL(k_karras, dim3((n - 1 + kBlock - 1) / kBlock), dim3(kBlock), 0, stream, n, keys, bt);
        // Bottom-up boxes without host reads: a build of at most kLastBlock primitives is one
        // workgroup. Otherwise a first pass (the leaf phase over 512-leaf runs, or the SAH pass
        // over the clusters) leaves the entries for the rest of the tree in ent[0] and their count
        // in dcount[0] (a scan of per-workgroup counts places them in order), then phases over
        // chunks of 256 entries (workgroup-scope handoffs) shrink the list, alternating between
        // ent[0] / dcount[0] and ent[1] / dcount[1], and one workgroup finishes from what remains.
        // The number of phases comes from an estimate of the first list's length (about 18
        // deferred entries per 512-leaf run, or 1 per SAH cluster of ~32 primitives, each phase
        // shrinking its list ~16x); it sets only how much the last workgroup has to do.
        const uint32_t phase_grid = (uint32_t)std::max(1, cu_count) * 4u;
        uint64_t estimate = 0;
        if (sah) {
            // clusters in position order -> ent[0]; each rebuilt by binned SAH, its primitives
            // reordered (ids, and the sorted records of a triangle build)
            const uint32_t nblocks = (n + kPhaseBlock - 1) / kPhaseBlock;
            hipLaunchKernelGGL(k_sah_cut, dim3(nblocks), dim3(kPhaseBlock), 0, stream, n, bt, pend, L.bcount);
            e = exclusive_scan_u32(stream, L.bcount, L.boff, nblocks);
            if (e != hipSuccess) return e;
            hipLaunchKernelGGL(k_gather_deferred, dim3(nblocks), dim3(kPhaseBlock), 0, stream, (const uint32_t *)pend,
                               (const uint32_t *)L.bcount, (const uint32_t *)L.boff, kPhaseBlock, nblocks, ent[0],
                               dcount);
            // one pass of resident workgroups (a second partial pass would double the time)
            static const int sah_blocks_per_cu = [] {
                int b = 0;
                if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_sah_clusters, 64, 0) != hipSuccess || b < 1) b = 1;
                return b;
            }();
            const uint32_t grid_sah = std::max<uint32_t>(
                1u, std::min<uint32_t>((uint32_t)std::max(1, cu_count) * (uint32_t)sah_blocks_per_cu, n / 2u));
            hipLaunchKernelGGL(k_sah_clusters, dim3(grid_sah), dim3(64), 0, stream, (const uint32_t *)ent[0],
                               (const uint32_t *)dcount, bt, ids, (const TriRecord *)prec, srec, boxes);
            estimate = n / 32u;
        } else if (n <= kLastBlock) {
            const uint32_t blk = (n + 63) / 64 * 64;
            hipLaunchKernelGGL(k_bounds_phase, dim3(1), dim3(blk), 0, stream, lb, n, bt, (const uint32_t *)nullptr, n,
                               pend, L.bcount, ctr, q0, 1u);
        } else {
            const uint32_t nblocks = (n + kLeafRun - 1) / kLeafRun;
            hipLaunchKernelGGL(k_bounds_leaves<kLeafRun>, dim3(nblocks), dim3(kLeafRun), 0, stream, lb,
                               (const TriRecord *)prec, srec, n, bt, pend, L.bcount);
            e = exclusive_scan_u32(stream, L.bcount, L.boff, nblocks);
            if (e != hipSuccess) return e;
            hipLaunchKernelGGL(k_gather_deferred, dim3(nblocks), dim3(kLeafRun), 0, stream, (const uint32_t *)pend,
                               (const uint32_t *)L.bcount, (const uint32_t *)L.boff, kLeafRun, nblocks, ent[0], dcount);
            estimate = (uint64_t)nblocks * 18u;
        }
        if (sah || n > kLastBlock) {
            uint32_t w = 0;
            for (uint64_t c = estimate; c > 4u * kLastBlock; c /= 16u) {
                hipLaunchKernelGGL(k_bounds_phase_dev, dim3(phase_grid), dim3(kPhaseBlock), 0, stream, lb, n, bt,
                                   (const uint32_t *)ent[w], (const uint32_t *)(dcount + w), pend, L.bcount, ctr, q0);
                e = exclusive_scan_u32_counted(stream, L.bcount, L.boff, dcount + w, kPhaseBlock);
                if (e != hipSuccess) return e;
                hipLaunchKernelGGL(k_gather_dev, dim3(phase_grid), dim3(kPhaseBlock), 0, stream, (const uint32_t *)pend,
                                   (const uint32_t *)L.bcount, (const uint32_t *)L.boff, (const uint32_t *)(dcount + w),
                                   ent[w ^ 1u], dcount + (w ^ 1u));
                w ^= 1u;
            }
            hipLaunchKernelGGL(k_bounds_last, dim3(1), dim3(kLastBlock), 0, stream, lb, n, bt, (const uint32_t *)ent[w],
                               (const uint32_t *)(dcount + w), ctr, q0);
        }
    }
    // triangle BLAS: leaf sorted positions go to `pend` (free after the bottom-up phases) and the
    // records are copied from the sorted records in one pass after the last level
    WideOut out{nodes, tris, tris ? pend : prim_out};
    // A root-to-leaf path of the radix tree crosses each of the 30 key bits and 32 index bits at
    // most once (a binned-SAH cluster adds at most 63 levels below its root), and every wide level
    // descends at least one binary level, so the wide depth stays below kCollapseLaunches. The
    // first `head` levels (by default the expected depth, log_width of the primitive count, + 4) are launched over
    // the whole GPU; a launch whose level is empty returns at once. One workgroup then walks any
    // deeper levels. No level count is read on the host; the host checks after the build that
    // level kCollapseLaunches holds no nodes.
    uint32_t head = 4;
    for (uint64_t cap = 1; cap < n; cap *= std::max<uint32_t>(width, 2u)) head++;
    if (collapse_head > 0) head = collapse_head; // set by the caller (lc_hip_build_set_collapse_head)
    head = std::min<uint32_t>(head, kCollapseLaunches);
    for (uint32_t level = 0; level < head; level++) {
        if (cs.block == 1024)
            hipLaunchKernelGGL((k_collapse<1024, false>), dim3(grid_collapse), dim3(1024), 0, stream, lb, bt, ctr, level,
                               q0, q1, out, split_small, leaf_max, width);
        else if (cs.block == 512)
            hipLaunchKernelGGL((k_collapse<512, false>), dim3(grid_collapse), dim3(512), 0, stream, lb, bt, ctr, level,
                               q0, q1, out, split_small, leaf_max, width);
        else
            hipLaunchKernelGGL((k_collapse<256, false>), dim3(grid_collapse), dim3(256), 0, stream, lb, bt, ctr, level,
                               q0, q1, out, split_small, leaf_max, width);
    }
    if (head < kCollapseLaunches)
        hipLaunchKernelGGL((k_collapse<512, true>), dim3(1), dim3(512), 0, stream, lb, bt, ctr, head, q0, q1, out,
                           split_small, leaf_max, width);
    hipLaunchKernelGGL(k_finish_counts, dim3(1), dim3(1), 0, stream, ctr);
    if (tris)
        hipLaunchKernelGGL(k_tri_records, dim3(std::min<uint32_t>(grid, 1u << 20)), dim3(kBlock), 0, stream, pend,
                           (const TriRecord *)srec, tris, n);
    return hipGetLastError();
}

hipError_t aabbs_to_boxes(hipStream_t stream, const float *aabb, uint32_t n, float4 *boxes) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_aabbs_to_boxes, dim3(std::min<uint32_t>((n + kBlock - 1) / kBlock, 1u << 16)), dim3(kBlock), 0,
                       stream, aabb, n, boxes);
    return hipGetLastError();
}

hipError_t proc_records(hipStream_t stream, const uint32_t *ids, const float4 *boxes, uint32_t n, ProcRecord *out) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_proc_records, dim3(std::min<uint32_t>((n + kBlock - 1) / kBlock, 1u << 16)), dim3(kBlock), 0,
                       stream, ids, boxes, n, out);
    return hipGetLastError();
}

namespace {
// The refit's launches over the build's counters (level sizes; the root box is rewritten) and an
// array of 2 float4 per wide node for the nodes' exact boxes.
hipError_t refit_launches(hipStream_t stream, BuildCounters *ctr, float4 *exact, const void *vbase, uint64_t vstride,
                          uint64_t nverts, const uint32_t *tri, uint32_t n, uint32_t depth, WideNode *nodes,
                          TriRecord *tris, int cu_count, uint32_t *error_flag) {
    static const int refit_blocks_per_cu = [] {
        int b = 0;
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b, k_refit_level, kRefitBlock, 0) != hipSuccess || b < 1) b = 1;
        return b;
    }();
    PrimSource src{static_cast<const char *>(vbase), vstride, nverts, tri, nullptr, error_flag};
    hipLaunchKernelGGL(k_refit_records, dim3(std::min<uint32_t>((n + kBlock - 1) / kBlock, 1u << 20)), dim3(kBlock), 0,
                       stream, src, tris, n);
    const int grid = std::max(1, cu_count) * refit_blocks_per_cu;
    for (uint32_t level = depth; level-- > 0;)
        hipLaunchKernelGGL(k_refit_level, dim3(grid), dim3(kRefitBlock), 0, stream, ctr, level, nodes, tris, exact);
    return hipGetLastError();
}
} // namespace

// Refits a triangle BLAS built by bvh_build over the same scratch (its counters hold the level
// sizes) for new vertex positions; depth = the build's wide depth.
hipError_t bvh_refit(hipStream_t stream, void *scratch, size_t scratch_bytes, const void *vbase, uint64_t vstride,
                     uint64_t nverts, const uint32_t *tri, uint32_t n, uint32_t depth, WideNode *nodes,
                     TriRecord *tris, BuildCounters **counters_out, int cu_count, uint32_t *error_flag) {
    if (scratch_bytes < bvh_scratch_bytes(n) || nverts == 0) return hipErrorInvalidValue;
    ScratchLayout L(scratch, n);
    *counters_out = L.ctr;
    if (n == 0) return hipSuccess;
    return refit_launches(stream, L.ctr, L.bt.cbox, vbase, vstride, nverts, tri, n, depth, nodes, tris, cu_count,
                          error_flag);
}

// What a refit keeps of a build whose scratch is not kept: its counters (in a slot rounded up to
// 256 B), then 2 float4 per wide node (32 B per node instead of the build scratch's ~240 B per
// primitive).
constexpr size_t kRefitCounterSlot = (sizeof(BuildCounters) + 255) & ~size_t(255);

size_t bvh_refit_state_bytes(uint32_t node_count) {
    return kRefitCounterSlot + 2 * (size_t)std::max<uint32_t>(node_count, 1u) * sizeof(float4);
}

BuildCounters *bvh_refit_state_counters(void *state) { return static_cast<BuildCounters *>(state); }

hipError_t bvh_refit_from_state(hipStream_t stream, void *state, size_t state_bytes, uint32_t node_count,
                                const void *vbase, uint64_t vstride, uint64_t nverts, const uint32_t *tri, uint32_t n,
                                uint32_t depth, WideNode *nodes, TriRecord *tris, int cu_count, uint32_t *error_flag) {
    if (!state || state_bytes < bvh_refit_state_bytes(node_count) || nverts == 0) return hipErrorInvalidValue;
    if (n == 0) return hipSuccess;
    float4 *exact = reinterpret_cast<float4 *>(static_cast<char *>(state) + kRefitCounterSlot);
    return refit_launches(stream, static_cast<BuildCounters *>(state), exact, vbase, vstride, nverts, tri, n, depth,
                          nodes, tris, cu_count, error_flag);
}

} // namespace lchip
