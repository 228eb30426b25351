// lc_hip_bvh.h -- in-HBM layouts shared by the BVH build and traversal kernels (gfx950).
//
// A BLAS (one Mesh) lives in two device arrays:
//   nodes : WideNode[node_count]   80 B each, 16-B aligned, root at index 0
//   tris  : TriRecord[tri_count]   48 B each, leaf triangles copied out of the user buffers
// A TLAS (one Accel) is the same WideNode layout whose leaf "triangles" are instance indices,
// plus an InstanceRecord array.
//
// WideNode is an 8-wide node with child boxes quantised to 8 bits per plane relative to the
// node origin p and a per-axis power-of-two scale 2^(e-127) (the compressed layout of Ylitie,
// Karras, Laine, "Efficient Incoherent Ray Traversal on GPUs Through Compressed Wide BVHs",
// HPG 2017). Slot s of a node holds the child lying toward octant s (bit k of s set = positive
// along axis k); a ray visits slots in order i ^ oct(ray), which approximates front to back.
#pragma once
// Included by the library's kernels and, through lc_hip_device.h, by shaders compiled at run time
// with hiprtc: plain layouts only.

#ifndef __HIPCC_RTC__
#include <stddef.h>
#endif
#include <stdint.h>
#ifndef offsetof
#define offsetof(t, m) __builtin_offsetof(t, m)
#endif

namespace lchip {

// Word layout of a WideNode as five 16-byte vectors:
//   w0 = { p.x, p.y, p.z (float bits), ex | ey << 8 | ez << 16 | imask << 24 }
//   w1 = { child_base, tri_base, meta[0..3], meta[4..7] }
//   w2 = { qlo_x[0..3], qlo_x[4..7], qlo_y[0..3], qlo_y[4..7] }
//   w3 = { qlo_z[0..3], qlo_z[4..7], qhi_x[0..3], qhi_x[4..7] }
//   w4 = { qhi_y[0..3], qhi_y[4..7], qhi_z[0..3], qhi_z[4..7] }
// imask bit s: slot s is an internal child, stored at child_base + popcount(imask & ((1<<s)-1)).
// meta[s] == 0: empty slot. Leaf slot: meta = count << 5 | offset, its triangles are
// tris[tri_base + offset .. + count). Internal slot: meta = 0xFF.
struct alignas(16) WideNode {
    uint32_t w[20];
};
static_assert(sizeof(WideNode) == 80, "WideNode must be 80 bytes");

// Leaf triangle: three vertices as float4; v0.w holds the mesh-local primitive index bits.
struct alignas(16) TriRecord {
    float v0[4];
    float v1[4];
    float v2[4];
};
static_assert(sizeof(TriRecord) == 48, "TriRecord must be 48 bytes");

// Leaf record of a procedural primitive (an AABB, ProceduralPrimitiveBuildCommand,
// include/luisa/runtime/rhi/command.h:489-510): lo.w holds the primitive index bits. A procedural
// BLAS has one primitive per leaf slot.
struct alignas(16) ProcRecord {
    float lo[4];
    float hi[4];
};
static_assert(sizeof(ProcRecord) == 32, "ProcRecord must be 32 bytes");

// Instance as seen by traversal: world->object affine (row-major 3x4), BLAS pointers.
struct alignas(16) InstanceRecord {
    float inv[12];       // world -> object
    float fwd[12];       // object -> world (reported by instance_transform)
    uint64_t blas_nodes; // device address of the BLAS WideNode array (0 = no geometry)
    uint64_t blas_tris;  // device address of the BLAS TriRecord (ProcRecord if procedural) array
    uint32_t vis_mask;   // instance visibility (default 0xff)
    uint32_t user_id;
    uint32_t flags;      // kInstOpaque | kInstProcedural
    uint32_t blas_depth; // wide depth of the BLAS
};
static_assert(sizeof(InstanceRecord) == 128, "InstanceRecord must be 128 bytes");

constexpr uint32_t kInstOpaque = 1u, kInstProcedural = 2u; // InstanceRecord::flags
constexpr uint32_t kLeafMax = 3;         // triangles per leaf slot (<= 3 keeps a node's tris in 24 bits)
constexpr uint32_t kCollapseLaunches = 98;   // wide levels launched per build (depth <= 96)
constexpr uint32_t kMaxCollapseLevels = 100; // per-level counter words
constexpr uint32_t kLeafBit = 0x80000000u; // binary-node reference: leaf (sorted position)

// Per-build device-side counters (one per BLAS/TLAS build).
struct BuildCounters {
    uint32_t node_count;  // wide nodes (root = 0), summed from the levels after the collapse
    uint32_t tri_count;   // triangle records, summed likewise
    uint32_t depth;       // wide levels that held work
    uint32_t root_ref;    // binary root reference
    int32_t cbounds[6];   // centroid bounds as order-preserving ints
    float bounds[6];      // root box lo.xyz, hi.xyz
    // wide level l: (nodes << 32) | triangle records of its leaves; one 64-bit atomic reserves
    // both. On the little-endian device the node count is the word's upper 32-bit half.
    uint64_t level[kMaxCollapseLevels];
};
static_assert(offsetof(BuildCounters, level) % 8 == 0, "level words must be 8-B aligned");

// Address of level l's node count (the upper half of its word), for 4-byte device-to-host reads.
inline uint32_t *level_nodes_word(BuildCounters *c, uint32_t l) {
    return reinterpret_cast<uint32_t *>(&c->level[l]) + 1;
}

// An accel's user-visible instance fields on the device (the table the in-kernel instance ops
// read and write). dirty: bit f set when field f was written since the last build.
enum { kInstanceTransform = 0, kInstanceUserId = 1, kInstanceVisibility = 2 };
struct InstanceState {
    float affine[12];     // object -> world, row-major 3x4
    uint32_t user_id;
    uint32_t vis_mask;
    uint32_t dirty;
    uint32_t _pad;
};
static_assert(sizeof(InstanceState) == 64, "InstanceState is 64 bytes");

// What a ray is traced against: one instance's BLAS (single-instance form) or TLAS + BLASes.
struct Geometry {
    // single-instance form
    const WideNode *blas_nodes;
    const TriRecord *blas_tris;
    float inst_inv[12];
    uint32_t inst_mask;
    // two-level form
    const WideNode *tlas_nodes;
    const uint32_t *tlas_prims;
    const InstanceRecord *instances;
    uint32_t empty;       // nothing to traverse: every ray misses
};

// The view of an accel a ray query walks (also the device library's traces).
struct QueryGeometry {
    const WideNode *tlas_nodes;      // two-level form
    const uint32_t *tlas_prims;
    const InstanceRecord *instances; // instance 0 alone when `two` is 0
    uint32_t two;
    uint32_t empty;                  // nothing to traverse: every query misses
};

// Device-side descriptor of a built accel: what a run-time shader's accel argument points at
// (lc_hip_device.h), rewritten by every accel build.
struct DeviceAccel {
    QueryGeometry geo;
    InstanceState *state; // instance table of the instance ops (first word of state[count]: error word)
    uint32_t state_count;
    uint32_t stack_need;  // stack entries a walk can need (the host checks it against the shader stack)
    uint32_t *error;      // device error word: bit 0 stack overflow
    uint32_t _pad[2];
};

// One slot of a bindless array (buffer at its bound offset, bytes from there to the end).
struct BindlessSlot {
    void *ptr;
    uint64_t size_bytes;
};
// A 2-D texture slot of a bindless array: the image's pixels and the slot's sampler code
// (filter << 2 | address, include/luisa/runtime/rhi/sampler.h:7-31); data null when empty.
struct BindlessTex2D {
    void *data;
    uint32_t width, height, storage, sampler;
    uint32_t _pad[2];
};

} // namespace lchip
