// lc_oracle.cpp -- CPU restatement of the reference CPU backend's ray-tracing semantics.
// TEST INFRASTRUCTURE ONLY (the checker); see lc_oracle.h for scope and citations.
//
// Build: g++ -O2 -ffp-contract=off -fno-fast-math (no FMA contraction, IEEE float), so the
// float operation sequence of the watertight test is exactly the one written here.

#include "lc_oracle.h"

#include <algorithm>
#include <array>
#include <atomic>
#include <cfloat>
#include <cmath>
#include <cstring>
#include <functional>
#include <limits>
#include <map>
#include <memory>
#include <thread>
#include <utility>
#include <vector>
#include <xmmintrin.h>
#include <pmmintrin.h>

namespace {

// ------------------------------------------------------------------------------------------
// PCG32 (O'Neill 2014). The reference seeds nothing for these workloads; the stream layout
// (initstate 42, initseq = workload number) is part of the workload definition.
// ------------------------------------------------------------------------------------------
constexpr uint64_t kPcgMult = 6364136223846793005ull;

inline uint32_t pcg_step(orc_pcg32 *r) {
    uint64_t old = r->state;
    r->state = old * kPcgMult + r->inc;
    uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((0u - rot) & 31u));
}

inline float pcg_float(orc_pcg32 *r) { return (float)(pcg_step(r) >> 8) * 0x1p-24f; }

void set_ftz_daz() {
    // The reference CPU backend's worker threads run with FTZ/DAZ (cpu/mod.rs:503-516).
    _MM_SET_FLUSH_ZERO_MODE(_MM_FLUSH_ZERO_ON);
    _MM_SET_DENORMALS_ZERO_MODE(_MM_DENORMALS_ZERO_ON);
}

void parallel_for(uint64_t n, uint64_t chunk, int nthreads, const std::function<void(uint64_t, uint64_t)> &fn) {
    if (nthreads <= 0) nthreads = (int)std::max(1u, std::thread::hardware_concurrency());
    uint64_t nchunks = (n + chunk - 1) / chunk;
    if (nchunks == 0) return;
    std::atomic<uint64_t> next{0};
    auto worker = [&]() {
        set_ftz_daz();
        for (;;) {
            uint64_t c = next.fetch_add(1);
            if (c >= nchunks) break;
            uint64_t b = c * chunk, e = std::min(n, b + chunk);
            fn(b, e);
        }
    };
    int nt = (int)std::min<uint64_t>((uint64_t)nthreads, nchunks);
    if (nt <= 1) {
        // run on the caller's thread without leaving FTZ/DAZ behind in it
        const unsigned int csr = _mm_getcsr();
        worker();
        _mm_setcsr(csr);
        return;
    }
    std::vector<std::thread> ts;
    for (int i = 0; i < nt; i++) ts.emplace_back(worker);
    for (auto &t : ts) t.join();
}

// ------------------------------------------------------------------------------------------
// Watertight ray/triangle test (Woop, Benthin, Wald, JCGT 2013), single precision with the
// paper's double-precision recomputation of U, V, W when any of them is exactly zero.
// Barycentrics follow the reference convention: bary.x weights vertex 1, bary.y vertex 2
// (src/dsl/rtx/hit.cpp:7-9), i.e. u = V/det, v = W/det with U weighting vertex 0.
// ------------------------------------------------------------------------------------------
struct WtRay {
    float o[3], d[3];
    int kx, ky, kz;
    float Sx, Sy, Sz;
};

inline void wt_setup(WtRay &r, const float o[3], const float d[3]) {
    for (int i = 0; i < 3; i++) { r.o[i] = o[i]; r.d[i] = d[i]; }
    float ax = std::fabs(d[0]), ay = std::fabs(d[1]), az = std::fabs(d[2]);
    int kz = (ax > ay) ? ((ax > az) ? 0 : 2) : ((ay > az) ? 1 : 2);
    int kx = kz + 1 == 3 ? 0 : kz + 1;
    int ky = kx + 1 == 3 ? 0 : kx + 1;
    if (d[kz] < 0.0f) std::swap(kx, ky);
    r.kx = kx; r.ky = ky; r.kz = kz;
    r.Sx = d[kx] / d[kz];
    r.Sy = d[ky] / d[kz];
    r.Sz = 1.0f / d[kz];
}

// Returns true and (t,u,v) when the ray hits the triangle with t in [tmin, tmax].
inline bool wt_intersect(const WtRay &r, const float *v0, const float *v1, const float *v2,
                         float tmin, float tmax, float &t_out, float &u_out, float &v_out) {
    const int kx = r.kx, ky = r.ky, kz = r.kz;
    const float A_x = v0[kx] - r.o[kx], A_y = v0[ky] - r.o[ky], A_z = v0[kz] - r.o[kz];
    const float B_x = v1[kx] - r.o[kx], B_y = v1[ky] - r.o[ky], B_z = v1[kz] - r.o[kz];
    const float C_x = v2[kx] - r.o[kx], C_y = v2[ky] - r.o[ky], C_z = v2[kz] - r.o[kz];
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

// ------------------------------------------------------------------------------------------
// Binned-SAH binary BVH over primitive boxes (triangles for a mesh, instances for a scene).
// ------------------------------------------------------------------------------------------
struct Box {
    float lo[3], hi[3];
    void empty() { for (int i = 0; i < 3; i++) { lo[i] = FLT_MAX; hi[i] = -FLT_MAX; } }
    void grow(const Box &b) { for (int i = 0; i < 3; i++) { lo[i] = std::min(lo[i], b.lo[i]); hi[i] = std::max(hi[i], b.hi[i]); } }
    void grow(const float *p) { for (int i = 0; i < 3; i++) { lo[i] = std::min(lo[i], p[i]); hi[i] = std::max(hi[i], p[i]); } }
    float area() const {
        float dx = hi[0] - lo[0], dy = hi[1] - lo[1], dz = hi[2] - lo[2];
        if (dx < 0 || dy < 0 || dz < 0) return 0.0f;
        return 2.0f * (dx * dy + dy * dz + dz * dx);
    }
};

struct Node {
    Box box;
    uint32_t first; // leaf: first index into prim order; internal: left child (right = left + 1)
    uint32_t count; // > 0 => leaf
};

struct Bvh {
    std::vector<Node> nodes;
    std::vector<uint32_t> order; // leaf slots -> primitive id
    std::atomic<uint32_t> node_count{0};
};

constexpr int kBins = 32;
constexpr uint32_t kMaxLeaf = 4;

struct BuildCtx {
    Bvh *bvh;
    const Box *boxes;
    const float *cent; // 3 per prim
    std::atomic<int> spawned{0};
    int max_spawn;
};

void build_rec(BuildCtx &ctx, uint32_t node_idx, uint32_t begin, uint32_t end, int depth) {
    Bvh &bvh = *ctx.bvh;
    Node &node = bvh.nodes[node_idx];
    Box nb, cb;
    nb.empty();
    cb.empty();
    for (uint32_t i = begin; i < end; i++) {
        uint32_t p = bvh.order[i];
        nb.grow(ctx.boxes[p]);
        cb.grow(ctx.cent + 3 * (size_t)p);
    }
    node.box = nb;
    uint32_t n = end - begin;
    if (n <= 2) { node.first = begin; node.count = n; return; }

    // SAH over all three axes with kBins bins on the centroid bounds.
    float best_cost = FLT_MAX;
    int best_axis = -1, best_split = -1;
    for (int axis = 0; axis < 3; axis++) {
        float ext = cb.hi[axis] - cb.lo[axis];
        if (!(ext > 0.0f)) continue;
        float scale = (float)kBins / ext;
        Box bb[kBins];
        uint32_t bc[kBins] = {};
        for (auto &b : bb) b.empty();
        for (uint32_t i = begin; i < end; i++) {
            uint32_t p = bvh.order[i];
            int bi = std::min(kBins - 1, (int)((ctx.cent[3 * (size_t)p + axis] - cb.lo[axis]) * scale));
            bb[bi].grow(ctx.boxes[p]);
            bc[bi]++;
        }
        float right_area[kBins];
        uint32_t right_cnt[kBins];
        Box acc;
        acc.empty();
        uint32_t cnt = 0;
        for (int b = kBins - 1; b > 0; b--) {
            acc.grow(bb[b]);
            cnt += bc[b];
            right_area[b] = acc.area();
            right_cnt[b] = cnt;
        }
        acc.empty();
        cnt = 0;
        for (int b = 0; b < kBins - 1; b++) {
            acc.grow(bb[b]);
            cnt += bc[b];
            if (cnt == 0 || right_cnt[b + 1] == 0) continue;
            float cost = acc.area() * (float)cnt + right_area[b + 1] * (float)right_cnt[b + 1];
            if (cost < best_cost) { best_cost = cost; best_axis = axis; best_split = b; }
        }
    }
    uint32_t mid;
    if (best_axis < 0) {
        if (n <= kMaxLeaf) { node.first = begin; node.count = n; return; }
        mid = begin + n / 2; // all centroids coincide: split by index
    } else {
        float leaf_cost = nb.area() * (float)n;
        if (n <= kMaxLeaf && leaf_cost <= best_cost + nb.area() /* traversal step */) {
            node.first = begin; node.count = n; return;
        }
        float ext = cb.hi[best_axis] - cb.lo[best_axis];
        float scale = (float)kBins / ext;
        auto it = std::partition(bvh.order.begin() + begin, bvh.order.begin() + end, [&](uint32_t p) {
            int bi = std::min(kBins - 1, (int)((ctx.cent[3 * (size_t)p + best_axis] - cb.lo[best_axis]) * scale));
            return bi <= best_split;
        });
        mid = (uint32_t)(it - bvh.order.begin());
        if (mid == begin || mid == end) mid = begin + n / 2;
    }
    uint32_t left = bvh.node_count.fetch_add(2);
    node.first = left;
    node.count = 0;
    if (n > 200000 && ctx.spawned.fetch_add(1) < ctx.max_spawn) {
        std::thread th([&ctx, left, begin, mid, depth]() { set_ftz_daz(); build_rec(ctx, left, begin, mid, depth + 1); });
        build_rec(ctx, left + 1, mid, end, depth + 1);
        th.join();
    } else {
        build_rec(ctx, left, begin, mid, depth + 1);
        build_rec(ctx, left + 1, mid, end, depth + 1);
    }
}

void build_bvh(Bvh &bvh, const std::vector<Box> &boxes, int nthreads) {
    uint32_t n = (uint32_t)boxes.size();
    std::vector<float> cent(3 * (size_t)n);
    for (uint32_t i = 0; i < n; i++)
        for (int a = 0; a < 3; a++) cent[3 * (size_t)i + a] = 0.5f * (boxes[i].lo[a] + boxes[i].hi[a]);
    bvh.order.resize(n);
    for (uint32_t i = 0; i < n; i++) bvh.order[i] = i;
    bvh.nodes.assign(std::max<size_t>(1, 2 * (size_t)n), Node{});
    bvh.node_count = 1;
    if (n == 0) { bvh.nodes[0].box.empty(); bvh.nodes[0].first = 0; bvh.nodes[0].count = 0; return; }
    BuildCtx ctx{&bvh, boxes.data(), cent.data(), {0}, std::max(0, (nthreads <= 0 ? (int)std::thread::hardware_concurrency() : nthreads) - 1)};
    build_rec(ctx, 0, 0, n, 0);
    bvh.nodes.resize(bvh.node_count.load());
}

// Conservative slab test. Ize 2013 scales the box's far distance by 1 + 2*gamma(3) so a box the
// ray touches survives rounding; the clip against the ray's [tmin, tmax] (and the closest-hit
// culling distance) must be widened too, or a triangle lying on the box face whose computed t
// is <= tmax is culled when the box's computed entry rounds above tmax. So the interval test
// is widened by 2^-12 of the larger distance on both sides: culling only skips work, and the
// triangle test alone decides what is hit.
struct SlabRay {
    float o[3], inv[3];
};

inline void slab_setup(SlabRay &s, const float o[3], const float d[3]) {
    for (int i = 0; i < 3; i++) {
        float dd = d[i];
        if (std::fabs(dd) < 1e-20f) dd = std::copysign(1e-20f, dd);
        s.o[i] = o[i];
        s.inv[i] = 1.0f / dd;
    }
}

inline bool slab_hit(const SlabRay &s, const Box &b, float tmin, float tmax, float &tnear) {
    float tn = tmin, tf = tmax;
    for (int a = 0; a < 3; a++) {
        float t0 = (b.lo[a] - s.o[a]) * s.inv[a];
        float t1 = (b.hi[a] - s.o[a]) * s.inv[a];
        if (t0 > t1) std::swap(t0, t1);
        tn = t0 > tn ? t0 : tn;
        tf = t1 < tf ? t1 : tf;
    }
    tnear = tn;
    const float m = 0x1p-12f * std::max(std::fabs(tn), std::fabs(tf));
    return tn - m <= tf + m;
}

struct Mesh {
    std::vector<float> v; // 3 floats per vertex (copied out of the strided input)
    std::vector<uint32_t> tri;
    Bvh bvh;
    Box bounds;
};

// Procedural primitive: AABBs (6 floats: min.xyz, max.xyz), include/luisa/runtime/rtx/aabb.h:7-10.
struct Procedural {
    std::vector<float> aabb;
};

struct Instance {
    const Mesh *mesh;       // null for a procedural instance
    const Procedural *proc; // null for a mesh instance
    float fwd[12]; // object -> world, row-major 3x4
    float inv[12]; // world -> object, row-major 3x4
    uint32_t mask;
    bool opaque;
};

struct Scene {
    std::vector<Instance> inst;
    Bvh tlas;
};

// Row-major 3x4 inverse in double precision, rounded to float.
void invert_affine(const float m[12], float out[12]) {
    double a = m[0], b = m[1], c = m[2], d = m[4], e = m[5], f = m[6], g = m[8], h = m[9], i = m[10];
    double A = e * i - f * h, B = -(d * i - f * g), C = d * h - e * g;
    double det = a * A + b * B + c * C;
    double id = det != 0.0 ? 1.0 / det : 0.0;
    double r[9] = {A * id, -(b * i - c * h) * id, (b * f - c * e) * id,
                   B * id, (a * i - c * g) * id, -(a * f - c * d) * id,
                   C * id, -(a * h - b * g) * id, (a * e - b * d) * id};
    double tx = m[3], ty = m[7], tz = m[11];
    for (int row = 0; row < 3; row++) {
        out[row * 4 + 0] = (float)r[row * 3 + 0];
        out[row * 4 + 1] = (float)r[row * 3 + 1];
        out[row * 4 + 2] = (float)r[row * 3 + 2];
        out[row * 4 + 3] = (float)(-(r[row * 3 + 0] * tx + r[row * 3 + 1] * ty + r[row * 3 + 2] * tz));
    }
}

inline void xform_point(const float m[12], const float p[3], float out[3]) {
    for (int r = 0; r < 3; r++) out[r] = m[4 * r + 0] * p[0] + m[4 * r + 1] * p[1] + m[4 * r + 2] * p[2] + m[4 * r + 3];
}

inline void xform_vector(const float m[12], const float p[3], float out[3]) {
    for (int r = 0; r < 3; r++) out[r] = m[4 * r + 0] * p[0] + m[4 * r + 1] * p[1] + m[4 * r + 2] * p[2];
}

struct TraceState {
    float t_best, t2;
    uint32_t inst, prim;
    float u, v;
    bool track2; // also track the second-nearest distinct primitive
};

// Traverse one mesh BVH; `cull` is the culling distance (t_best, or t2 when tracking two).
template<bool ANY>
bool trace_mesh(const Mesh &m, const float o[3], const float d[3], float tmin, float tmax,
                uint32_t inst_id, TraceState &st) {
    if (m.tri.empty()) return false;
    WtRay wr;
    wt_setup(wr, o, d);
    SlabRay sr;
    slab_setup(sr, o, d);
    uint32_t stack[128];
    int sp = 0;
    stack[sp++] = 0;
    const Node *nodes = m.bvh.nodes.data();
    while (sp > 0) {
        const Node &n = nodes[stack[--sp]];
        float cull = ANY ? tmax : (st.track2 ? std::min(tmax, st.t2) : std::min(tmax, st.t_best));
        float tn;
        if (!slab_hit(sr, n.box, tmin, cull, tn)) continue;
        if (n.count > 0) {
            for (uint32_t k = 0; k < n.count; k++) {
                uint32_t p = m.bvh.order[n.first + k];
                const uint32_t *ix = &m.tri[3 * (size_t)p];
                float t, u, v;
                if (!wt_intersect(wr, &m.v[3 * (size_t)ix[0]], &m.v[3 * (size_t)ix[1]], &m.v[3 * (size_t)ix[2]],
                                  tmin, tmax, t, u, v))
                    continue;
                if (ANY) return true;
                // nearer, or as near with a smaller (instance, primitive) key: the result does
                // not depend on the BVH or the visiting order
                const bool nearer = t != st.t_best ? t < st.t_best
                                                   : (inst_id != st.inst ? inst_id < st.inst : p < st.prim);
                if (nearer) {
                    if (st.track2) st.t2 = st.t_best;
                    st.t_best = t; st.inst = inst_id; st.prim = p; st.u = u; st.v = v;
                } else if (st.track2 && t < st.t2) {
                    st.t2 = t;
                }
            }
        } else {
            const Node &l = nodes[n.first], &r = nodes[n.first + 1];
            float tl, tr;
            float c2 = ANY ? tmax : (st.track2 ? std::min(tmax, st.t2) : std::min(tmax, st.t_best));
            bool hl = slab_hit(sr, l.box, tmin, c2, tl);
            bool hr = slab_hit(sr, r.box, tmin, c2, tr);
            if (hl && hr) {
                if (tl <= tr) { stack[sp++] = n.first + 1; stack[sp++] = n.first; }
                else { stack[sp++] = n.first; stack[sp++] = n.first + 1; }
            } else if (hl) {
                stack[sp++] = n.first;
            } else if (hr) {
                stack[sp++] = n.first + 1;
            }
        }
    }
    return false;
}

template<bool ANY>
bool trace_scene(const Scene &s, const float *ray, uint32_t mask, TraceState &st) {
    const float o[3] = {ray[0], ray[1], ray[2]};
    const float d[3] = {ray[4], ray[5], ray[6]};
    const float tmin = ray[3], tmax = ray[7];
    if (s.inst.empty()) return false;
    SlabRay sr;
    slab_setup(sr, o, d);
    uint32_t stack[128];
    int sp = 0;
    stack[sp++] = 0;
    const Node *nodes = s.tlas.nodes.data();
    while (sp > 0) {
        const Node &n = nodes[stack[--sp]];
        float cull = ANY ? tmax : (st.track2 ? std::min(tmax, st.t2) : std::min(tmax, st.t_best));
        float tn;
        if (!slab_hit(sr, n.box, tmin, cull, tn)) continue;
        if (n.count > 0) {
            for (uint32_t k = 0; k < n.count; k++) {
                uint32_t id = s.tlas.order[n.first + k];
                const Instance &in = s.inst[id];
                // procedural primitives answer ray queries only (cpu/accel.rs:476-484: plain
                // traces run without intersection functions)
                if (!in.mesh || (in.mask & mask & 0xffu) == 0u) continue;
                float lo[3], ld[3];
                xform_point(in.inv, o, lo);
                xform_vector(in.inv, d, ld);
                if (trace_mesh<ANY>(*in.mesh, lo, ld, tmin, tmax, id, st) && ANY) return true;
            }
        } else {
            stack[sp++] = n.first + 1;
            stack[sp++] = n.first;
        }
    }
    return false;
}

} // namespace

extern "C" {

void orc_pcg32_seed(orc_pcg32 *rng, uint64_t initstate, uint64_t initseq) {
    rng->state = 0u;
    rng->inc = (initseq << 1u) | 1u;
    pcg_step(rng);
    rng->state += initstate;
    pcg_step(rng);
}

uint32_t orc_pcg32_next(orc_pcg32 *rng) { return pcg_step(rng); }

float orc_pcg32_float(orc_pcg32 *rng) { return pcg_float(rng); }

void orc_pcg32_advance(orc_pcg32 *rng, uint64_t delta) {
    // Brown, "Random number generation with arbitrary strides" (1994): jump an LCG by delta.
    uint64_t cur_mult = kPcgMult, cur_plus = rng->inc, acc_mult = 1u, acc_plus = 0u;
    while (delta > 0) {
        if (delta & 1u) {
            acc_mult *= cur_mult;
            acc_plus = acc_plus * cur_mult + cur_plus;
        }
        cur_plus = (cur_mult + 1u) * cur_plus;
        cur_mult *= cur_mult;
        delta >>= 1u;
    }
    rng->state = acc_mult * rng->state + acc_plus;
}

void orc_icosphere_counts(int level, uint32_t *nverts, uint32_t *ntris) {
    uint64_t f = 20, v = 12;
    for (int l = 0; l < level; l++) { v += f * 3 / 2; f *= 4; }
    *nverts = (uint32_t)v;
    *ntris = (uint32_t)f;
}

void orc_icosphere(int level, float *verts4, uint32_t *tris) {
    const double p = (1.0 + std::sqrt(5.0)) / 2.0;
    std::vector<std::array<double, 3>> v = {
        {-1, p, 0}, {1, p, 0}, {-1, -p, 0}, {1, -p, 0}, {0, -1, p}, {0, 1, p},
        {0, -1, -p}, {0, 1, -p}, {p, 0, -1}, {p, 0, 1}, {-p, 0, -1}, {-p, 0, 1}};
    for (auto &x : v) {
        double l = std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
        for (auto &c : x) c /= l;
    }
    std::vector<std::array<uint32_t, 3>> f = {
        {0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11}, {1, 5, 9}, {5, 11, 4},
        {11, 10, 2}, {10, 7, 6}, {7, 1, 8}, {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8},
        {3, 8, 9}, {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};
    for (int l = 0; l < level; l++) {
        std::map<std::pair<uint32_t, uint32_t>, uint32_t> mid;
        auto midpoint = [&](uint32_t a, uint32_t b) {
            auto key = std::make_pair(std::min(a, b), std::max(a, b));
            auto it = mid.find(key);
            if (it != mid.end()) return it->second;
            std::array<double, 3> m;
            for (int c = 0; c < 3; c++) m[c] = 0.5 * (v[a][c] + v[b][c]);
            double len = std::sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
            for (auto &c : m) c /= len;
            v.push_back(m);
            uint32_t id = (uint32_t)v.size() - 1;
            mid.emplace(key, id);
            return id;
        };
        std::vector<std::array<uint32_t, 3>> nf;
        nf.reserve(f.size() * 4);
        for (auto &t : f) {
            uint32_t ab = midpoint(t[0], t[1]);
            uint32_t bc = midpoint(t[1], t[2]);
            uint32_t ca = midpoint(t[2], t[0]);
            nf.push_back({t[0], ab, ca});
            nf.push_back({t[1], bc, ab});
            nf.push_back({t[2], ca, bc});
            nf.push_back({ab, bc, ca});
        }
        f.swap(nf);
    }
    for (size_t i = 0; i < v.size(); i++) {
        verts4[4 * i + 0] = (float)v[i][0];
        verts4[4 * i + 1] = (float)v[i][1];
        verts4[4 * i + 2] = (float)v[i][2];
        verts4[4 * i + 3] = 0.0f;
    }
    for (size_t i = 0; i < f.size(); i++)
        for (int c = 0; c < 3; c++) tris[3 * i + c] = f[i][c];
}

void orc_rays_orbit_to_ball(uint64_t initstate, uint64_t initseq, uint64_t nrays, float orbit, float *rays) {
    orc_pcg32 rng;
    orc_pcg32_seed(&rng, initstate, initseq);
    const double two_pi = 6.283185307179586476925286766559;
    for (uint64_t i = 0; i < nrays; i++) {
        double z = 1.0 - 2.0 * (double)pcg_float(&rng);
        double phi = two_pi * (double)pcg_float(&rng);
        double r = std::sqrt(std::max(0.0, 1.0 - z * z));
        double o[3] = {orbit * r * std::cos(phi), orbit * r * std::sin(phi), orbit * z};
        double b[3];
        for (;;) {
            for (int c = 0; c < 3; c++) b[c] = 2.0 * (double)pcg_float(&rng) - 1.0;
            if (b[0] * b[0] + b[1] * b[1] + b[2] * b[2] <= 1.0) break;
        }
        double dv[3] = {b[0] - o[0], b[1] - o[1], b[2] - o[2]};
        double len = std::sqrt(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
        float *r8 = rays + 8 * i;
        r8[0] = (float)o[0]; r8[1] = (float)o[1]; r8[2] = (float)o[2]; r8[3] = 0.0f;
        r8[4] = (float)(dv[0] / len); r8[5] = (float)(dv[1] / len); r8[6] = (float)(dv[2] / len);
        r8[7] = 1e30f;
    }
}

void orc_soup(uint64_t initstate, uint64_t initseq, uint64_t ntris, float extent, float sigma,
              float *verts4, uint32_t *tris) {
    const double two_pi = 6.283185307179586476925286766559;
    parallel_for(ntris, 1u << 16, 0, [&](uint64_t b, uint64_t e) {
        orc_pcg32 rng;
        orc_pcg32_seed(&rng, initstate, initseq);
        orc_pcg32_advance(&rng, 13u * b);
        for (uint64_t i = b; i < e; i++) {
            double c[3];
            for (int k = 0; k < 3; k++) c[k] = (double)extent * (double)pcg_float(&rng);
            double nrm[10];
            for (int k = 0; k < 5; k++) {
                double u1 = 1.0 - (double)pcg_float(&rng);
                double u2 = (double)pcg_float(&rng);
                double r = std::sqrt(-2.0 * std::log(u1));
                nrm[2 * k] = r * std::cos(two_pi * u2);
                nrm[2 * k + 1] = r * std::sin(two_pi * u2);
            }
            for (int vtx = 0; vtx < 3; vtx++) {
                float *p = verts4 + 4 * (3 * i + vtx);
                for (int k = 0; k < 3; k++) p[k] = (float)(c[k] + (double)sigma * nrm[3 * vtx + k]);
                p[3] = 0.0f;
                tris[3 * i + vtx] = (uint32_t)(3 * i + vtx);
            }
        }
    });
}

void orc_rays_box_sphere(uint64_t initstate, uint64_t initseq, uint64_t skip, uint64_t nrays,
                         float lo, float hi, float *rays) {
    const double two_pi = 6.283185307179586476925286766559;
    parallel_for(nrays, 1u << 16, 0, [&](uint64_t b, uint64_t e) {
        orc_pcg32 rng;
        orc_pcg32_seed(&rng, initstate, initseq);
        orc_pcg32_advance(&rng, skip + 5u * b);
        for (uint64_t i = b; i < e; i++) {
            double o[3];
            for (int k = 0; k < 3; k++) o[k] = (double)lo + ((double)hi - (double)lo) * (double)pcg_float(&rng);
            double z = 1.0 - 2.0 * (double)pcg_float(&rng);
            double phi = two_pi * (double)pcg_float(&rng);
            double r = std::sqrt(std::max(0.0, 1.0 - z * z));
            float *r8 = rays + 8 * i;
            r8[0] = (float)o[0]; r8[1] = (float)o[1]; r8[2] = (float)o[2]; r8[3] = 0.0f;
            r8[4] = (float)(r * std::cos(phi)); r8[5] = (float)(r * std::sin(phi)); r8[6] = (float)z;
            r8[7] = 1e30f;
        }
    });
}

void orc_instances(uint64_t initstate, uint64_t initseq, uint32_t count, uint32_t mesh_count, float extent, float slo,
                   float shi, float *affines, uint32_t *mesh_ids) {
    const double two_pi = 6.283185307179586476925286766559;
    orc_pcg32 rng;
    orc_pcg32_seed(&rng, initstate, initseq);
    for (uint32_t i = 0; i < count; i++) {
        double m = (double)pcg_float(&rng) * mesh_count;
        mesh_ids[i] = std::min<uint32_t>(mesh_count - 1, (uint32_t)m);
        // Shoemake, "Uniform random rotations", Graphics Gems III (1992)
        double u1 = pcg_float(&rng), u2 = pcg_float(&rng), u3 = pcg_float(&rng);
        double q[4] = {std::sqrt(1.0 - u1) * std::sin(two_pi * u2), std::sqrt(1.0 - u1) * std::cos(two_pi * u2),
                       std::sqrt(u1) * std::sin(two_pi * u3), std::sqrt(u1) * std::cos(two_pi * u3)};
        double tr[3];
        for (int k = 0; k < 3; k++) tr[k] = (double)extent * (2.0 * (double)pcg_float(&rng) - 1.0);
        double sc = (double)slo + ((double)shi - (double)slo) * (double)pcg_float(&rng);
        const double x = q[0], y = q[1], z = q[2], w = q[3];
        const double rot[9] = {1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w),
                               2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w),
                               2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)};
        float *a = affines + 12 * (size_t)i;
        for (int r = 0; r < 3; r++) {
            a[4 * r + 0] = (float)(sc * rot[3 * r + 0]);
            a[4 * r + 1] = (float)(sc * rot[3 * r + 1]);
            a[4 * r + 2] = (float)(sc * rot[3 * r + 2]);
            a[4 * r + 3] = (float)tr[r];
        }
    }
}

void orc_batch_sizes(uint64_t initstate, uint64_t initseq, uint64_t skip, uint32_t count, uint32_t lo, uint32_t hi,
                     uint32_t *sizes) {
    orc_pcg32 rng;
    orc_pcg32_seed(&rng, initstate, initseq);
    orc_pcg32_advance(&rng, skip);
    const double la = std::log((double)lo), lb = std::log((double)hi);
    for (uint32_t i = 0; i < count; i++) {
        double v = std::exp(la + (lb - la) * (double)pcg_float(&rng));
        v = std::floor(v + 0.5);
        sizes[i] = (uint32_t)std::min<double>(hi, std::max<double>(lo, v));
    }
}

void *orc_mesh_build(const float *verts, uint64_t vstride_bytes, uint64_t nverts, const uint32_t *tris,
                     uint64_t ntris, int nthreads) {
    auto *m = new Mesh;
    m->v.resize(3 * nverts);
    const char *base = reinterpret_cast<const char *>(verts);
    for (uint64_t i = 0; i < nverts; i++) {
        const float *p = reinterpret_cast<const float *>(base + i * vstride_bytes);
        m->v[3 * i] = p[0]; m->v[3 * i + 1] = p[1]; m->v[3 * i + 2] = p[2];
    }
    m->tri.assign(tris, tris + 3 * ntris);
    std::vector<Box> boxes(ntris);
    parallel_for(ntris, 1u << 16, nthreads, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; i++) {
            boxes[i].empty();
            for (int c = 0; c < 3; c++) boxes[i].grow(&m->v[3 * (size_t)m->tri[3 * i + c]]);
        }
    });
    build_bvh(m->bvh, boxes, nthreads);
    m->bounds = m->bvh.nodes[0].box;
    return m;
}

void orc_mesh_free(void *mesh) { delete static_cast<Mesh *>(mesh); }

uint64_t orc_mesh_node_count(const void *mesh) { return static_cast<const Mesh *>(mesh)->bvh.nodes.size(); }

void *orc_scene_create(void) { return new Scene; }

void orc_scene_add_ex(void *scene, const void *mesh, const void *procedural, const float affine[12], uint32_t mask,
                      uint32_t opaque) {
    auto *s = static_cast<Scene *>(scene);
    Instance in;
    in.mesh = static_cast<const Mesh *>(mesh);
    in.proc = static_cast<const Procedural *>(procedural);
    std::memcpy(in.fwd, affine, sizeof(in.fwd));
    invert_affine(affine, in.inv);
    in.mask = mask;
    in.opaque = opaque != 0;
    s->inst.push_back(in);
}

void orc_scene_add(void *scene, const void *mesh, const float affine[12], uint32_t mask) {
    orc_scene_add_ex(scene, mesh, nullptr, affine, mask, 1u);
}

void orc_scene_commit(void *scene) {
    auto *s = static_cast<Scene *>(scene);
    std::vector<Box> boxes(s->inst.size());
    for (size_t i = 0; i < s->inst.size(); i++) {
        const Instance &in = s->inst[i];
        boxes[i].empty();
        if (!in.mesh || in.mesh->tri.empty()) continue; // procedural: not in the plain-trace TLAS
        const Box &b = in.mesh->bounds;
        for (int c = 0; c < 8; c++) {
            float p[3] = {(c & 1) ? b.hi[0] : b.lo[0], (c & 2) ? b.hi[1] : b.lo[1], (c & 4) ? b.hi[2] : b.lo[2]};
            float w[3];
            xform_point(in.fwd, p, w);
            boxes[i].grow(w);
        }
        // widen by a relative epsilon: the world box must contain every transformed vertex
        for (int a = 0; a < 3; a++) {
            float pad = 1e-5f * std::max({std::fabs(boxes[i].lo[a]), std::fabs(boxes[i].hi[a]), 1.0f});
            boxes[i].lo[a] -= pad;
            boxes[i].hi[a] += pad;
        }
    }
    build_bvh(s->tlas, boxes, 1);
}

void orc_scene_free(void *scene) { delete static_cast<Scene *>(scene); }

void orc_trace_closest(const void *scene, const float *rays, uint64_t nrays, uint32_t mask, orc_hit *hits, int nthreads) {
    const Scene &s = *static_cast<const Scene *>(scene);
    parallel_for(nrays, 64, nthreads, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; i++) {
            const float *r = rays + 8 * i;
            TraceState st{std::numeric_limits<float>::infinity(), std::numeric_limits<float>::infinity(),
                          UINT32_MAX, UINT32_MAX, 0.0f, 0.0f, false};
            trace_scene<false>(s, r, mask, st);
            orc_hit &h = hits[i];
            if (st.prim != UINT32_MAX) {
                h.inst = st.inst; h.prim = st.prim; h.u = st.u; h.v = st.v; h.t = st.t_best;
            } else {
                h.inst = UINT32_MAX; h.prim = UINT32_MAX; h.u = 0.0f; h.v = 0.0f; h.t = r[7];
            }
            h.pad = 0;
        }
    });
}

void orc_trace_closest2(const void *scene, const float *rays, uint64_t nrays, uint32_t mask, orc_hit2 *hits, int nthreads) {
    const Scene &s = *static_cast<const Scene *>(scene);
    parallel_for(nrays, 64, nthreads, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; i++) {
            const float *r = rays + 8 * i;
            TraceState st{std::numeric_limits<float>::infinity(), std::numeric_limits<float>::infinity(),
                          UINT32_MAX, UINT32_MAX, 0.0f, 0.0f, true};
            trace_scene<false>(s, r, mask, st);
            orc_hit2 &h = hits[i];
            if (st.prim != UINT32_MAX) {
                h.inst = st.inst; h.prim = st.prim; h.u = st.u; h.v = st.v; h.t = st.t_best;
            } else {
                h.inst = UINT32_MAX; h.prim = UINT32_MAX; h.u = 0.0f; h.v = 0.0f; h.t = r[7];
            }
            h.t2 = st.t2;
        }
    });
}

void orc_trace_any(const void *scene, const float *rays, uint64_t nrays, uint32_t mask, uint8_t *hit, int nthreads) {
    const Scene &s = *static_cast<const Scene *>(scene);
    parallel_for(nrays, 64, nthreads, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; i++) {
            TraceState st{std::numeric_limits<float>::infinity(), std::numeric_limits<float>::infinity(),
                          UINT32_MAX, UINT32_MAX, 0.0f, 0.0f, false};
            hit[i] = trace_scene<true>(s, rays + 8 * i, mask, st) ? 1 : 0;
        }
    });
}

void orc_brute_closest_f64(const float *verts, uint64_t vstride_bytes, const uint32_t *tris, uint64_t ntris,
                           const float *rays, uint64_t nrays, orc_bf_hit *out, int nthreads) {
    const char *base = reinterpret_cast<const char *>(verts);
    auto vtx = [&](uint32_t i, double p[3]) {
        const float *q = reinterpret_cast<const float *>(base + (uint64_t)i * vstride_bytes);
        p[0] = q[0]; p[1] = q[1]; p[2] = q[2];
    };
    parallel_for(nrays, 16, nthreads, [&](uint64_t b, uint64_t e) {
        for (uint64_t r = b; r < e; r++) {
            const float *ray = rays + 8 * r;
            double o[3] = {ray[0], ray[1], ray[2]}, d[3] = {ray[4], ray[5], ray[6]};
            double tmin = ray[3], tmax = ray[7];
            orc_bf_hit best{UINT32_MAX, 0, std::numeric_limits<double>::infinity(), 0, 0,
                            std::numeric_limits<double>::infinity()};
            for (uint64_t k = 0; k < ntris; k++) {
                double a[3], bb[3], c[3];
                vtx(tris[3 * k], a); vtx(tris[3 * k + 1], bb); vtx(tris[3 * k + 2], c);
                double e1[3] = {bb[0] - a[0], bb[1] - a[1], bb[2] - a[2]};
                double e2[3] = {c[0] - a[0], c[1] - a[1], c[2] - a[2]};
                double p[3] = {d[1] * e2[2] - d[2] * e2[1], d[2] * e2[0] - d[0] * e2[2], d[0] * e2[1] - d[1] * e2[0]};
                double det = e1[0] * p[0] + e1[1] * p[1] + e1[2] * p[2];
                if (det == 0.0) continue;
                double inv = 1.0 / det;
                double s[3] = {o[0] - a[0], o[1] - a[1], o[2] - a[2]};
                double u = (s[0] * p[0] + s[1] * p[1] + s[2] * p[2]) * inv;
                if (u < 0.0 || u > 1.0) continue;
                double q[3] = {s[1] * e1[2] - s[2] * e1[1], s[2] * e1[0] - s[0] * e1[2], s[0] * e1[1] - s[1] * e1[0]};
                double v = (d[0] * q[0] + d[1] * q[1] + d[2] * q[2]) * inv;
                if (v < 0.0 || u + v > 1.0) continue;
                double t = (e2[0] * q[0] + e2[1] * q[1] + e2[2] * q[2]) * inv;
                if (t < tmin || t > tmax) continue;
                if (t < best.t) {
                    best.t2 = best.t;
                    best.t = t; best.u = u; best.v = v; best.prim = (uint32_t)k;
                } else if (t < best.t2) {
                    best.t2 = t;
                }
            }
            out[r] = best;
        }
    });
}

int orc_watertight(const float o[3], const float d[3], const float v0[3], const float v1[3], const float v2[3],
                   float tmin, float tmax, float *t, float *u, float *v) {
    const unsigned int csr = _mm_getcsr();
    set_ftz_daz();
    WtRay r;
    wt_setup(r, o, d);
    const int hit = wt_intersect(r, v0, v1, v2, tmin, tmax, *t, *u, *v) ? 1 : 0;
    _mm_setcsr(csr);
    return hit;
}

} // extern "C"

// ------------------------------------------------------------------------------------------
// Path tracer: the DSL program of src/tests/test_path_tracing.cpp:165-258 restated on the
// scene above (camera ray, closest hit, triangle_interpolate, geometric normal, emission with
// the balance heuristic, one light sample through offset_ray_origin plus a shadow ray, a
// cosine-weighted bounce in the make_onb frame, Russian roulette with q = max(luminance, 0.05)),
// with the fixed choices stated in lc_oracle.h. Single precision, no contraction, one operation
// at a time in a fixed order.
namespace {

constexpr uint32_t kPtDrawsPerSample = 64;
constexpr float kPtInvPi = 0.318309886183790672f;

struct F3 {
    float x, y, z;
};
inline F3 f3(float x, float y, float z) { return F3{x, y, z}; }
inline F3 f3sub(F3 a, F3 b) { return f3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline float f3dot(F3 a, F3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline F3 f3cross(F3 a, F3 b) { return f3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
inline F3 f3normalize(F3 a) {
    const float l = std::sqrt(f3dot(a, a));
    return f3(a.x / l, a.y / l, a.z / l);
}

// sin(2 pi u), cos(2 pi u), u in [0,1): quadrant k = floor(4u) (exact), theta = (4u - k) pi/2,
// Taylor series of sin to degree 11 and cos to degree 12 (truncation below 6e-8 on [0, pi/2)).
inline void pt_sincos_2pi(float u, float &s, float &c) {
    const float x = u * 4.0f;
    const int k = (int)x;
    const float f = x - (float)k;
    const float t = f * 1.57079632679489662f;
    const float t2 = t * t;
    const float ps = t * (1.0f + t2 * (-1.0f / 6.0f + t2 * (1.0f / 120.0f + t2 * (-1.0f / 5040.0f +
                     t2 * (1.0f / 362880.0f + t2 * (-1.0f / 39916800.0f))))));
    const float pc = 1.0f + t2 * (-1.0f / 2.0f + t2 * (1.0f / 24.0f + t2 * (-1.0f / 720.0f +
                     t2 * (1.0f / 40320.0f + t2 * (-1.0f / 3628800.0f + t2 * (1.0f / 479001600.0f))))));
    if ((k & 3) == 0) { s = ps; c = pc; }
    else if ((k & 3) == 1) { s = pc; c = -ps; }
    else if ((k & 3) == 2) { s = -ps; c = -pc; }
    else { s = -pc; c = ps; }
}

// offset_ray_origin (src/dsl/rtx/ray.cpp:16-24)
inline float pt_offset1(float p, float n) {
    const int32_t of = (int32_t)(256.0f * n);
    uint32_t bits;
    std::memcpy(&bits, &p, 4);
    bits += (uint32_t)(p < 0.0f ? -of : of);
    float pi;
    std::memcpy(&pi, &bits, 4);
    return std::fabs(p) < (1.0f / 32.0f) ? p + (1.0f / 65536.0f) * n : pi;
}
inline F3 pt_offset(F3 p, F3 n) { return f3(pt_offset1(p.x, n.x), pt_offset1(p.y, n.y), pt_offset1(p.z, n.z)); }

struct PtLight {
    F3 pos, u, v, e, n;
    float area;
};

F3 pt_vertex(const Instance &in, uint32_t idx) {
    const float *p = &in.mesh->v[3 * (size_t)idx];
    float w[3];
    xform_point(in.fwd, p, w);
    return f3(w[0], w[1], w[2]);
}

// One sample of pixel (x, y); returns the path's radiance.
// Optional record of every ray a path traces (8 floats each).
struct PtRecord {
    std::vector<float> closest, shadow;
};

F3 pt_sample(const Scene &sc, const orc_pt_material *mats, const orc_pt_params &P, const PtLight &Lt, uint32_t x,
             uint32_t y, uint32_t sample, uint64_t &n_closest, uint64_t &n_shadow, PtRecord *rec = nullptr) {
    const uint32_t pixel = y * P.width + x;
    orc_pcg32 rng;
    orc_pcg32_seed(&rng, P.seed, P.sequence + pixel);
    orc_pcg32_advance(&rng, (uint64_t)kPtDrawsPerSample * sample);
    // camera ray
    const float rx = pcg_float(&rng);
    const float ry = pcg_float(&rng);
    const float frame = (float)std::min(P.width, P.height);
    const float fx = (((float)x + rx) * 2.0f - (float)P.width) / frame;
    const float fy = (((float)y + ry) * 2.0f - (float)P.height) / frame;
    F3 d = f3normalize(f3(fx * P.camera_right[0] - fy * P.camera_up[0] + P.camera_forward[0],
                          fx * P.camera_right[1] - fy * P.camera_up[1] + P.camera_forward[1],
                          fx * P.camera_right[2] - fy * P.camera_up[2] + P.camera_forward[2]));
    F3 o = f3(P.camera_origin[0], P.camera_origin[1], P.camera_origin[2]);
    F3 beta = f3(1.0f, 1.0f, 1.0f), L = f3(0.0f, 0.0f, 0.0f);
    float pdf_bsdf = 0.0f;
    for (uint32_t depth = 0; depth < P.max_depth; depth++) {
        const float ray[8] = {o.x, o.y, o.z, 0.0f, d.x, d.y, d.z, std::numeric_limits<float>::max()};
        TraceState st{std::numeric_limits<float>::infinity(), std::numeric_limits<float>::infinity(), UINT32_MAX,
                      UINT32_MAX, 0.0f, 0.0f, false};
        trace_scene<false>(sc, ray, 0xFFu, st);
        n_closest++;
        if (rec) rec->closest.insert(rec->closest.end(), ray, ray + 8);
        if (st.prim == UINT32_MAX) break;
        const Instance &in = sc.inst[st.inst];
        const uint32_t *tri = &in.mesh->tri[3 * (size_t)st.prim];
        const F3 p0 = pt_vertex(in, tri[0]), p1 = pt_vertex(in, tri[1]), p2 = pt_vertex(in, tri[2]);
        const float bu = st.u, bv = st.v;
        const float w0 = 1.0f - bu - bv;
        const F3 p = f3(w0 * p0.x + bu * p1.x + bv * p2.x, w0 * p0.y + bu * p1.y + bv * p2.y,
                        w0 * p0.z + bu * p1.z + bv * p2.z);
        const F3 n = f3normalize(f3cross(f3sub(p1, p0), f3sub(p2, p0)));
        const float cos_wo = -d.x * n.x + -d.y * n.y + -d.z * n.z;
        if (cos_wo < 1e-4f) break;
        const orc_pt_material &mat = mats[st.inst];
        if (mat.emissive) {
            if (depth == 0) {
                L = f3(L.x + Lt.e.x, L.y + Lt.e.y, L.z + Lt.e.z);
            } else {
                const F3 dp = f3sub(p, o);
                const float pdf_light = f3dot(dp, dp) / (Lt.area * cos_wo);
                const float mis = pdf_bsdf / std::fmax(pdf_bsdf + pdf_light, 1e-4f);
                L = f3(L.x + mis * beta.x * Lt.e.x, L.y + mis * beta.y * Lt.e.y, L.z + mis * beta.z * Lt.e.z);
            }
            break;
        }
        const float ulx = pcg_float(&rng);
        const float uly = pcg_float(&rng);
        const F3 pl = f3(Lt.pos.x + ulx * Lt.u.x + uly * Lt.v.x, Lt.pos.y + ulx * Lt.u.y + uly * Lt.v.y,
                         Lt.pos.z + ulx * Lt.u.z + uly * Lt.v.z);
        const F3 pp = pt_offset(p, n);
        const F3 ppl = pt_offset(pl, Lt.n);
        const F3 dl = f3sub(ppl, pp);
        const float d_light = std::sqrt(f3dot(dl, dl));
        const F3 wi = f3normalize(dl);
        const float cos_wi_light = f3dot(wi, n);
        const float cos_light = -f3dot(Lt.n, wi);
        const F3 albedo = f3(mat.albedo[0], mat.albedo[1], mat.albedo[2]);
        if (cos_wi_light > 1e-4f && cos_light > 1e-4f) {
            const float pdf_light = (d_light * d_light) / (Lt.area * cos_light);
            const float pdf_b = cos_wi_light * kPtInvPi;
            const float mis = pdf_light / std::fmax(pdf_light + pdf_b, 1e-4f);
            const float den = std::fmax(pdf_light, 1e-4f);
            const F3 c = f3(beta.x * (albedo.x * kPtInvPi * cos_wi_light) * mis * Lt.e.x / den,
                            beta.y * (albedo.y * kPtInvPi * cos_wi_light) * mis * Lt.e.y / den,
                            beta.z * (albedo.z * kPtInvPi * cos_wi_light) * mis * Lt.e.z / den);
            const F3 so = pt_offset(pp, n);
            const float sray[8] = {so.x, so.y, so.z, 0.0f, wi.x, wi.y, wi.z, d_light};
            TraceState sst{std::numeric_limits<float>::infinity(), std::numeric_limits<float>::infinity(), UINT32_MAX,
                           UINT32_MAX, 0.0f, 0.0f, false};
            const bool occluded = trace_scene<true>(sc, sray, 0xFFu, sst);
            n_shadow++;
            if (rec) rec->shadow.insert(rec->shadow.end(), sray, sray + 8);
            if (!occluded) L = f3(L.x + c.x, L.y + c.y, L.z + c.z);
        }
        const float ux = pcg_float(&rng);
        const float uy = pcg_float(&rng);
        const float r = std::sqrt(ux);
        float sp, cp;
        pt_sincos_2pi(uy, sp, cp);
        const F3 wl = f3(r * cp, r * sp, std::sqrt(1.0f - ux));
        const F3 bn = f3normalize(std::fabs(n.x) > std::fabs(n.z) ? f3(-n.y, n.x, 0.0f) : f3(0.0f, -n.z, n.y));
        const F3 tn = f3normalize(f3cross(bn, n));
        d = f3(wl.x * tn.x + wl.y * bn.x + wl.z * n.x, wl.x * tn.y + wl.y * bn.y + wl.z * n.y,
               wl.x * tn.z + wl.y * bn.z + wl.z * n.z);
        o = pp;
        pdf_bsdf = std::fabs(wl.z) * kPtInvPi;
        beta = f3(beta.x * albedo.x, beta.y * albedo.y, beta.z * albedo.z);
        if (depth >= P.rr_depth) {
            const float l = 0.212671f * beta.x + 0.715160f * beta.y + 0.072169f * beta.z;
            if (l == 0.0f) break;
            const float q = std::fmax(l, 0.05f);
            const float u_rr = pcg_float(&rng);
            if (u_rr >= q) break;
            const float iq = 1.0f / q;
            beta = f3(beta.x * iq, beta.y * iq, beta.z * iq);
        }
    }
    if (std::isnan(L.x) || std::isnan(L.y) || std::isnan(L.z)) L = f3(0.0f, 0.0f, 0.0f);
    return L;
}

} // namespace

extern "C" void orc_sincos_2pi(float u, float *s, float *c) { pt_sincos_2pi(u, *s, *c); }

extern "C" void orc_offset_ray_origin(const float p[3], const float n[3], float out[3]) {
    const F3 o = pt_offset(f3(p[0], p[1], p[2]), f3(n[0], n[1], n[2]));
    out[0] = o.x; out[1] = o.y; out[2] = o.z;
}

extern "C" void orc_pathtrace_rays(const void *scene, const orc_pt_material *mats, const orc_pt_params *params,
                                   uint64_t pixel_begin, uint64_t pixel_end, float *closest, uint64_t closest_cap,
                                   uint64_t *closest_count, float *shadow, uint64_t shadow_cap, uint64_t *shadow_count) {
    const unsigned int csr = _mm_getcsr();
    set_ftz_daz();
    const Scene &sc = *static_cast<const Scene *>(scene);
    const orc_pt_params P = *params;
    PtLight Lt;
    Lt.pos = f3(P.light_position[0], P.light_position[1], P.light_position[2]);
    Lt.u = f3(P.light_u[0], P.light_u[1], P.light_u[2]);
    Lt.v = f3(P.light_v[0], P.light_v[1], P.light_v[2]);
    Lt.e = f3(P.light_emission[0], P.light_emission[1], P.light_emission[2]);
    const F3 c = f3cross(Lt.u, Lt.v);
    Lt.area = std::sqrt(f3dot(c, c));
    Lt.n = f3(c.x / Lt.area, c.y / Lt.area, c.z / Lt.area);
    PtRecord rec;
    uint64_t nc = 0, ns = 0;
    const uint64_t npix = std::min<uint64_t>(pixel_end, (uint64_t)P.width * P.height);
    for (uint64_t pixel = pixel_begin; pixel < npix; pixel++)
        for (uint32_t s = 0; s < P.spp; s++)
            pt_sample(sc, mats, P, Lt, (uint32_t)(pixel % P.width), (uint32_t)(pixel / P.width), P.sample_offset + s, nc,
                      ns, &rec);
    _mm_setcsr(csr);
    *closest_count = rec.closest.size() / 8;
    *shadow_count = rec.shadow.size() / 8;
    std::memcpy(closest, rec.closest.data(), sizeof(float) * std::min<uint64_t>(rec.closest.size(), 8 * closest_cap));
    std::memcpy(shadow, rec.shadow.data(), sizeof(float) * std::min<uint64_t>(rec.shadow.size(), 8 * shadow_cap));
}

extern "C" void orc_pathtrace(const void *scene, const orc_pt_material *mats, const orc_pt_params *params,
                              uint64_t pixel_begin, uint64_t pixel_end, float *image, uint64_t counts[2],
                              int nthreads) {
    const Scene &sc = *static_cast<const Scene *>(scene);
    const orc_pt_params P = *params;
    PtLight Lt;
    Lt.pos = f3(P.light_position[0], P.light_position[1], P.light_position[2]);
    Lt.u = f3(P.light_u[0], P.light_u[1], P.light_u[2]);
    Lt.v = f3(P.light_v[0], P.light_v[1], P.light_v[2]);
    Lt.e = f3(P.light_emission[0], P.light_emission[1], P.light_emission[2]);
    const F3 c = f3cross(Lt.u, Lt.v);
    Lt.area = std::sqrt(f3dot(c, c));
    Lt.n = f3(c.x / Lt.area, c.y / Lt.area, c.z / Lt.area);
    std::atomic<uint64_t> nc{0}, ns{0};
    const uint64_t npix = std::min<uint64_t>(pixel_end, (uint64_t)P.width * P.height);
    if (pixel_begin >= npix) return;
    parallel_for(npix - pixel_begin, 16, nthreads, [&](uint64_t b, uint64_t e) {
        uint64_t n_closest = 0, n_shadow = 0;
        for (uint64_t k = b; k < e; k++) {
            const uint64_t pixel = pixel_begin + k;
            const uint32_t x = (uint32_t)(pixel % P.width), y = (uint32_t)(pixel / P.width);
            F3 sum = f3(0.0f, 0.0f, 0.0f);
            for (uint32_t s = 0; s < P.spp; s++) {
                const F3 L = pt_sample(sc, mats, P, Lt, x, y, P.sample_offset + s, n_closest, n_shadow);
                sum = f3(sum.x + L.x, sum.y + L.y, sum.z + L.z);
            }
            const float fs = (float)P.spp;
            float *out = image + 4 * pixel;
            out[0] = sum.x / fs;
            out[1] = sum.y / fs;
            out[2] = sum.z / fs;
            out[3] = 1.0f;
        }
        nc += n_closest;
        ns += n_shadow;
    });
    if (counts) {
        counts[0] += nc.load();
        counts[1] += ns.load();
    }
}

// ------------------------------------------------------------------------------------------
// Ray queries (AccelImpl::ray_query, src/rust/luisa_compute_backend_impl/src/cpu/accel.rs:582-800)
// with the built-in candidate callbacks (src/tests/test_procedural.cpp:86-120), restated by
// enumerating every candidate of every instance whose mask passes (no BVH): a surface candidate
// is a watertight hit with t in [tmin, tmax], committed for opaque instances, otherwise per the
// surface callback (cpu/accel.rs:378-393 disable the filter for opaque instances); a procedural
// candidate commits the callback's distance if it lies in [tmin, tmax) (cpu/accel.rs:711-717).
// The query's answer is the nearest committed candidate, ties to the smaller (instance, primitive).
namespace {

inline float q_len2(float a, float b) { return std::sqrt(a * a + b * b); }

bool q_surface(const orc_query_policy &P, float u, float v) {
    if (P.surface_mode == 0u) return true;  // commit
    if (P.surface_mode != 2u) return false; // ignore
    const float w = (1.0f - u) - v;          // cutout
    return q_len2(w, u) < P.cutout && q_len2(u, v) < P.cutout && q_len2(v, w) < P.cutout;
}

// sphere inscribed in the AABB (centre = box centre, radius = half the smallest extent), met by
// the object-space ray o + t d: L = c - o, tc = dot(L, d) / dot(d, d) > 0, squared distance of the
// ray from the centre dot(L, L) - tc^2 dot(d, d) <= r^2, t = tc - sqrt((r^2 - that) / dot(d, d))
bool q_procedural(const orc_query_policy &P, const float *box, const float o[3], const float d[3], float &t) {
    if (P.procedural_mode != 1u) return false;
    const float c[3] = {(box[0] + box[3]) * 0.5f, (box[1] + box[4]) * 0.5f, (box[2] + box[5]) * 0.5f};
    const float r = 0.5f * std::fmin(std::fmin(box[3] - box[0], box[4] - box[1]), box[5] - box[2]);
    const float L[3] = {c[0] - o[0], c[1] - o[1], c[2] - o[2]};
    const float dd = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2];
    const float ld = (L[0] * d[0] + L[1] * d[1]) + L[2] * d[2];
    const float ll = (L[0] * L[0] + L[1] * L[1]) + L[2] * L[2];
    const float tc = ld / dd;
    if (!(tc > 0.0f)) return false;
    const float q = ll - (tc * tc) * dd;
    const float rr = r * r;
    if (!(q <= rr)) return false;
    t = tc - std::sqrt((rr - q) / dd);
    return true;
}

// Evaluates candidate `prim` of instance `id`; true with (t, u, v) when it is committed.
bool q_candidate(const Instance &in, const orc_query_policy &P, const float *ray, uint32_t prim, float &t, float &u,
                 float &v) {
    const float o[3] = {ray[0], ray[1], ray[2]}, d[3] = {ray[4], ray[5], ray[6]};
    const float tmin = ray[3], tmax = ray[7];
    float lo[3], ld[3];
    xform_point(in.inv, o, lo);
    xform_vector(in.inv, d, ld);
    if (in.proc) {
        if (!q_procedural(P, &in.proc->aabb[6 * (size_t)prim], lo, ld, t)) return false;
        u = v = 0.0f;
        return t >= tmin && t < tmax;
    }
    WtRay wr;
    wt_setup(wr, lo, ld);
    const uint32_t *ix = &in.mesh->tri[3 * (size_t)prim];
    if (!wt_intersect(wr, &in.mesh->v[3 * (size_t)ix[0]], &in.mesh->v[3 * (size_t)ix[1]],
                      &in.mesh->v[3 * (size_t)ix[2]], tmin, tmax, t, u, v))
        return false;
    return in.opaque || q_surface(P, u, v);
}

uint32_t q_prim_count(const Instance &in) {
    return in.proc ? (uint32_t)(in.proc->aabb.size() / 6) : (uint32_t)(in.mesh ? in.mesh->tri.size() / 3 : 0);
}

} // namespace

extern "C" {

void *orc_procedural_create(const float *aabbs, uint64_t n) {
    auto *p = new Procedural;
    p->aabb.assign(aabbs, aabbs + 6 * n);
    return p;
}

void orc_procedural_free(void *procedural) { delete static_cast<Procedural *>(procedural); }

void orc_query(const void *scene, const float *rays, uint64_t nrays, uint32_t mask, const orc_query_policy *policy,
               orc_committed_hit *out, int nthreads) {
    const Scene &s = *static_cast<const Scene *>(scene);
    const orc_query_policy P = *policy;
    parallel_for(nrays, 16, nthreads, [&](uint64_t b, uint64_t e) {
        for (uint64_t i = b; i < e; i++) {
            const float *r = rays + 8 * i;
            orc_committed_hit best{UINT32_MAX, UINT32_MAX, 0.0f, 0.0f, 0u, r[7]};
            float t_best = std::numeric_limits<float>::infinity();
            for (uint32_t id = 0; id < (uint32_t)s.inst.size(); id++) {
                const Instance &in = s.inst[id];
                if ((in.mask & mask & 0xffu) == 0u) continue;
                const uint32_t np = q_prim_count(in);
                for (uint32_t p = 0; p < np; p++) {
                    float t, u, v;
                    if (!q_candidate(in, P, r, p, t, u, v)) continue;
                    const bool nearer = t != t_best ? t < t_best : (id != best.inst ? id < best.inst : p < best.prim);
                    if (!nearer) continue;
                    t_best = t;
                    best = orc_committed_hit{id, p, u, v, in.proc ? 2u : 1u, t};
                }
            }
            out[i] = best;
        }
    });
}

int orc_query_candidate(const void *scene, const float *ray, uint32_t mask, const orc_query_policy *policy,
                        uint32_t inst, uint32_t prim, uint32_t hit_type, float *t, float *u, float *v) {
    const Scene &s = *static_cast<const Scene *>(scene);
    if (inst >= s.inst.size()) return 0;
    const Instance &in = s.inst[inst];
    if ((in.mask & mask & 0xffu) == 0u || prim >= q_prim_count(in)) return 0;
    if (hit_type != (in.proc ? 2u : 1u)) return 0;
    const unsigned int csr = _mm_getcsr();
    set_ftz_daz();
    const bool ok = q_candidate(in, *policy, ray, prim, *t, *u, *v);
    _mm_setcsr(csr);
    return ok ? 1 : 0;
}

} // extern "C"
