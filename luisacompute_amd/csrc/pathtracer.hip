// pathtracer.hip -- the box-room path tracer on gfx950, in two forms over the same shading code:
//
//   megakernel : one persistent launch; a wave takes an 8x8 pixel tile, a lane runs its pixel's
//                samples one after another, tracing through the Walker of traverse.h;
//   wavefront  : samples are rendered in waves of paths; a wave holds up to kMaxWaveJobs jobs
//                (an image's sample range each: one render's chunk, or several requests of
//                the serving loop). A path's state (ray, beta, pdf, radiance, RNG, path id)
//                travels inside the queue entry, so every launch reads and writes its queue
//                sequentially. Per bounce: a persistent closest-hit launch over the queue (the
//                hit carries its BVH leaf record, so shading reads one 48-B record instead of
//                gathering indices and vertices), a shading launch that appends continuation
//                and shadow rays to compacted queues (one atomic per block iteration), a
//                persistent any-hit launch over the shadow queue, a resolve launch, and
//                optionally a bucket sort that physically reorders the continuation queue by
//                origin cell and direction octant. Queue lengths and the wave's job table stay
//                in device memory, so a wave's launches are captured once in a hipGraph and
//                replayed for every wave.
//
// The integrator restates the DSL program of src/tests/test_path_tracing.cpp:165-258
// (camera ray, closest hit, triangle_interpolate, geometric normal, emission with the balance
// heuristic, one light sample through offset_ray_origin with a shadow ray, cosine-weighted
// hemisphere sampling in the make_onb frame, Russian roulette with q = max(luminance, 0.05)),
// with these fixed choices: a new pixel jitter per sample, Russian roulette from bounce
// rr_depth on, at most max_depth closest-hit rays per path, shadow rays traced only when the
// light sample can contribute (the reference traces them and drops the result), and PCG32
// per pixel (stream initseq = sequence + pixel, sample s starting 64 s draws into the stream).
// make_ray / offset_ray_origin / triangle_interpolate follow src/dsl/rtx/ray.cpp:7-24 and
// src/dsl/rtx/hit.cpp:15-17. All arithmetic is IEEE single precision without contraction, with
// sine and cosine from fixed polynomials, so an image is the same bit for bit in both forms and
// in the CPU restatement.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cstdint>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

#include "lc_hip_bvh.h"
#include "lc_hip_internal.h"
#include "lc_hip_rtx.h"
#include "traverse.h"

namespace lchip {
namespace pt {

using namespace rtx;

constexpr uint32_t kDrawsPerSample = 64; // RNG draws reserved per sample: 2 + 5 per bounce
constexpr float kInvPi = 0.318309886183790672f;
constexpr uint32_t kMiss = 0xFFFFFFFFu;
constexpr uint32_t kEnded = 0x80000000u; // shadow target: the path ended (index into the final radiance)

// ---- PCG32 (O'Neill 2014: pcg32_srandom_r, pcg32_random_r; jump-ahead of Brown 1994) ----
struct Pcg {
    uint64_t state, inc;
};
__device__ __forceinline__ uint32_t pcg_next(Pcg &r) {
    const uint64_t old = r.state;
    r.state = old * 6364136223846793005ull + r.inc;
    const uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    const uint32_t rot = (uint32_t)(old >> 59u);
    return (xs >> rot) | (xs << ((0u - rot) & 31u));
}
__device__ __forceinline__ void pcg_seed(Pcg &r, uint64_t initstate, uint64_t initseq) {
    r.state = 0u;
    r.inc = (initseq << 1u) | 1u;
    pcg_next(r);
    r.state += initstate;
    pcg_next(r);
}
__device__ __forceinline__ void pcg_advance(Pcg &r, uint64_t delta) {
    uint64_t cur_mult = 6364136223846793005ull, cur_plus = r.inc, acc_mult = 1u, acc_plus = 0u;
    while (delta > 0u) {
        if (delta & 1u) {
            acc_mult *= cur_mult;
            acc_plus = acc_plus * cur_mult + cur_plus;
        }
        cur_plus = (cur_mult + 1u) * cur_plus;
        cur_mult *= cur_mult;
        delta >>= 1u;
    }
    r.state = acc_mult * r.state + acc_plus;
}
__device__ __forceinline__ float pcg_float(Pcg &r) { return (float)(pcg_next(r) >> 8) * 0x1p-24f; }

// sin(2 pi u), cos(2 pi u) for u in [0, 1): quadrant from 4u (exact), then Taylor polynomials of
// degree 11 / 12 on [0, pi/2) (truncation error below 6e-8).
__device__ __forceinline__ void sincos_2pi(float u, float &s, float &c) {
    const float x = u * 4.0f;
    const int k = (int)x;
    const float f = x - (float)k;
    const float t = f * 1.57079632679489662f;
    const float t2 = t * t;
    const float ps = t * (1.0f + t2 * (-1.0f / 6.0f + t2 * (1.0f / 120.0f + t2 * (-1.0f / 5040.0f +
                     t2 * (1.0f / 362880.0f + t2 * (-1.0f / 39916800.0f))))));
    const float pc = 1.0f + t2 * (-1.0f / 2.0f + t2 * (1.0f / 24.0f + t2 * (-1.0f / 720.0f +
                     t2 * (1.0f / 40320.0f + t2 * (-1.0f / 3628800.0f + t2 * (1.0f / 479001600.0f))))));
    switch (k & 3) {
    case 0: s = ps; c = pc; break;
    case 1: s = pc; c = -ps; break;
    case 2: s = -ps; c = -pc; break;
    default: s = -pc; c = ps; break;
    }
}

__device__ __forceinline__ V3 to_world(const float *m, float4 v) {
    return v3(m[0] * v.x + m[1] * v.y + m[2] * v.z + m[3], m[4] * v.x + m[5] * v.y + m[6] * v.z + m[7],
              m[8] * v.x + m[9] * v.y + m[10] * v.z + m[11]);
}

// Camera ray through pixel (x, y) with a jitter of two draws; the film spans [-1, 1] along the
// shorter image side: direction = normalize(fx * right - fy * up + forward).
__device__ __forceinline__ void camera_ray(const PtScene &S, uint32_t width, uint32_t height, uint32_t x, uint32_t y,
                                           Pcg &rng, V3 &o, V3 &d) {
    const float rx = pcg_float(rng);
    const float ry = pcg_float(rng);
    const float frame = (float)(width < height ? width : height);
    const float fx = (((float)x + rx) * 2.0f - (float)width) / frame;
    const float fy = (((float)y + ry) * 2.0f - (float)height) / frame;
    const V3 w = v3(fx * S.cam_r[0] - fy * S.cam_u[0] + S.cam_f[0], fx * S.cam_r[1] - fy * S.cam_u[1] + S.cam_f[1],
                    fx * S.cam_r[2] - fy * S.cam_u[2] + S.cam_f[2]);
    d = normalize(w);
    o = v3(S.cam_o[0], S.cam_o[1], S.cam_o[2]);
}

// What one shading step leaves to do.
struct Bounce {
    bool shadow; // trace (so, 0, wi, d_light); add contrib to L when unoccluded
    bool cont;   // continue with the ray (no, 0, nd, FLT_MAX)
    V3 so, wi;
    float d_light;
    V3 contrib;
    V3 no, nd;
};

// One path vertex at bounce `depth` (0: the camera ray's hit); `rec` is the hit triangle's leaf
// record in the instance's BLAS. Emission is added to L here; the light sample's contribution
// is returned for the caller to add after the shadow ray.
__device__ __forceinline__ void shade(const PtScene &S, uint32_t depth, V3 ro, V3 rd, uint32_t inst, uint32_t rec,
                                      float bu, float bv, Pcg &rng, V3 &beta, float &pdf_bsdf, V3 &L, Bounce &out) {
    out.shadow = false;
    out.cont = false;
    if (inst == kMiss) return;
    const PtInstance &I = S.inst[inst];
    const float4 *tr = reinterpret_cast<const float4 *>(I.tris + rec);
    const float4 a = tr[0], b = tr[1], c = tr[2];
    const V3 p0 = to_world(I.fwd, a), p1 = to_world(I.fwd, b), p2 = to_world(I.fwd, c);
    const V3 p = triangle_interpolate(bu, bv, p0, p1, p2);
    const V3 n = normalize(cross(sub(p1, p0), sub(p2, p0)));
    const float cos_wo = -rd.x * n.x + -rd.y * n.y + -rd.z * n.z;
    if (cos_wo < 1e-4f) return;
    if (I.emissive) {
        if (depth == 0u) {
            L = v3(L.x + S.light_e[0], L.y + S.light_e[1], L.z + S.light_e[2]);
        } else {
            const V3 dp = sub(p, ro);
            const float pdf_light = dot(dp, dp) / (S.light_area * cos_wo);
            const float mis = pdf_bsdf / fmaxf(pdf_bsdf + pdf_light, 1e-4f);
            L = v3(L.x + mis * beta.x * S.light_e[0], L.y + mis * beta.y * S.light_e[1],
                   L.z + mis * beta.z * S.light_e[2]);
        }
        return;
    }
    // light sample
    const float ulx = pcg_float(rng);
    const float uly = pcg_float(rng);
    const V3 pl = v3(S.light_pos[0] + ulx * S.light_u[0] + uly * S.light_v[0],
                     S.light_pos[1] + ulx * S.light_u[1] + uly * S.light_v[1],
                     S.light_pos[2] + ulx * S.light_u[2] + uly * S.light_v[2]);
    const V3 ln = v3(S.light_n[0], S.light_n[1], S.light_n[2]);
    const V3 pp = offset_ray_origin(p, n);
    const V3 ppl = offset_ray_origin(pl, ln);
    const V3 dl = sub(ppl, pp);
    const float d_light = sqrtf(dot(dl, dl));
    const V3 wi = normalize(dl);
    const float cos_wi_light = dot(wi, n);
    const float cos_light = -dot(ln, wi);
    const V3 albedo = v3(I.albedo[0], I.albedo[1], I.albedo[2]);
    if (cos_wi_light > 1e-4f && cos_light > 1e-4f) {
        const float pdf_light = (d_light * d_light) / (S.light_area * cos_light);
        const float pdf_b = cos_wi_light * kInvPi;
        const float mis = pdf_light / fmaxf(pdf_light + pdf_b, 1e-4f);
        const float den = fmaxf(pdf_light, 1e-4f);
        out.contrib = v3(beta.x * (albedo.x * kInvPi * cos_wi_light) * mis * S.light_e[0] / den,
                         beta.y * (albedo.y * kInvPi * cos_wi_light) * mis * S.light_e[1] / den,
                         beta.z * (albedo.z * kInvPi * cos_wi_light) * mis * S.light_e[2] / den);
        out.so = offset_ray_origin(pp, n);
        out.wi = wi;
        out.d_light = d_light;
        out.shadow = true;
    }
    // cosine-weighted hemisphere sample in the make_onb frame
    const float ux = pcg_float(rng);
    const float uy = pcg_float(rng);
    const float r = sqrtf(ux);
    float sp, cp;
    sincos_2pi(uy, sp, cp);
    const V3 wl = v3(r * cp, r * sp, sqrtf(1.0f - ux));
    const V3 bn = normalize(fabsf(n.x) > fabsf(n.z) ? v3(-n.y, n.x, 0.0f) : v3(0.0f, -n.z, n.y));
    const V3 tn = normalize(cross(bn, n));
    out.nd = v3(wl.x * tn.x + wl.y * bn.x + wl.z * n.x, wl.x * tn.y + wl.y * bn.y + wl.z * n.y,
                wl.x * tn.z + wl.y * bn.z + wl.z * n.z);
    out.no = pp;
    pdf_bsdf = fabsf(wl.z) * kInvPi;
    beta = v3(beta.x * albedo.x, beta.y * albedo.y, beta.z * albedo.z);
    if (depth >= S.rr_depth) {
        const float l = 0.212671f * beta.x + 0.715160f * beta.y + 0.072169f * beta.z;
        if (l == 0.0f) return;
        const float q = fmaxf(l, 0.05f);
        const float rr = pcg_float(rng);
        if (rr >= q) return;
        const float iq = 1.0f / q;
        beta = v3(beta.x * iq, beta.y * iq, beta.z * iq);
    }
    out.cont = depth + 1u < S.max_depth;
}

__device__ __forceinline__ bool has_nan(float x, float y, float z) { return isnan(x) || isnan(y) || isnan(z); }

// ---------------------------------------------------------------------------------------------
// megakernel
template<bool TWO, int BLOCK, int SLDS>
__global__ __launch_bounds__(BLOCK) void k_pt_mega(const PtScene S, const Geometry G, float4 *image,
                                                   uint32_t *tile_counter, uint2 *spill, uint32_t spill_depth,
                                                   uint32_t *error, unsigned long long *counts) {
    __shared__ uint2 stk[SLDS][BLOCK];
    const uint32_t tid = threadIdx.x, lane = tid & 63u;
    const uint64_t glane = (uint64_t)blockIdx.x * BLOCK + tid;
    trav::Walker<false, TWO, false, BLOCK, SLDS> Wc;
    trav::Walker<true, TWO, false, BLOCK, SLDS> Wa;
    Wc.init(G, stk, tid, spill + glane * spill_depth, spill_depth, error);
    Wa.init(G, stk, tid, spill + glane * spill_depth, spill_depth, error);
    const uint32_t tiles_x = (S.width + 7u) / 8u, tiles = tiles_x * ((S.height + 7u) / 8u);
    unsigned long long n_closest = 0, n_shadow = 0;
    for (;;) {
        uint32_t tile = 0;
        if (lane == 0) tile = atomicAdd(tile_counter, 1u);
        tile = __shfl(tile, 0, 64);
        if (tile >= tiles) break;
        const uint32_t x = (tile % tiles_x) * 8u + (lane & 7u), y = (tile / tiles_x) * 8u + (lane >> 3);
        if (x < S.width && y < S.height) {
            const uint32_t pixel = y * S.width + x;
            V3 sum = v3(0.0f, 0.0f, 0.0f);
            for (uint32_t s = 0; s < S.spp; s++) {
                Pcg rng;
                pcg_seed(rng, S.seed, S.sequence + pixel);
                pcg_advance(rng, (uint64_t)kDrawsPerSample * (S.sample_offset + s));
                V3 o, d;
                camera_ray(S, S.width, S.height, x, y, rng, o, d);
                V3 beta = v3(1.0f, 1.0f, 1.0f), L = v3(0.0f, 0.0f, 0.0f);
                float pdf = 0.0f;
                for (uint32_t depth = 0; depth < S.max_depth; depth++) {
                    Wc.start(G, o.x, o.y, o.z, 0.0f, d.x, d.y, d.z, FLT_MAX, 0xFFu);
                    while (!Wc.run(G, 64)) {
                    }
                    n_closest++;
                    Bounce b;
                    shade(S, depth, o, d, Wc.b_inst, Wc.b_rec, Wc.bu, Wc.bv, rng, beta, pdf, L, b);
                    if (b.shadow) {
                        Wa.start(G, b.so.x, b.so.y, b.so.z, 0.0f, b.wi.x, b.wi.y, b.wi.z, b.d_light, 0xFFu);
                        while (!Wa.run(G, 64)) {
                        }
                        n_shadow++;
                        if (!Wa.any_hit) L = v3(L.x + b.contrib.x, L.y + b.contrib.y, L.z + b.contrib.z);
                    }
                    if (!b.cont) break;
                    o = b.no;
                    d = b.nd;
                }
                if (has_nan(L.x, L.y, L.z)) L = v3(0.0f, 0.0f, 0.0f);
                sum = v3(sum.x + L.x, sum.y + L.y, sum.z + L.z);
            }
            const float fs = (float)S.spp;
            image[pixel] = make_float4(sum.x / fs, sum.y / fs, sum.z / fs, 1.0f);
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        n_closest += __shfl_xor(n_closest, o, 64);
        n_shadow += __shfl_xor(n_shadow, o, 64);
    }
    if (lane == 0) {
        atomicAdd(counts, n_closest);
        atomicAdd(counts + 1, n_shadow);
    }
}

// ---------------------------------------------------------------------------------------------
// wavefront
struct Queue {         // one ray queue, one array per field (coalesced per field)
    float4 *ray;       // 2 per entry: origin, tmin / direction, tmax
    float4 *beta;      // beta.xyz, pdf_bsdf
    float4 *rad;       // radiance so far .xyz, path id bits in .w
    uint4 *rng;        // PCG32 state (x, y) and increment (z, w)
    uint16_t *key;     // sort bucket (continuation queues)
};

struct Queues {
    Queue q[3];        // rotating: shade reads one, appends to the next, the sort writes the spare
    uint2 *hits;       // closest hits of the current queue (leaf record in the padding word)
    float4 *sr;        // shadow rays, 2 per entry
    float4 *sc;        // contribution .xyz; .w bits: continuation slot, or path | kEnded
    uint8_t *occ;      // any-hit results of the shadow queue
    float4 *Lf;        // final radiance per path of the wave
    uint32_t *hist;    // [bin][block] counts, then offsets
    uint32_t *bin_total, *bin_base;
    uint32_t *ctr;     // [0..2] queue lengths, [3] shadow queue length, [4] megakernel tiles,
                       // [kCtrClosest..) / [kCtrAny..) the trace launches' segment counters
};
constexpr uint32_t kCtrClosest = 64, kCtrAny = kCtrClosest + kRayCounterWords;
constexpr uint32_t kCtrWords = kCtrAny + kRayCounterWords;

constexpr uint32_t kBins = 4096;     // 9-bit origin cell (3 bits per axis, Morton order) x octant
constexpr uint32_t kSortBlocks = 1024;

struct SortBox {
    float lo[3], scale[3]; // cell = (o - lo) * scale, 8 cells per axis
};

__device__ __forceinline__ uint32_t spread3(uint32_t v) { // 3 bits -> positions 0, 3, 6
    return (v & 1u) | ((v & 2u) << 2) | ((v & 4u) << 4);
}

__device__ __forceinline__ uint32_t bucket_of(V3 o, V3 d, const SortBox &B) {
    const float oo[3] = {o.x, o.y, o.z};
    uint32_t c[3];
    for (int a = 0; a < 3; a++) {
        float f = (oo[a] - B.lo[a]) * B.scale[a];
        f = fminf(fmaxf(f, 0.0f), 7.0f);
        c[a] = (uint32_t)f;
    }
    const uint32_t oct = (d.x < 0.0f ? 1u : 0u) | (d.y < 0.0f ? 2u : 0u) | (d.z < 0.0f ? 4u : 0u);
    return ((spread3(c[0]) | (spread3(c[1]) << 1) | (spread3(c[2]) << 2)) << 3) | oct;
}

// Last k in [0, n) with key[k] <= v (key ascending, key[0] = 0).
__device__ __forceinline__ uint32_t search_le(const uint32_t *key, uint32_t n, uint32_t v) {
    uint32_t lo = 0;
    for (uint32_t step = kMaxWaveJobs / 2; step > 0; step >>= 1)
        if (lo + step < n && key[lo + step] <= v) lo += step;
    return lo;
}

constexpr int kGridBlock = 256;

// Path p of the wave is, within its job, sample sample0 + local / npix of pixel local % npix.
__global__ __launch_bounds__(kGridBlock) void k_pt_camera(const PtScene S, const PtWaveTable *T, const Queues Q) {
    __shared__ uint32_t base[kMaxWaveJobs];
    const uint32_t njobs = T->njobs, paths = T->paths;
    for (uint32_t k = threadIdx.x; k < njobs; k += blockDim.x) base[k] = T->jobs[k].path_base;
    __syncthreads();
    if (blockIdx.x == 0 && threadIdx.x == 0) Q.ctr[0] = paths;
    const Queue &A = Q.q[0];
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < paths; p += gridDim.x * blockDim.x) {
        const PtJob &J = T->jobs[search_le(base, njobs, p)];
        const uint32_t npix = J.width * J.height;
        const uint32_t local = p - J.path_base;
        const uint32_t s_local = local / npix, pixel = local % npix;
        const uint32_t x = pixel % J.width, y = pixel / J.width;
        Pcg rng;
        pcg_seed(rng, J.seed, J.sequence + pixel);
        pcg_advance(rng, (uint64_t)kDrawsPerSample * (J.sample0 + s_local));
        V3 o, d;
        camera_ray(S, J.width, J.height, x, y, rng, o, d);
        A.ray[2 * (uint64_t)p] = make_float4(o.x, o.y, o.z, 0.0f);
        A.ray[2 * (uint64_t)p + 1] = make_float4(d.x, d.y, d.z, FLT_MAX);
        A.beta[p] = make_float4(1.0f, 1.0f, 1.0f, 0.0f);
        A.rad[p] = make_float4(0.0f, 0.0f, 0.0f, __uint_as_float(p));
        A.rng[p] = make_uint4((uint32_t)rng.state, (uint32_t)(rng.state >> 32), (uint32_t)rng.inc,
                              (uint32_t)(rng.inc >> 32));
    }
}

// Appends are reserved once per block iteration (one atomic per queue for kShadeBlock items):
// a single counter address serialises its atomics, so per-wave reservations would dominate.
constexpr int kShadeBlock = 512;
constexpr int kShadeWaves = kShadeBlock / 64;

__global__ __launch_bounds__(kShadeBlock) void k_pt_shade(const PtScene S, uint32_t depth, const Queues Q,
                                                          uint32_t cur, uint32_t nxt, int keys, SortBox B) {
    __shared__ uint32_t s_off[2][2][kShadeWaves]; // [buffer][continuation, shadow][wave] offsets
    __shared__ uint32_t s_base[2][2];
    const uint32_t count = Q.ctr[cur];
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    const uint64_t lt = (1ull << lane) - 1ull;
    const uint32_t stride = gridDim.x * blockDim.x;
    const Queue A = Q.q[cur], N = Q.q[nxt];
    uint32_t buf = 0;
    for (uint32_t base = blockIdx.x * blockDim.x; base < count; base += stride, buf ^= 1u) {
        const uint32_t i = base + threadIdx.x;
        const bool valid = i < count;
        Bounce b;
        b.shadow = false;
        b.cont = false;
        uint32_t path = 0;
        V3 beta = v3(0.0f, 0.0f, 0.0f), L = v3(0.0f, 0.0f, 0.0f);
        float pdf = 0.0f;
        Pcg rng{0u, 0u};
        if (valid) {
            const float4 r0 = A.ray[2 * (uint64_t)i], r1 = A.ray[2 * (uint64_t)i + 1];
            const float4 bp = A.beta[i], lp = A.rad[i];
            const uint4 rs = A.rng[i];
            const uint2 h0 = Q.hits[3 * (uint64_t)i], h1 = Q.hits[3 * (uint64_t)i + 1], h2 = Q.hits[3 * (uint64_t)i + 2];
            path = __float_as_uint(lp.w);
            rng = Pcg{(uint64_t)rs.x | ((uint64_t)rs.y << 32), (uint64_t)rs.z | ((uint64_t)rs.w << 32)};
            beta = v3(bp.x, bp.y, bp.z);
            pdf = bp.w;
            L = v3(lp.x, lp.y, lp.z);
            shade(S, depth, v3(r0.x, r0.y, r0.z), v3(r1.x, r1.y, r1.z), h0.x, h2.y, __uint_as_float(h1.x),
                  __uint_as_float(h1.y), rng, beta, pdf, L, b);
        }
        // block-wide reservation of the continuation and shadow slots (double-buffered offsets,
        // so one barrier pair per iteration suffices)
        const uint64_t mc = __ballot(b.cont), ms = __ballot(b.shadow);
        if (lane == 0) {
            s_off[buf][0][wave] = (uint32_t)__popcll(mc);
            s_off[buf][1][wave] = (uint32_t)__popcll(ms);
        }
        __syncthreads();
        if (threadIdx.x < 2) {
            uint32_t run = 0;
            for (int w = 0; w < kShadeWaves; w++) {
                const uint32_t c = s_off[buf][threadIdx.x][w];
                s_off[buf][threadIdx.x][w] = run;
                run += c;
            }
            s_base[buf][threadIdx.x] = run ? atomicAdd(Q.ctr + (threadIdx.x == 0 ? nxt : 3u), run) : 0u;
        }
        __syncthreads();
        // continuation first: the shadow entry names its slot
        const uint32_t k = s_base[buf][0] + s_off[buf][0][wave] + (uint32_t)__popcll(mc & lt);
        if (b.cont) {
            N.ray[2 * (uint64_t)k] = make_float4(b.no.x, b.no.y, b.no.z, 0.0f);
            N.ray[2 * (uint64_t)k + 1] = make_float4(b.nd.x, b.nd.y, b.nd.z, FLT_MAX);
            N.beta[k] = make_float4(beta.x, beta.y, beta.z, pdf);
            N.rad[k] = make_float4(L.x, L.y, L.z, __uint_as_float(path));
            N.rng[k] = make_uint4((uint32_t)rng.state, (uint32_t)(rng.state >> 32), (uint32_t)rng.inc,
                                  (uint32_t)(rng.inc >> 32));
            if (keys) N.key[k] = (uint16_t)bucket_of(b.no, b.nd, B);
        } else if (valid) {
            Q.Lf[path] = make_float4(L.x, L.y, L.z, 0.0f);
        }
        if (b.shadow) {
            const uint64_t s = s_base[buf][1] + s_off[buf][1][wave] + (uint32_t)__popcll(ms & lt);
            Q.sr[2 * s] = make_float4(b.so.x, b.so.y, b.so.z, 0.0f);
            Q.sr[2 * s + 1] = make_float4(b.wi.x, b.wi.y, b.wi.z, b.d_light);
            Q.sc[s] = make_float4(b.contrib.x, b.contrib.y, b.contrib.z, __uint_as_float(b.cont ? k : (path | kEnded)));
        }
    }
}

// Unoccluded light samples add their contribution to the continuation entry or, for a path
// that ended this bounce, to its final radiance (one shadow ray per path per bounce: no races).
__global__ void k_pt_resolve(const Queues Q, uint32_t nxt) {
    const uint32_t count = Q.ctr[3];
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
        if (Q.occ[i]) continue;
        const float4 c = Q.sc[i];
        const uint32_t t = __float_as_uint(c.w);
        float4 *dst = (t & kEnded) ? Q.Lf + (t & ~kEnded) : Q.q[nxt].rad + t;
        float4 l = *dst;
        l.x = l.x + c.x;
        l.y = l.y + c.y;
        l.z = l.z + c.z;
        *dst = l;
    }
}

// Queue lengths: clear the next queue and the shadow queue; tally the rays traced.
__global__ void k_pt_counters(const Queues Q, uint32_t cur, uint32_t nxt, unsigned long long *counts, int phase) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (phase == 0) {
        counts[0] += Q.ctr[cur];
        Q.ctr[nxt] = 0u;
        Q.ctr[3] = 0u;
    } else {
        counts[1] += Q.ctr[3];
    }
}

// Adds each job's samples of this wave to its accumulator in sample order (the order of the
// megakernel's per-pixel sum); the job's last wave writes the mean.
__global__ __launch_bounds__(kGridBlock) void k_pt_accumulate(const PtWaveTable *T, const float4 *Lf) {
    __shared__ uint32_t base[kMaxWaveJobs];
    const uint32_t njobs = T->njobs, pixels = T->pixels;
    for (uint32_t k = threadIdx.x; k < njobs; k += blockDim.x) base[k] = T->jobs[k].pix_base;
    __syncthreads();
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < pixels; t += gridDim.x * blockDim.x) {
        const PtJob &J = T->jobs[search_le(base, njobs, t)];
        const uint32_t npix = J.width * J.height;
        const uint32_t pixel = t - J.pix_base;
        float4 a = J.first ? make_float4(0.0f, 0.0f, 0.0f, 0.0f) : J.acc[pixel];
        for (uint32_t s = 0; s < J.nsamples; s++) {
            const float4 l = Lf[J.path_base + (uint64_t)s * npix + pixel];
            if (has_nan(l.x, l.y, l.z)) continue; // a NaN path adds nothing
            a.x = a.x + l.x;
            a.y = a.y + l.y;
            a.z = a.z + l.z;
        }
        if (J.finish_spp) {
            const float fs = (float)J.finish_spp;
            J.out[pixel] = make_float4(a.x / fs, a.y / fs, a.z / fs, 1.0f);
        } else {
            J.acc[pixel] = a;
        }
    }
}

__global__ void k_fill_u64(unsigned long long *p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0ull;
}

__global__ void k_fill_u32(uint32_t *p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0u;
}

// ---- bucket sort of the continuation queue (queue length read on the GPU) ----
// Block b takes queue entries [b * seg, (b + 1) * seg), seg = ceil(count / gridDim.x).
__global__ __launch_bounds__(256) void k_pt_hist(const Queues Q, uint32_t nxt) {
    __shared__ uint32_t h[kBins];
    for (uint32_t k = threadIdx.x; k < kBins; k += blockDim.x) h[k] = 0u;
    __syncthreads();
    const uint32_t count = Q.ctr[nxt];
    const uint32_t seg = (count + gridDim.x - 1u) / gridDim.x;
    const uint32_t begin = min(blockIdx.x * seg, count), end = min(begin + seg, count);
    for (uint32_t i = begin + threadIdx.x; i < end; i += blockDim.x) atomicAdd(&h[Q.q[nxt].key[i]], 1u);
    __syncthreads();
    for (uint32_t k = threadIdx.x; k < kBins; k += blockDim.x) Q.hist[(uint64_t)k * gridDim.x + blockIdx.x] = h[k];
}

// One block per bin: exclusive scan of the bin's per-block counts (kSortBlocks of them).
__global__ __launch_bounds__(256) void k_pt_scan_bins(const Queues Q) {
    __shared__ uint32_t part[256];
    constexpr uint32_t per = kSortBlocks / 256u;
    uint32_t *row = Q.hist + (uint64_t)blockIdx.x * kSortBlocks;
    uint32_t v[per], s = 0;
    for (uint32_t j = 0; j < per; j++) {
        v[j] = row[threadIdx.x * per + j];
        s += v[j];
    }
    part[threadIdx.x] = s;
    __syncthreads();
    for (uint32_t off = 1; off < 256u; off <<= 1) {
        const uint32_t add = threadIdx.x >= off ? part[threadIdx.x - off] : 0u;
        __syncthreads();
        part[threadIdx.x] += add;
        __syncthreads();
    }
    uint32_t run = part[threadIdx.x] - s; // exclusive prefix of this thread's run
    for (uint32_t j = 0; j < per; j++) {
        row[threadIdx.x * per + j] = run;
        run += v[j];
    }
    if (threadIdx.x == 255u) Q.bin_total[blockIdx.x] = part[255];
}

// One block: exclusive scan of the bin totals.
__global__ __launch_bounds__(1024) void k_pt_scan_total(const Queues Q) {
    __shared__ uint32_t part[1024];
    constexpr uint32_t per = kBins / 1024u;
    uint32_t v[per], s = 0;
    for (uint32_t j = 0; j < per; j++) {
        v[j] = Q.bin_total[threadIdx.x * per + j];
        s += v[j];
    }
    part[threadIdx.x] = s;
    __syncthreads();
    for (uint32_t off = 1; off < 1024u; off <<= 1) {
        const uint32_t add = threadIdx.x >= off ? part[threadIdx.x - off] : 0u;
        __syncthreads();
        part[threadIdx.x] += add;
        __syncthreads();
    }
    uint32_t run = part[threadIdx.x] - s;
    for (uint32_t j = 0; j < per; j++) {
        Q.bin_base[threadIdx.x * per + j] = run;
        run += v[j];
    }
}

// Moves each entry of queue `from` to its bucket position in queue `to`.
__global__ __launch_bounds__(256) void k_pt_scatter(const Queues Q, uint32_t from, uint32_t to) {
    __shared__ uint32_t pos[kBins];
    for (uint32_t k = threadIdx.x; k < kBins; k += blockDim.x)
        pos[k] = Q.bin_base[k] + Q.hist[(uint64_t)k * gridDim.x + blockIdx.x];
    __syncthreads();
    const uint32_t count = Q.ctr[from];
    if (blockIdx.x == 0 && threadIdx.x == 0) Q.ctr[to] = count;
    const uint32_t seg = (count + gridDim.x - 1u) / gridDim.x;
    const uint32_t begin = min(blockIdx.x * seg, count), end = min(begin + seg, count);
    const Queue F = Q.q[from], T = Q.q[to];
    for (uint32_t i = begin + threadIdx.x; i < end; i += blockDim.x) {
        const uint32_t at = atomicAdd(&pos[F.key[i]], 1u);
        T.ray[2 * (uint64_t)at] = F.ray[2 * (uint64_t)i];
        T.ray[2 * (uint64_t)at + 1] = F.ray[2 * (uint64_t)i + 1];
        T.beta[at] = F.beta[i];
        T.rad[at] = F.rad[i];
        T.rng[at] = F.rng[i];
    }
}

// ---------------------------------------------------------------------------------------------
// host side
constexpr int kStaging = 8; // pinned wave-table staging slots (ring)

struct Workspace {
    size_t cap = 0;         // paths per wave
    void *mem = nullptr;
    Queues Q{};
    uint2 *spill = nullptr;
    size_t spill_bytes = 0;
    PtInstance *inst = nullptr;
    size_t inst_cap = 0;
    std::vector<PtInstance> inst_host; // what the device table holds
    PtWaveTable *table = nullptr;       // the device wave table the launches read
    PtWaveTable *staging = nullptr;     // pinned host ring of kStaging tables
    hipEvent_t staged[kStaging] = {};   // copy out of each staging slot done
    int next_stage = 0;
    unsigned long long *counts = nullptr; // [closest rays, shadow rays] when the caller passes none
    uint64_t generation = 0;
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    std::vector<uint8_t> key;
};

size_t round_up(size_t x) { return (x + 255u) & ~(size_t)255u; }

hipError_t workspace_reserve(Workspace &W, size_t cap) {
    if (cap <= W.cap && W.mem) return hipSuccess;
    if (W.mem) {
        hipError_t e = hipFree(W.mem);
        if (e != hipSuccess) return e;
        W.mem = nullptr;
    }
    // per queue: ray 32 + beta 16 + rad 16 + rng 16 + key 2 bytes per entry
    const size_t sizes[] = {32 * cap, 16 * cap, 16 * cap, 16 * cap, 2 * cap,   // q[0]
                            32 * cap, 16 * cap, 16 * cap, 16 * cap, 2 * cap,   // q[1]
                            32 * cap, 16 * cap, 16 * cap, 16 * cap, 2 * cap,   // q[2]
                            24 * cap, 32 * cap, 16 * cap, cap, 16 * cap,       // hits, sr, sc, occ, Lf
                            4ull * kBins * kSortBlocks, 4ull * kBins, 4ull * kBins, 4 * kCtrWords};
    size_t total = 0;
    for (size_t s : sizes) total += round_up(s);
    hipError_t e = hipMalloc(&W.mem, total);
    if (e != hipSuccess) {
        W.mem = nullptr;
        W.cap = 0;
        return e;
    }
    char *p = static_cast<char *>(W.mem);
    size_t k = 0;
    auto take = [&](auto *&dst) {
        dst = reinterpret_cast<std::remove_reference_t<decltype(dst)>>(p);
        p += round_up(sizes[k++]);
    };
    for (int q = 0; q < 3; q++) {
        take(W.Q.q[q].ray); take(W.Q.q[q].beta); take(W.Q.q[q].rad); take(W.Q.q[q].rng); take(W.Q.q[q].key);
    }
    take(W.Q.hits); take(W.Q.sr); take(W.Q.sc); take(W.Q.occ); take(W.Q.Lf);
    take(W.Q.hist); take(W.Q.bin_total); take(W.Q.bin_base); take(W.Q.ctr);
    W.cap = cap;
    W.generation++;
    return hipSuccess;
}

hipError_t grow(void *&ptr, size_t &have, size_t need, uint64_t &generation) {
    if (need <= have && ptr) return hipSuccess;
    if (ptr) {
        hipError_t e = hipFree(ptr);
        if (e != hipSuccess) return e;
        ptr = nullptr;
    }
    have = 0;
    hipError_t e = hipMalloc(&ptr, need);
    if (e != hipSuccess) {
        ptr = nullptr;
        return e;
    }
    have = need;
    generation++;
    return hipSuccess;
}

void drop_graph(Workspace &W) {
    if (W.exec) hipGraphExecDestroy(W.exec);
    if (W.graph) hipGraphDestroy(W.graph);
    W.exec = nullptr;
    W.graph = nullptr;
    W.key.clear();
}

constexpr int kMegaBlock = 128, kMegaLds = 16;

template<bool TWO>
int mega_blocks_per_cu() {
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, (const void *)k_pt_mega<TWO, kMegaBlock, kMegaLds>, kMegaBlock, 0) !=
        hipSuccess)
        return 1;
    return nb > 0 ? nb : 1;
}

} // namespace pt

using namespace pt;

// One workspace (queues, spill, graph) per stream: renders on different streams never share
// device memory; renders on one stream are ordered by the stream.
struct PtState {
    std::map<hipStream_t, Workspace *> ws;
};

static void free_workspace(Workspace *W) {
    drop_graph(*W);
    for (hipEvent_t &e : W->staged)
        if (e) hipEventDestroy(e);
    if (W->mem) hipFree(W->mem);
    if (W->spill) hipFree(W->spill);
    if (W->inst) hipFree(W->inst);
    if (W->counts) hipFree(W->counts);
    if (W->table) hipFree(W->table);
    if (W->staging) hipHostFree(W->staging);
    delete W;
}

void pt_release(Device *dev) {
    PtState *st = static_cast<PtState *>(dev->pt);
    if (!st) return;
    for (auto &kv : st->ws) free_workspace(kv.second);
    delete st;
    dev->pt = nullptr;
}

void pt_release_stream(Device *dev, hipStream_t s) {
    std::lock_guard<std::mutex> lk(dev->mu);
    PtState *st = static_cast<PtState *>(dev->pt);
    if (!st) return;
    auto it = st->ws.find(s);
    if (it == st->ws.end()) return;
    free_workspace(it->second);
    st->ws.erase(it);
}

#define PT_TRY(x)                                                 \
    do {                                                          \
        hipError_t e_ = (x);                                      \
        if (e_ != hipSuccess) {                                   \
            err = std::string(#x) + ": " + hipGetErrorString(e_); \
            return -1;                                            \
        }                                                         \
    } while (0)

// The stream's workspace with the instance table uploaded (re-uploaded only when it changes,
// after the stream has drained the previous copy). Caller holds dev->mu.
static int prepare(Device *dev, hipStream_t s, const std::vector<PtInstance> &inst, Workspace *&out, std::string &err) {
    if (!dev->pt) dev->pt = new PtState();
    PtState &state = *static_cast<PtState *>(dev->pt);
    Workspace *&slot = state.ws[s];
    if (!slot) slot = new Workspace();
    Workspace &W = *slot;
    out = slot;
    void *p = W.inst;
    size_t have = W.inst_cap;
    uint64_t gen = W.generation;
    PT_TRY(grow(p, have, std::max<size_t>(1, inst.size()) * sizeof(PtInstance), gen));
    if (p != W.inst) W.inst_host.clear();
    W.inst = static_cast<PtInstance *>(p);
    W.inst_cap = have;
    W.generation = gen;
    if (W.inst_host.size() != inst.size() ||
        std::memcmp(W.inst_host.data(), inst.data(), inst.size() * sizeof(PtInstance)) != 0) {
        PT_TRY(hipStreamSynchronize(s));
        W.inst_host = inst;
        PT_TRY(hipMemcpy(W.inst, W.inst_host.data(), inst.size() * sizeof(PtInstance), hipMemcpyHostToDevice));
    }
    if (!W.counts) PT_TRY(hipMalloc(&W.counts, 2 * sizeof(unsigned long long)));
    if (!W.table) {
        PT_TRY(hipMalloc(&W.table, sizeof(PtWaveTable)));
        PT_TRY(hipHostMalloc(reinterpret_cast<void **>(&W.staging), kStaging * sizeof(PtWaveTable), hipHostMallocDefault));
        for (hipEvent_t &e : W.staged) PT_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
    return 0;
}

// Enqueues one wave: the wave table is staged through the pinned ring and copied to the device
// table in stream order, then the wave's launches run (from the cached hipGraph when `graph`).
static int run_wave(Device *dev, hipStream_t s, Workspace &W, const PtBinding &Bd, const PtScene &S,
                    const PtJob *jobs, uint32_t njobs, size_t cap, const PtOptions &opt, unsigned long long *counts,
                    std::string &err) {
    if (njobs == 0) return 0;
    if (njobs > kMaxWaveJobs) {
        err = "too many jobs in one wave";
        return -2;
    }
    if (cap >= (size_t)kEnded) {
        err = "a wavefront wave must hold fewer than 2^31 paths";
        return -2;
    }
    // table: path and pixel bases
    {
        const int slot = W.next_stage;
        W.next_stage = (W.next_stage + 1) % kStaging;
        PT_TRY(hipEventSynchronize(W.staged[slot])); // that slot's previous copy has been consumed
        PtWaveTable &T = W.staging[slot];
        uint64_t paths = 0, pixels = 0;
        for (uint32_t j = 0; j < njobs; j++) {
            T.jobs[j] = jobs[j];
            T.jobs[j].path_base = (uint32_t)paths;
            T.jobs[j].pix_base = (uint32_t)pixels;
            const uint64_t npix = (uint64_t)jobs[j].width * jobs[j].height;
            paths += npix * jobs[j].nsamples;
            pixels += npix;
        }
        if (paths > cap) {
            err = "wave exceeds its path capacity";
            return -2;
        }
        T.njobs = njobs;
        T.paths = (uint32_t)paths;
        T.pixels = (uint32_t)pixels;
        T._pad = 0;
        const size_t bytes = offsetof(PtWaveTable, jobs) + njobs * sizeof(PtJob);
        PT_TRY(hipMemcpyAsync(W.table, &T, bytes, hipMemcpyHostToDevice, s));
        PT_TRY(hipEventRecord(W.staged[slot], s));
    }
    PT_TRY(workspace_reserve(W, cap));
    const TraceConfig cc = dev->cfg[0], ca = dev->cfg[1];
    const bool two = Bd.two;
    const int bpc_c = cc.blocks_per_cu > 0 ? cc.blocks_per_cu : trace_blocks_per_cu(false, two, false, cc.block, cc.slds);
    const int bpc_a = ca.blocks_per_cu > 0 ? ca.blocks_per_cu : trace_blocks_per_cu(true, two, false, ca.block, ca.slds);
    const int grid_c = bpc_c * dev->cu_count, grid_a = bpc_a * dev->cu_count;
    const uint32_t sd_c = Bd.stack_need > (uint32_t)cc.slds ? Bd.stack_need - cc.slds : 0u;
    const uint32_t sd_a = Bd.stack_need > (uint32_t)ca.slds ? Bd.stack_need - ca.slds : 0u;
    {
        const size_t need = std::max<size_t>({16, (size_t)grid_c * cc.block * sd_c * sizeof(uint2),
                                              (size_t)grid_a * ca.block * sd_a * sizeof(uint2)});
        void *sp = W.spill;
        size_t have = W.spill_bytes;
        PT_TRY(grow(sp, have, need, W.generation));
        W.spill = static_cast<uint2 *>(sp);
        W.spill_bytes = have;
    }
    SortBox SB;
    for (int a = 0; a < 3; a++) {
        const float ext = Bd.bounds[3 + a] - Bd.bounds[a];
        SB.lo[a] = Bd.bounds[a];
        SB.scale[a] = ext > 0.0f ? 8.0f / ext : 0.0f;
    }
    // graph cache key: everything the captured launches depend on (the wave table is data)
    std::vector<uint8_t> key;
    auto put = [&](const void *p, size_t n) {
        const uint8_t *b = static_cast<const uint8_t *>(p);
        key.insert(key.end(), b, b + n);
    };
    put(&Bd.G, sizeof(Bd.G)); put(&S, sizeof(S)); put(&opt, sizeof(opt)); put(&counts, sizeof(counts));
    put(&W.generation, sizeof(W.generation)); put(&cc, sizeof(cc)); put(&ca, sizeof(ca)); put(&SB, sizeof(SB));
    put(&cap, sizeof(cap)); put(&two, sizeof(two)); put(&Bd.stack_need, sizeof(Bd.stack_need));
    put(&dev->error, sizeof(dev->error));
    if (opt.graph && W.exec && key == W.key) {
        PT_TRY(hipGraphLaunch(W.exec, s));
        return 0;
    }
    if (opt.graph) {
        drop_graph(W);
        PT_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    }
    const int grid_shade_blocks = std::max(1, dev->cu_count * 4);
    const int grid_small = std::max(1, dev->cu_count * 8);
    auto enqueue = [&]() -> hipError_t {
        const Queues &Q = W.Q;
        hipLaunchKernelGGL(k_pt_camera, dim3(grid_small), dim3(kGridBlock), 0, s, S, W.table, Q);
        uint32_t cur = 0, nxt = 1, spare = 2;
        for (uint32_t depth = 0; depth < S.max_depth; depth++) {
            TraceParams P;
            std::memset(&P, 0, sizeof(P));
            P.batches.n = 1;
            P.batches.b[0] = RayBatch{Q.q[cur].ray, Q.hits, 0u, 0xFFu};
            P.count = (uint32_t)cap;
            P.count_ptr = Q.ctr + cur;
            P.ray_counter = Q.ctr + kCtrClosest;
            P.order = nullptr;
            P.geo = Bd.G;
            P.spill = W.spill;
            P.spill_depth = sd_c;
            P.steps = cc.steps;
            P.fetch_chunk = opt.fetch_chunk;
            P.fetch_segments = opt.fetch_segments;
            P.tri_quarters = cc.tri_quarters;
            P.refill_below = cc.refill_below;
            P.stats = nullptr;
            P.error = dev->error;
            P.hit_record = 1u;
            hipError_t e = trace_launch(P, s, false, two, false, grid_c, cc.block, cc.slds);
            if (e != hipSuccess) return e;
            hipLaunchKernelGGL(k_pt_counters, dim3(1), dim3(64), 0, s, Q, cur, nxt, counts, 0);
            const bool sort = opt.sort && depth + 1u < S.max_depth;
            hipLaunchKernelGGL(k_pt_shade, dim3(grid_shade_blocks), dim3(kShadeBlock), 0, s, S, depth, Q, cur, nxt,
                               sort ? 1 : 0, SB);
            TraceParams A = P;
            A.batches.b[0] = RayBatch{Q.sr, Q.occ, 0u, 0xFFu};
            A.count_ptr = Q.ctr + 3;
            A.ray_counter = Q.ctr + kCtrAny;
            A.spill_depth = sd_a;
            A.steps = ca.steps;
            A.refill_below = ca.refill_below;
            A.hit_record = 0u;
            e = trace_launch(A, s, true, two, false, grid_a, ca.block, ca.slds);
            if (e != hipSuccess) return e;
            hipLaunchKernelGGL(k_pt_counters, dim3(1), dim3(64), 0, s, Q, cur, nxt, counts, 1);
            hipLaunchKernelGGL(k_pt_resolve, dim3(grid_small), dim3(256), 0, s, Q, nxt);
            if (sort) {
                hipLaunchKernelGGL(k_pt_hist, dim3(kSortBlocks), dim3(256), 0, s, Q, nxt);
                hipLaunchKernelGGL(k_pt_scan_bins, dim3(kBins), dim3(256), 0, s, Q);
                hipLaunchKernelGGL(k_pt_scan_total, dim3(1), dim3(1024), 0, s, Q);
                hipLaunchKernelGGL(k_pt_scatter, dim3(kSortBlocks), dim3(256), 0, s, Q, nxt, spare);
                const uint32_t old = cur;
                cur = spare;
                spare = nxt;
                nxt = old;
            } else {
                std::swap(cur, nxt);
            }
        }
        hipLaunchKernelGGL(k_pt_accumulate, dim3(grid_small), dim3(kGridBlock), 0, s, W.table, Q.Lf);
        return hipGetLastError();
    };
    hipError_t e = enqueue();
    if (opt.graph) {
        hipGraph_t g = nullptr;
        hipError_t e2 = hipStreamEndCapture(s, &g);
        if (e != hipSuccess || e2 != hipSuccess) {
            if (g) hipGraphDestroy(g);
            err = std::string("wavefront capture: ") + hipGetErrorString(e != hipSuccess ? e : e2);
            return -1;
        }
        W.graph = g;
        PT_TRY(hipGraphInstantiate(&W.exec, W.graph, nullptr, nullptr, 0));
        W.key = key;
        PT_TRY(hipGraphLaunch(W.exec, s));
        return 0;
    }
    if (e != hipSuccess) {
        err = std::string("wavefront launch: ") + hipGetErrorString(e);
        return -1;
    }
    return 0;
}

int pt_wave(Device *dev, hipStream_t s, const PtBinding &B, const PtJob *jobs, uint32_t njobs, size_t capacity,
            const PtOptions &opt, unsigned long long *counts, std::string &err) {
    std::lock_guard<std::mutex> lk(dev->mu);
    Workspace *W = nullptr;
    int r = prepare(dev, s, B.inst, W, err);
    if (r != 0) return r;
    PtScene S = B.S;
    S.inst = W->inst;
    return run_wave(dev, s, *W, B, S, jobs, njobs, capacity, opt, counts ? counts : W->counts, err);
}

int pt_render(Device *dev, hipStream_t s, const Geometry &G, bool two, uint32_t stack_need, const float bounds[6],
              const PtScene &scene_in, const std::vector<PtInstance> &inst, const PtOptions &opt, float4 *image,
              unsigned long long *counts, std::string &err) {
    std::lock_guard<std::mutex> lk(dev->mu);
    Workspace *Wp = nullptr;
    int r = prepare(dev, s, inst, Wp, err);
    if (r != 0) return r;
    Workspace &W = *Wp;
    if (!counts) counts = W.counts;
    PtScene S = scene_in;
    S.inst = W.inst;
    const uint32_t npix = S.width * S.height;
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, s, counts, 2u);
    if (!opt.wavefront) {
        const int bpc = two ? mega_blocks_per_cu<true>() : mega_blocks_per_cu<false>();
        const int grid = bpc * dev->cu_count;
        const uint32_t spill_depth = stack_need > (uint32_t)kMegaLds ? stack_need - kMegaLds : 0u;
        void *sp = W.spill;
        size_t have = W.spill_bytes;
        PT_TRY(grow(sp, have, std::max<size_t>(16, (size_t)grid * kMegaBlock * spill_depth * sizeof(uint2)), W.generation));
        W.spill = static_cast<uint2 *>(sp);
        W.spill_bytes = have;
        PT_TRY(workspace_reserve(W, std::max<size_t>(W.cap, 1)));
        hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, s, W.Q.ctr, 16u);
        if (two)
            hipLaunchKernelGGL((k_pt_mega<true, kMegaBlock, kMegaLds>), dim3(grid), dim3(kMegaBlock), 0, s, S, G, image,
                               W.Q.ctr + 4, W.spill, spill_depth, dev->error, counts);
        else
            hipLaunchKernelGGL((k_pt_mega<false, kMegaBlock, kMegaLds>), dim3(grid), dim3(kMegaBlock), 0, s, S, G, image,
                               W.Q.ctr + 4, W.spill, spill_depth, dev->error, counts);
        PT_TRY(hipGetLastError());
        return 0;
    }
    // wavefront: one job per wave, chunk_spp samples per wave, accumulated in place in the image
    uint32_t chunk = opt.chunk_spp;
    if (chunk == 0) chunk = (uint32_t)std::max<uint64_t>(1, (32ull << 20) / std::max<uint32_t>(1, npix));
    chunk = std::min(chunk, S.spp);
    const size_t cap = (size_t)npix * chunk;
    PtBinding Bd;
    Bd.G = G;
    Bd.two = two;
    Bd.stack_need = stack_need;
    std::memcpy(Bd.bounds, bounds, sizeof(Bd.bounds));
    for (uint32_t base = 0; base < S.spp; base += chunk) {
        PtJob J;
        std::memset(&J, 0, sizeof(J));
        J.acc = image;
        J.out = image;
        J.width = S.width;
        J.height = S.height;
        J.sample0 = S.sample_offset + base;
        J.nsamples = std::min(chunk, S.spp - base);
        J.first = base == 0 ? 1u : 0u;
        J.finish_spp = base + J.nsamples == S.spp ? S.spp : 0u;
        J.seed = S.seed;
        J.sequence = S.sequence;
        // the per-render fields do not reach the wave launches (the graph key stays the same)
        PtScene Sw = S;
        Sw.width = Sw.height = Sw.spp = Sw.sample_offset = 0;
        Sw.seed = Sw.sequence = 0;
        r = run_wave(dev, s, W, Bd, Sw, &J, 1, cap, opt, counts, err);
        if (r != 0) return r;
    }
    return 0;
}

#undef PT_TRY

} // namespace lchip
