/*
 * lc_oracle.h -- CPU restatement of the reference CPU backend's ray-tracing semantics.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker for the HIP backend: only the
 * test suite, the smoke check and the benchmark's CPU-baseline leg may load it. The product
 * (luisacompute_amd/) never links, calls or falls back to it.
 *
 * What it restates (reference paths relative to the LuisaCompute checkout):
 *   - hit semantics of AccelImpl::trace_closest / trace_any
 *       src/rust/luisa_compute_backend_impl/src/cpu/accel.rs:449-535
 *     miss -> inst = prim = 0xFFFFFFFF, bary = (0,0), t = ray.tmax; any-hit -> bool.
 *   - mesh input interpretation (FLOAT3 at a byte stride, UINT3 indices)  cpu/accel.rs:215-237
 *   - instance semantics (instance id = index, default mask 0xff, row-major 3x4 affine,
 *     world->object ray transform without renormalisation)               cpu/accel.rs:324-447
 *   - FTZ/DAZ on the worker threads                                      cpu/mod.rs:503-516
 * The reference delegates the arithmetic to Embree 4 (embree_sys 0.1.11, un-vendored), which
 * cannot be built offline. This file restates the published algorithms instead:
 *   - watertight ray/triangle test: Woop, Benthin, Wald, "Watertight Ray/Triangle
 *     Intersection", JCGT 2(1), 2013 (with the double-precision edge fallback);
 *   - binned-SAH BVH (Wald 2007 "On fast construction of SAH-based BVHs");
 *   - robust slab test: Ize, "Robust BVH Ray Traversal", JCGT 2(2), 2013;
 *   - PCG32: O'Neill 2014, pcg32_srandom_r / pcg32_random_r / pcg32_advance_r.
 * plus an independent double-precision Moller-Trumbore brute force used to pin it.
 *
 * Pinning status per function is listed in oracle/README.md.
 */
#ifndef LC_ORACLE_H
#define LC_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- PCG32 (O'Neill 2014, pcg-c-basic) ---- */
typedef struct orc_pcg32 { uint64_t state; uint64_t inc; } orc_pcg32;
void     orc_pcg32_seed(orc_pcg32 *rng, uint64_t initstate, uint64_t initseq);
uint32_t orc_pcg32_next(orc_pcg32 *rng);
void     orc_pcg32_advance(orc_pcg32 *rng, uint64_t delta);
float    orc_pcg32_float(orc_pcg32 *rng); /* (x >> 8) * 2^-24 */

/* ---- synthetic workload generators (restated independently of the product's) ---- */
/* icosphere: level-L subdivided icosahedron, radius 1. verts: 4 floats per vertex (16-B stride). */
void orc_icosphere_counts(int level, uint32_t *nverts, uint32_t *ntris);
void orc_icosphere(int level, float *verts4, uint32_t *tris);
/* rays from a sphere of radius `orbit` toward uniform points in the unit ball (8 floats/ray). */
void orc_rays_orbit_to_ball(uint64_t initstate, uint64_t initseq, uint64_t nrays, float orbit,
                            float *rays);
/* triangle soup: centroid U[0,extent]^3, vertices centroid + N(0, sigma^2) per axis.
   Stream offsets: triangle i consumes draws [13 i, 13 i + 13). */
void orc_soup(uint64_t initstate, uint64_t initseq, uint64_t ntris, float extent, float sigma,
              float *verts4, uint32_t *tris);
/* rays with origin U[lo,hi]^3 and direction uniform on S^2; ray j consumes draws
   [skip + 5 j, skip + 5 j + 5) of the stream. */
void orc_rays_box_sphere(uint64_t initstate, uint64_t initseq, uint64_t skip, uint64_t nrays,
                         float lo, float hi, float *rays);

/* instances: per instance 8 draws -- mesh floor(u * mesh_count), Shoemake uniform rotation,
   translation U[-extent, extent]^3, scale U[slo, shi]; affines row-major 3x4 (12 floats). */
void orc_instances(uint64_t initstate, uint64_t initseq, uint32_t count, uint32_t mesh_count, float extent, float slo,
                   float shi, float *affines, uint32_t *mesh_ids);
/* log-uniform batch sizes in [lo, hi], one draw each after `skip` draws */
void orc_batch_sizes(uint64_t initstate, uint64_t initseq, uint64_t skip, uint32_t count, uint32_t lo, uint32_t hi,
                     uint32_t *sizes);

/* ---- geometry / scene ---- */
typedef struct orc_hit { uint32_t inst; uint32_t prim; float u; float v; float t; uint32_t pad; } orc_hit;
typedef struct orc_hit2 { uint32_t inst; uint32_t prim; float u; float v; float t; float t2; } orc_hit2;

void *orc_mesh_build(const float *verts, uint64_t vstride_bytes, uint64_t nverts,
                     const uint32_t *tris, uint64_t ntris, int nthreads);
void  orc_mesh_free(void *mesh);
uint64_t orc_mesh_node_count(const void *mesh);

void *orc_scene_create(void);
/* affine: row-major 3x4 (object -> world); mask: instance visibility (default 0xff). */
void  orc_scene_add(void *scene, const void *mesh, const float affine[12], uint32_t mask);
void  orc_scene_commit(void *scene);
void  orc_scene_free(void *scene);

/* rays: 8 floats each {ox,oy,oz,tmin,dx,dy,dz,tmax}. */
void orc_trace_closest(const void *scene, const float *rays, uint64_t nrays, uint32_t mask,
                       orc_hit *hits, int nthreads);
/* closest hit plus the second-nearest distinct-primitive hit distance t2 (inf if none). */
void orc_trace_closest2(const void *scene, const float *rays, uint64_t nrays, uint32_t mask,
                        orc_hit2 *hits, int nthreads);
void orc_trace_any(const void *scene, const float *rays, uint64_t nrays, uint32_t mask,
                   uint8_t *hit, int nthreads);

/* ---- procedural primitives and ray queries (cpu/accel.rs:582-800) ----
   A procedural primitive is a list of AABBs (6 floats each: min.xyz, max.xyz). Plain traces do
   not see procedural instances; ray queries report their candidates to the procedural callback.
   orc_scene_add(...) == orc_scene_add_ex(scene, mesh, NULL, affine, mask, 1). */
void *orc_procedural_create(const float *aabbs, uint64_t n);
void  orc_procedural_free(void *procedural);
void  orc_scene_add_ex(void *scene, const void *mesh, const void *procedural, const float affine[12], uint32_t mask,
                       uint32_t opaque);
/* built-in callbacks: surface_mode 0 commit, 1 ignore, 2 cutout (|(w,u)|, |(u,v)|, |(v,w)| all
   < cutout with w = 1 - u - v); procedural_mode 0 ignore, 1 the sphere inscribed in the AABB. */
typedef struct orc_query_policy { uint32_t surface_mode; float cutout; uint32_t procedural_mode; uint32_t terminate_on_first; } orc_query_policy;
typedef struct orc_committed_hit { uint32_t inst; uint32_t prim; float u; float v; uint32_t hit_type; float t; } orc_committed_hit;
/* nearest committed candidate per ray (every candidate enumerated, no BVH); terminate_on_first
   is ignored: any committed candidate is a valid answer of traverse_any, checked one at a time
   with orc_query_candidate. Miss: inst = prim = UINT32_MAX, u = v = 0, hit_type 0, t = tmax. */
void orc_query(const void *scene, const float *rays, uint64_t nrays, uint32_t mask, const orc_query_policy *policy,
               orc_committed_hit *out, int nthreads);
/* 1 with (t, u, v) if candidate (inst, prim) of kind hit_type (1 surface, 2 procedural) is
   committed for `ray` (8 floats), else 0. */
int orc_query_candidate(const void *scene, const float *ray, uint32_t mask, const orc_query_policy *policy,
                        uint32_t inst, uint32_t prim, uint32_t hit_type, float *t, float *u, float *v);

/* Independent check: double-precision Moller-Trumbore brute force over one mesh (no BVH).
   out per ray: prim (UINT32_MAX on miss), t, u, v (doubles) and second-nearest t2. */
typedef struct orc_bf_hit { uint32_t prim; uint32_t pad; double t, u, v, t2; } orc_bf_hit;
void orc_brute_closest_f64(const float *verts, uint64_t vstride_bytes, const uint32_t *tris,
                           uint64_t ntris, const float *rays, uint64_t nrays, orc_bf_hit *out,
                           int nthreads);

/* ---- path tracer: the DSL program of src/tests/test_path_tracing.cpp:165-258 on a scene whose
   instance i is shaded with mats[i]. Fixed choices (the product's path tracer makes the same):
   pixel (x, y) looks along normalize(fx right - fy up + forward), fx = (2 (x + jx) - width) /
   min(width, height), fy likewise, with a fresh jitter (jx, jy) per sample; pixel i draws from
   PCG32 stream (seed, sequence + i), sample s starting 64 s draws into it; per bounce two draws
   for the light sample, two for the cosine-weighted direction, one for Russian roulette from
   bounce rr_depth on (q = max(luminance(beta), 0.05)); at most max_depth closest-hit rays per
   path; a shadow ray only when the light sample can contribute; sin/cos of the sampled angle
   from the degree 11/12 Taylor polynomials after an exact quadrant reduction; a NaN path
   counts as zero. image: 4 floats per pixel (mean radiance over the samples, 1); only pixels
   [pixel_begin, pixel_end) are written. counts[0] += closest-hit rays, counts[1] += shadow rays. */
typedef struct orc_pt_material { float albedo[3]; uint32_t emissive; } orc_pt_material;
typedef struct orc_pt_params {
    float light_position[3], light_u[3], light_v[3], light_emission[3];
    float camera_origin[3], camera_right[3], camera_up[3], camera_forward[3];
    uint32_t width, height, spp, sample_offset, max_depth, rr_depth;
    uint64_t seed, sequence;
} orc_pt_params;
void orc_pathtrace(const void *scene, const orc_pt_material *mats, const orc_pt_params *params,
                   uint64_t pixel_begin, uint64_t pixel_end, float *image, uint64_t counts[2], int nthreads);
/* every ray the path tracer traces for pixels [pixel_begin, pixel_end) (8 floats each, in path
   order): closest-hit rays and shadow rays; counts are the totals (arrays hold up to cap). */
void orc_pathtrace_rays(const void *scene, const orc_pt_material *mats, const orc_pt_params *params,
                        uint64_t pixel_begin, uint64_t pixel_end, float *closest, uint64_t closest_cap,
                        uint64_t *closest_count, float *shadow, uint64_t shadow_cap, uint64_t *shadow_count);
/* offset_ray_origin (src/dsl/rtx/ray.cpp:16-24), as the path tracer applies it */
void orc_offset_ray_origin(const float p[3], const float n[3], float out[3]);
/* the path tracer's sin(2 pi u), cos(2 pi u) (exposed for the accuracy test) */
void orc_sincos_2pi(float u, float *s, float *c);

/* single watertight test exposed for unit tests (returns 1 on hit inside [tmin,tmax]). */
int orc_watertight(const float o[3], const float d[3], const float v0[3], const float v1[3],
                   const float v2[3], float tmin, float tmax, float *t, float *u, float *v);

#ifdef __cplusplus
}
#endif
#endif
