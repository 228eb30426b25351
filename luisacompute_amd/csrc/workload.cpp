// workload.cpp -- synthetic scenes and rays for the benchmark workloads, generated on the host
// from PCG32 streams (O'Neill 2014: pcg32_srandom_r / pcg32_random_r, and Brown 1994 LCG
// jump-ahead for parallel generation). Stream layout: initstate 42, initseq = workload number.
//   icosphere  : icosahedron (12 vertices / 20 faces, outward winding) subdivided L times,
//                midpoints shared between neighbouring faces and projected to the unit sphere.
//   orbit rays : origin on a sphere of radius `orbit` (z = 1 - 2u, phi = 2 pi u), direction
//                toward a point uniform in the unit ball (cube rejection), tmin 0, tmax 1e30.
//   soup       : 13 draws per triangle: centroid U[0,extent]^3, then 5 Box-Muller pairs of which
//                9 normals offset the three vertices by sigma each.
//   box rays   : 5 draws per ray: origin U[lo,hi]^3, direction uniform on S^2.
// Floats from a draw x are (x >> 8) * 2^-24; geometry is computed in double, stored in float.
#include "lc_hip.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <map>
#include <thread>
#include <utility>
#include <vector>

namespace {

struct Pcg {
    uint64_t state = 0, inc = 1;
    static constexpr uint64_t mult = 6364136223846793005ull;
    Pcg(uint64_t initstate, uint64_t initseq) {
        inc = (initseq << 1u) | 1u;
        next();
        state += initstate;
        next();
    }
    uint32_t next() {
        uint64_t old = state;
        state = old * mult + inc;
        uint32_t x = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = (uint32_t)(old >> 59u);
        return (x >> rot) | (x << ((0u - rot) & 31u));
    }
    double uniform() { return (double)((float)(next() >> 8) * 0x1p-24f); }
    void skip(uint64_t delta) {
        uint64_t m = mult, c = inc, am = 1, ac = 0;
        for (; delta; delta >>= 1u) {
            if (delta & 1u) {
                am *= m;
                ac = ac * m + c;
            }
            c = (m + 1u) * c;
            m *= m;
        }
        state = am * state + ac;
    }
};

constexpr double kTwoPi = 6.283185307179586476925286766559;

template<typename F>
void chunked(uint64_t n, F &&fn) {
    const uint64_t chunk = 1u << 16;
    uint64_t nchunks = (n + chunk - 1) / chunk;
    unsigned nt = std::max(1u, std::min<unsigned>(std::thread::hardware_concurrency(), 16u));
    nt = (unsigned)std::min<uint64_t>(nt, std::max<uint64_t>(nchunks, 1));
    std::vector<std::thread> ts;
    for (unsigned t = 0; t < nt; t++)
        ts.emplace_back([&, t] {
            for (uint64_t c = t; c < nchunks; c += nt) fn(c * chunk, std::min(n, (c + 1) * chunk));
        });
    for (auto &t : ts) t.join();
}

} // namespace

extern "C" {

void lc_hip_workload_icosphere_counts(int level, uint32_t *vertex_count, uint32_t *triangle_count) {
    uint64_t v = 12, f = 20;
    for (int i = 0; i < level; i++) {
        v += f * 3 / 2; // one new vertex per edge; edges = 3F/2
        f *= 4;
    }
    if (vertex_count) *vertex_count = (uint32_t)v;
    if (triangle_count) *triangle_count = (uint32_t)f;
}

void lc_hip_workload_icosphere(int level, float *vertices4, uint32_t *triangles) {
    const double g = (1.0 + std::sqrt(5.0)) * 0.5;
    std::vector<std::array<double, 3>> vert{{-1, g, 0}, {1, g, 0}, {-1, -g, 0}, {1, -g, 0}, {0, -1, g}, {0, 1, g},
                                           {0, -1, -g}, {0, 1, -g}, {g, 0, -1}, {g, 0, 1}, {-g, 0, -1}, {-g, 0, 1}};
    auto unit = [](std::array<double, 3> &p) {
        double l = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
        for (double &c : p) c /= l;
    };
    for (auto &p : vert) unit(p);
    std::vector<std::array<uint32_t, 3>> face{{0, 11, 5}, {0, 5, 1}, {0, 1, 7}, {0, 7, 10}, {0, 10, 11},
                                              {1, 5, 9}, {5, 11, 4}, {11, 10, 2}, {10, 7, 6}, {7, 1, 8},
                                              {3, 9, 4}, {3, 4, 2}, {3, 2, 6}, {3, 6, 8}, {3, 8, 9},
                                              {4, 9, 5}, {2, 4, 11}, {6, 2, 10}, {8, 6, 7}, {9, 8, 1}};
    for (int l = 0; l < level; l++) {
        std::map<std::pair<uint32_t, uint32_t>, uint32_t> edge_mid;
        auto mid = [&](uint32_t a, uint32_t b) -> uint32_t {
            std::pair<uint32_t, uint32_t> key{std::min(a, b), std::max(a, b)};
            auto it = edge_mid.find(key);
            if (it != edge_mid.end()) return it->second;
            std::array<double, 3> m{0.5 * (vert[a][0] + vert[b][0]), 0.5 * (vert[a][1] + vert[b][1]),
                                    0.5 * (vert[a][2] + vert[b][2])};
            unit(m);
            vert.push_back(m);
            uint32_t id = (uint32_t)(vert.size() - 1);
            edge_mid[key] = id;
            return id;
        };
        std::vector<std::array<uint32_t, 3>> next;
        next.reserve(face.size() * 4);
        for (const auto &f : face) {
            uint32_t ab = mid(f[0], f[1]), bc = mid(f[1], f[2]), ca = mid(f[2], f[0]);
            next.push_back({f[0], ab, ca});
            next.push_back({f[1], bc, ab});
            next.push_back({f[2], ca, bc});
            next.push_back({ab, bc, ca});
        }
        face.swap(next);
    }
    for (size_t i = 0; i < vert.size(); i++) {
        for (int c = 0; c < 3; c++) vertices4[4 * i + c] = (float)vert[i][c];
        vertices4[4 * i + 3] = 0.0f;
    }
    for (size_t i = 0; i < face.size(); i++)
        for (int c = 0; c < 3; c++) triangles[3 * i + c] = face[i][c];
}

void lc_hip_workload_rays_orbit_to_ball(uint64_t initstate, uint64_t initseq, uint64_t ray_count, float orbit,
                                        lc_hip_ray *rays) {
    Pcg rng(initstate, initseq);
    for (uint64_t i = 0; i < ray_count; i++) {
        double z = 1.0 - 2.0 * rng.uniform();
        double phi = kTwoPi * rng.uniform();
        double rr = std::sqrt(std::max(0.0, 1.0 - z * z));
        double o[3] = {orbit * rr * std::cos(phi), orbit * rr * std::sin(phi), orbit * z};
        double b[3];
        do {
            for (double &c : b) c = 2.0 * rng.uniform() - 1.0;
        } while (b[0] * b[0] + b[1] * b[1] + b[2] * b[2] > 1.0);
        double d[3] = {b[0] - o[0], b[1] - o[1], b[2] - o[2]};
        double len = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        lc_hip_ray &r = rays[i];
        for (int c = 0; c < 3; c++) {
            r.origin[c] = (float)o[c];
            r.direction[c] = (float)(d[c] / len);
        }
        r.t_min = 0.0f;
        r.t_max = 1e30f;
    }
}

void lc_hip_workload_soup(uint64_t initstate, uint64_t initseq, uint64_t triangle_count, float extent, float sigma,
                          float *vertices4, uint32_t *triangles) {
    chunked(triangle_count, [&](uint64_t b, uint64_t e) {
        Pcg rng(initstate, initseq);
        rng.skip(13u * b);
        for (uint64_t i = b; i < e; i++) {
            double c[3];
            for (double &x : c) x = (double)extent * rng.uniform();
            double g[10];
            for (int k = 0; k < 5; k++) {
                double u1 = 1.0 - rng.uniform(), u2 = rng.uniform();
                double r = std::sqrt(-2.0 * std::log(u1));
                g[2 * k] = r * std::cos(kTwoPi * u2);
                g[2 * k + 1] = r * std::sin(kTwoPi * u2);
            }
            for (int v = 0; v < 3; v++) {
                float *p = vertices4 + 4 * (3 * i + v);
                for (int k = 0; k < 3; k++) p[k] = (float)(c[k] + (double)sigma * g[3 * v + k]);
                p[3] = 0.0f;
                triangles[3 * i + v] = (uint32_t)(3 * i + v);
            }
        }
    });
}

void lc_hip_workload_instances(uint64_t initstate, uint64_t initseq, uint32_t instance_count, uint32_t mesh_count,
                               float translate_extent, float scale_lo, float scale_hi, float *affines,
                               uint32_t *mesh_ids) {
    // 8 draws per instance: mesh, rotation (Shoemake 1992 uniform quaternion), translation, scale
    Pcg rng(initstate, initseq);
    for (uint32_t i = 0; i < instance_count; i++) {
        const double um = rng.uniform();
        mesh_ids[i] = std::min<uint32_t>(mesh_count - 1, (uint32_t)(um * mesh_count));
        const double u1 = rng.uniform(), u2 = rng.uniform(), u3 = rng.uniform();
        const double x = std::sqrt(1.0 - u1) * std::sin(kTwoPi * u2), y = std::sqrt(1.0 - u1) * std::cos(kTwoPi * u2);
        const double z = std::sqrt(u1) * std::sin(kTwoPi * u3), w = std::sqrt(u1) * std::cos(kTwoPi * u3);
        double t[3];
        for (double &c : t) c = (double)translate_extent * (2.0 * rng.uniform() - 1.0);
        const double s = (double)scale_lo + ((double)scale_hi - (double)scale_lo) * rng.uniform();
        const double R[3][3] = {{1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)},
                                {2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)},
                                {2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)}};
        float *a = affines + 12 * (size_t)i;
        for (int r = 0; r < 3; r++) {
            for (int c = 0; c < 3; c++) a[4 * r + c] = (float)(s * R[r][c]);
            a[4 * r + 3] = (float)t[r];
        }
    }
}

void lc_hip_workload_batch_sizes(uint64_t initstate, uint64_t initseq, uint64_t skip, uint32_t batch_count,
                                 uint32_t lo, uint32_t hi, uint32_t *sizes) {
    // log-uniform in [lo, hi], rounded
    Pcg rng(initstate, initseq);
    rng.skip(skip);
    const double a = std::log((double)lo), b = std::log((double)hi);
    for (uint32_t i = 0; i < batch_count; i++) {
        const double v = std::exp(a + (b - a) * rng.uniform());
        sizes[i] = (uint32_t)std::min<double>(hi, std::max<double>(lo, std::floor(v + 0.5)));
    }
}

void lc_hip_workload_rays_box_sphere(uint64_t initstate, uint64_t initseq, uint64_t skip, uint64_t ray_count, float lo,
                                     float hi, lc_hip_ray *rays) {
    chunked(ray_count, [&](uint64_t b, uint64_t e) {
        Pcg rng(initstate, initseq);
        rng.skip(skip + 5u * b);
        for (uint64_t i = b; i < e; i++) {
            lc_hip_ray &r = rays[i];
            for (int k = 0; k < 3; k++) r.origin[k] = (float)((double)lo + ((double)hi - (double)lo) * rng.uniform());
            double z = 1.0 - 2.0 * rng.uniform();
            double phi = kTwoPi * rng.uniform();
            double rr = std::sqrt(std::max(0.0, 1.0 - z * z));
            r.direction[0] = (float)(rr * std::cos(phi));
            r.direction[1] = (float)(rr * std::sin(phi));
            r.direction[2] = (float)z;
            r.t_min = 0.0f;
            r.t_max = 1e30f;
        }
    });
}

} // extern "C"
