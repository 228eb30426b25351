"""Writes the golden fixtures of tests/golden/ from the CPU oracle.

Each fixture holds the inputs (vertices, triangles, instance affines, rays) and the oracle's
expected closest hits (inst, prim, u, v, t, second-nearest t2) and any-hit results. Before a
fixture is written, its single-mesh hits are cross-checked against the double-precision
brute force (prim ids outside the 1e-6 tie margin, t/u/v within 1e-5).

Run: python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import binding as ob  # noqa: E402

MISS = 0xFFFFFFFF


def cross_check(v, t, rays, h):
    bf = ob.brute_closest_f64(v, t, rays)
    hit = bf["prim"] != MISS
    with np.errstate(invalid="ignore"):
        tie = hit & ((bf["t2"] - bf["t"]) <= 1e-6 * np.abs(bf["t"]))
    ok = ~tie
    assert np.array_equal(h["prim"][ok], bf["prim"][ok])
    m = ok & hit
    assert np.all(np.abs(h["t"][m] - bf["t"][m]) <= 1e-5 * bf["t"][m] + 1e-6)
    assert np.all(np.abs(h["u"][m] - bf["u"][m]) <= 1e-5) and np.all(np.abs(h["v"][m] - bf["v"][m]) <= 1e-5)


def save(name, verts, tris, affines, masks, rays, ray_mask=0xFF):
    meshes = ob.Mesh(verts, tris)
    sc = ob.Scene()
    for a, mk in zip(affines, masks):
        sc.add(meshes, a, int(mk))
    sc.commit()
    h = sc.trace_closest2(rays, mask=ray_mask)
    anyh = sc.trace_any(rays, mask=ray_mask)
    if len(affines) == 1 and np.array_equal(affines[0], ob.IDENTITY):
        cross_check(verts, tris, rays, h)
    np.savez_compressed(os.path.join(HERE, name), verts=verts, tris=tris, affines=np.asarray(affines, np.float32),
                        masks=np.asarray(masks, np.uint32), rays=rays, ray_mask=np.uint32(ray_mask),
                        inst=h["inst"], prim=h["prim"], u=h["u"], v=h["v"], t=h["t"], t2=h["t2"], any=anyh)
    print(name, "rays", len(rays), "hits", int((h["prim"] != MISS).sum()))


def main():
    # 1. icosphere recipe scaled down: level 3 (1,280 triangles), 4,096 rays, PCG32 (42, 1)
    v, t = ob.icosphere(3)
    save("icosphere_l3_4096.npz", v, t, [ob.IDENTITY], [0xFF], ob.rays_orbit_to_ball(4096))
    # 2. soup recipe scaled down: 2,000 triangles in [0,10]^3, 4,096 rays
    v, t = ob.soup(2000, initstate=42, initseq=2, extent=10.0, sigma=0.5)
    save("soup_2000_4096.npz", v, t, [ob.IDENTITY], [0xFF],
         ob.rays_box_sphere(4096, initstate=42, initseq=2, skip=13 * 2000, lo=0.0, hi=10.0))
    # 3. src/tests/test_rtx.cpp:24-28,101-107 scene: one triangle, two instances, 64 x 64 rays
    verts = np.array([[-0.5, -0.5, 0, 0], [0.5, -0.5, 0, 0], [0, 0.5, 0, 0]], np.float32)
    tris = np.array([[0, 1, 2]], np.uint32)
    c, s = np.cos(0.5), np.sin(0.5)
    a0 = np.array([1.5, 0, 0, 0, 0, 1.5, 0, 0, 0, 0, 1.5, 0], np.float32)
    a1 = np.array([c, -s, 0, -0.25, s, c, 0, 0, 0, 0, 1, 0.1], np.float32)
    n = 64
    xs = ((np.arange(n) + 0.5) / n * 2 - 1).astype(np.float32)
    X, Y = np.meshgrid(xs, xs)
    rays = np.zeros((n * n, 8), np.float32)
    rays[:, 0], rays[:, 1], rays[:, 2] = X.ravel(), Y.ravel(), 1.0
    rays[:, 6] = -1.0
    rays[:, 7] = np.finfo(np.float32).max
    save("rtx_two_instances_64x64.npz", verts, tris, [a0, a1], [0xFF, 0xFF], rays)


if __name__ == "__main__":
    main()
