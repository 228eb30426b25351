"""Pins the CPU oracle: published known answers, closed forms, invariants and an independent
double-precision brute force. None of these re-derive the oracle's own formulas."""
import numpy as np
import pytest

import luisacompute_amd as lc
from oracle import binding as ob
from parity import compare_closest

MISS = 0xFFFFFFFF


# ---------------------------------------------------------------- PCG32
def test_pcg32_published_vector():
    """pcg32-demo (pcg-c-basic, O'Neill 2014): pcg32_srandom_r(42, 54), first six outputs."""
    r = ob.Rng(42, 54)
    assert [r.next() for _ in range(6)] == [0xA15C02B7, 0x7B47F409, 0xBA1D3330, 0x83D2F293, 0xBFA4784B, 0xCBED606E]


@pytest.mark.parametrize("delta", [0, 1, 2, 3, 7, 100, 1000, 4097])
def test_pcg32_advance_equals_stepping(delta):
    a, b = ob.Rng(42, 2), ob.Rng(42, 2)
    for _ in range(delta):
        a.next()
    b.advance(delta)
    assert [a.next() for _ in range(4)] == [b.next() for _ in range(4)]


def test_pcg32_float_is_top_24_bits():
    a, b = ob.Rng(7, 9), ob.Rng(7, 9)
    for _ in range(100):
        x = a.next()
        assert b.float() == np.float32((x >> 8) * 2.0 ** -24)


# ---------------------------------------------------------------- generators
@pytest.mark.parametrize("level", [0, 1, 3, 6])
def test_icosphere_topology(level):
    v, t = ob.icosphere(level)
    assert len(t) == 20 * 4 ** level and len(v) == 10 * 4 ** level + 2
    assert np.allclose(np.linalg.norm(v[:, :3].astype(np.float64), axis=1), 1.0, atol=1e-6)
    # closed orientable 2-manifold: every directed edge once, every undirected edge twice
    e = np.concatenate([t[:, [0, 1]], t[:, [1, 2]], t[:, [2, 0]]])
    key = e[:, 0].astype(np.int64) * len(v) + e[:, 1]
    assert len(np.unique(key)) == len(key)
    rev = e[:, 1].astype(np.int64) * len(v) + e[:, 0]
    assert np.array_equal(np.sort(key), np.sort(rev))
    nedges = len(key) // 2
    assert len(v) - nedges + len(t) == 2  # Euler characteristic of the sphere
    p = v[:, :3].astype(np.float64)
    n = np.cross(p[t[:, 1]] - p[t[:, 0]], p[t[:, 2]] - p[t[:, 0]])
    c = (p[t[:, 0]] + p[t[:, 1]] + p[t[:, 2]]) / 3
    assert np.all(np.sum(n * c, 1) > 0)  # outward winding


def test_soup_statistics():
    n = 200000
    v, t = ob.soup(n)
    assert np.array_equal(t.reshape(-1), np.arange(3 * n, dtype=np.uint32))
    p = v[:, :3].astype(np.float64).reshape(n, 3, 3)
    c = p.mean(1)
    off = (p - p.mean(1, keepdims=True))
    assert np.allclose(c.mean(0), 50.0, atol=0.3)
    assert np.allclose(c.std(0), 100 / np.sqrt(12), rtol=0.01)
    # vertex = centroid + N(0, 0.5^2): deviations from the triangle mean have variance sigma^2 * 2/3
    assert np.allclose(off.reshape(-1, 3).std(0), 0.5 * np.sqrt(2 / 3), rtol=0.01)
    z = off.reshape(-1) / off.std()
    assert abs(np.mean(z ** 4) - 3.0) < 0.05  # Gaussian kurtosis


def test_ray_generators():
    r = ob.rays_orbit_to_ball(50000)
    o, d = r[:, :3].astype(np.float64), r[:, 4:7].astype(np.float64)
    assert np.allclose(np.linalg.norm(o, axis=1), 3.0, atol=1e-5)
    assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-6)
    # the line passes through a point of the unit ball
    dist = np.linalg.norm(np.cross(o, d), axis=1)
    assert np.all(dist <= 1.0 + 1e-5)
    assert np.all(np.sum(-o * d, 1) > 0)
    assert np.all(r[:, 3] == 0) and np.all(r[:, 7] == np.float32(1e30))
    r2 = ob.rays_box_sphere(100000)
    o2, d2 = r2[:, :3], r2[:, 4:7].astype(np.float64)
    assert o2.min() >= 0 and o2.max() <= 100
    assert np.allclose(np.linalg.norm(d2, axis=1), 1.0, atol=1e-6)
    assert np.allclose(d2.mean(0), 0, atol=0.01) and np.allclose((d2 ** 2).mean(0), 1 / 3, atol=0.01)


def test_product_generators_match_oracle():
    v, t = lc.icosphere(5)
    v2, t2 = ob.icosphere(5)
    assert np.array_equal(v, v2) and np.array_equal(t, t2)
    assert np.array_equal(lc.rays_as_floats(lc.rays_orbit_to_ball(20000)), ob.rays_orbit_to_ball(20000))
    v, t = lc.soup(100000)
    v2, t2 = ob.soup(100000)
    assert np.array_equal(v, v2) and np.array_equal(t, t2)
    a = lc.rays_as_floats(lc.rays_box_sphere(30000, skip=13 * 100000))
    assert np.array_equal(a, ob.rays_box_sphere(30000, skip=13 * 100000))


# ---------------------------------------------------------------- watertight test
T0 = ([0, 0, 0], [1, 0, 0], [0, 1, 0])


def test_watertight_closed_form_and_bary_convention():
    """bary.x weights vertex 1, bary.y vertex 2 (src/dsl/rtx/hit.cpp:7-9)."""
    hit, t, u, v = ob.watertight([0.25, 0.125, 1], [0, 0, -1], *T0)
    assert hit and t == 1.0 and u == 0.25 and v == 0.125
    hit, t, u, v = ob.watertight([0.25, 0.125, -2], [0, 0, 1], *T0)  # back side: no culling
    assert hit and t == 2.0 and u == 0.25 and v == 0.125
    hit, t, u, v = ob.watertight([1, 0, 1], [0, 0, -1], *T0)  # exactly on vertex 1
    assert hit and (u, v) == (1.0, 0.0)


def test_watertight_range_and_misses():
    assert not ob.watertight([0.25, 0.25, 1], [0, 0, -1], *T0, tmax=0.999)[0]
    assert not ob.watertight([0.25, 0.25, 1], [0, 0, -1], *T0, tmin=1.001)[0]
    assert ob.watertight([0.25, 0.25, 1], [0, 0, -1], *T0, tmin=1.0, tmax=1.0)[0]  # inclusive
    assert not ob.watertight([0.8, 0.8, 1], [0, 0, -1], *T0)[0]   # outside the hypotenuse
    assert not ob.watertight([0.25, 0.25, 1], [1, 0, 0], *T0)[0]  # parallel to the plane
    assert not ob.watertight([0.25, 0.25, 1], [0, 0, 1], *T0)[0]  # pointing away


def test_watertight_shared_edge_never_leaks():
    """Two triangles of a quad share the diagonal; rays through points on it hit one or both."""
    a, b, c, d = [0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]
    rng = np.random.default_rng(0)
    for x in rng.random(2000).astype(np.float32):
        o = [float(x), float(x), 1.0]
        dd = [float(rng.uniform(-0.3, 0.3)) * 0, 0.0, -1.0]
        h1 = ob.watertight(o, dd, a, b, c)[0]
        h2 = ob.watertight(o, dd, a, c, d)[0]
        assert h1 or h2
    # oblique rays through exactly representable edge points
    for k in range(1, 64):
        p = k / 64.0
        for dirv in ([0.3, -0.2, -1.0], [-0.5, 0.4, -1.0]):
            o = [p - dirv[0], p - dirv[1], 1.0]
            assert ob.watertight(o, dirv, a, b, c)[0] or ob.watertight(o, dirv, a, c, d)[0]


# ---------------------------------------------------------------- BVH traversal vs brute force
@pytest.mark.parametrize("ntris,seed", [(1, 3), (2, 4), (5, 5), (64, 6), (1000, 7), (4000, 8)])
def test_bvh_matches_double_brute_force(ntris, seed):
    v, t = ob.soup(ntris, initseq=seed, extent=10.0, sigma=1.0)
    rays = ob.rays_box_sphere(4000, initseq=seed + 50, lo=-2.0, hi=12.0)
    sc = ob.single_mesh_scene(v, t)
    h = sc.trace_closest2(rays)
    bf = ob.brute_closest_f64(v, t, rays)
    hit = bf["prim"] != MISS
    with np.errstate(invalid="ignore"):
        tie = hit & ((bf["t2"] - bf["t"]) <= 1e-6 * np.abs(bf["t"]))
    strict = ~tie
    # grazing rays (a double-precision hit at the very edge) may flip; none occur in these sets
    assert np.array_equal(h["prim"][strict], bf["prim"][strict])
    m = strict & hit
    # single precision resolves t to a few ulp of the coordinates over the incidence cosine,
    # which dominates 1e-5 relative only for origins very close to the surface
    p = v[:, :3].astype(np.float64)
    tri = t[bf["prim"][m]]
    nrm = np.cross(p[tri[:, 1]] - p[tri[:, 0]], p[tri[:, 2]] - p[tri[:, 0]])
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    cos = np.abs(np.sum(nrm * rays[m, 4:7], 1))
    scale = np.max(np.abs(p[tri] - rays[m, None, :3]), axis=(1, 2))
    tol = 1e-5 * bf["t"][m] + 16 * 2.0 ** -24 * scale / cos
    assert np.all(np.abs(h["t"][m] - bf["t"][m]) <= tol)
    assert np.all(np.abs(h["u"][m] - bf["u"][m]) <= 1e-5)
    assert np.all(np.abs(h["v"][m] - bf["v"][m]) <= 1e-5)
    assert np.all(h["t"][~hit] == rays[~hit, 7]) and np.all(h["inst"][~hit] == MISS)
    a = sc.trace_any(rays)
    assert np.array_equal(a.astype(bool), h["prim"] != MISS)


def test_icosphere_closed_form():
    """Rays whose line passes inside the polyhedron's inscribed sphere must hit between the unit
    sphere and the inscribed sphere; lines missing the unit sphere must miss."""
    v, t = ob.icosphere(4)
    p = v[:, :3].astype(np.float64)
    nrm = np.cross(p[t[:, 1]] - p[t[:, 0]], p[t[:, 2]] - p[t[:, 0]])
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    r_in = np.min(np.sum(nrm * p[t[:, 0]], 1))
    rng = np.random.default_rng(2)
    n = 20000
    o = rng.normal(size=(n, 3))
    o = 3 * o / np.linalg.norm(o, axis=1, keepdims=True)
    target = rng.uniform(-1.3, 1.3, (n, 3))
    d = target - o
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((n, 8), np.float32)
    rays[:, :3], rays[:, 4:7], rays[:, 7] = o, d, 1e30
    h = ob.single_mesh_scene(v, t).trace_closest(rays)
    o64, d64 = rays[:, :3].astype(np.float64), rays[:, 4:7].astype(np.float64)
    b = np.sum(o64 * d64, 1)
    dist2 = np.sum(o64 * o64, 1) - b * b
    inside = dist2 < (r_in * 0.999) ** 2
    outside = dist2 > 1.001 ** 2
    assert inside.sum() > 1000 and outside.sum() > 1000
    assert np.all(h["prim"][inside] != MISS) and np.all(h["prim"][outside] == MISS)
    t_unit = -b - np.sqrt(np.maximum(1 - dist2, 0))
    t_in = -b - np.sqrt(np.maximum(r_in ** 2 - dist2, 0))
    tt = h["t"][inside].astype(np.float64)
    assert np.all(tt >= t_unit[inside] - 1e-5) and np.all(tt <= t_in[inside] + 1e-5)


def test_rtx_scene_closed_form_oracle():
    """src/tests/test_rtx.cpp:24-28,101-107 scene through the oracle's two-level path."""
    verts = np.array([[-0.5, -0.5, 0, 0], [0.5, -0.5, 0, 0], [0, 0.5, 0, 0]], np.float32)
    tris = np.array([[0, 1, 2]], np.uint32)
    c, s = np.cos(0.5), np.sin(0.5)
    a0 = np.array([1.5, 0, 0, 0, 0, 1.5, 0, 0, 0, 0, 1.5, 0], np.float32)
    a1 = np.array([c, -s, 0, -0.25, s, c, 0, 0, 0, 0, 1, 0.1], np.float32)
    m = ob.Mesh(verts, tris)
    sc = ob.Scene().add(m, a0).add(m, a1).commit()
    rays = np.zeros((3, 8), np.float32)
    rays[:, 2] = 1.0
    rays[:, 6] = -1.0
    rays[:, 7] = 1e30
    rays[0, :2] = [-0.25, 0.0]   # inside both (instance 1 is nearer)
    rays[1, :2] = [0.6, -0.6]    # only the scaled instance 0
    rays[2, :2] = [0.9, 0.9]     # neither
    h = sc.trace_closest(rays)
    assert h["inst"][0] == 1 and abs(h["t"][0] - 0.9) < 1e-6
    assert h["inst"][1] == 0 and h["t"][1] == 1.0
    assert h["inst"][2] == MISS and h["t"][2] == np.float32(1e30)
    assert list(sc.trace_any(rays)) == [1, 1, 0]
    # visibility: masking instance 1 out exposes instance 0 behind it
    sc2 = ob.Scene().add(m, a0).add(m, a1, mask=0x01).commit()
    assert sc2.trace_closest(rays[:1], mask=0x02)["inst"][0] == 0


def test_permutation_and_clipping_invariants():
    v, t = ob.soup(3000, initseq=13, extent=8.0, sigma=0.8)
    rays = ob.rays_box_sphere(5000, initseq=14, lo=-1, hi=9)
    h = ob.single_mesh_scene(v, t).trace_closest2(rays)
    perm = np.random.default_rng(1).permutation(len(t))
    hp = ob.single_mesh_scene(v, t[perm]).trace_closest2(rays)
    hit = h["prim"] != MISS
    tie = hit & ((h["t2"] - h["t"]) <= 1e-6 * h["t"])
    ok = hit & ~tie
    assert np.array_equal(perm[hp["prim"][ok]], h["prim"][ok])
    assert np.array_equal(hp["t"][ok], h["t"][ok])
    # clipping tmax just below the hit removes that primitive
    r2 = rays.copy()
    r2[ok, 7] = h["t"][ok] * np.float32(1 - 1e-6)
    h2 = ob.single_mesh_scene(v, t).trace_closest(r2)
    assert not np.any(h2["prim"][ok] == h["prim"][ok])


def test_instance_transform_invariance_oracle():
    v, t = ob.icosphere(3)
    a = np.array([0.0, -1.5, 0, 3.0, 1.5, 0.0, 0, -2.0, 0, 0, 1.5, 1.0], np.float32)
    M = np.vstack([a.reshape(3, 4).astype(np.float64), [0, 0, 0, 1]])
    vw = (M @ np.c_[v[:, :3].astype(np.float64), np.ones(len(v))].T).T[:, :3]
    v4 = np.c_[vw, np.zeros(len(v))].astype(np.float32)
    rays = ob.rays_orbit_to_ball(5000, initseq=5)
    o = np.c_[rays[:, :3].astype(np.float64), np.ones(len(rays))]
    rays[:, :3] = (M @ o.T).T[:, :3]
    dw = (M[:3, :3] @ rays[:, 4:7].astype(np.float64).T).T
    rays[:, 4:7] = dw / np.linalg.norm(dw, axis=1, keepdims=True)
    hi = ob.Scene().add(ob.Mesh(v, t), a).commit().trace_closest(rays)
    hw = ob.single_mesh_scene(v4, t).trace_closest(rays)
    both = (hi["prim"] != MISS) & (hw["prim"] != MISS)
    assert both.sum() > 4500
    assert np.mean(hi["prim"][both] == hw["prim"][both]) > 0.999
    assert np.allclose(hi["t"][both], hw["t"][both], rtol=1e-4)


def test_instance_generator():
    """Shoemake rotations: s * R with R orthonormal and det +1; translations and scales in range;
    the product generator equals the oracle's."""
    a, m = ob.instances(1024, 8)
    M = a.reshape(-1, 3, 4)[:, :, :3].astype(np.float64)
    s = np.cbrt(np.linalg.det(M))
    assert np.all(s > 0.5 - 1e-5) and np.all(s < 2.0 + 1e-5)
    R = M / s[:, None, None]
    assert np.allclose(R @ R.transpose(0, 2, 1), np.eye(3), atol=1e-5)
    t = a.reshape(-1, 3, 4)[:, :, 3]
    assert t.min() >= -50 and t.max() <= 50
    assert set(np.unique(m)) == set(range(8))
    # uniform rotations: the rotated x axis is uniform on the sphere (mean ~ 0)
    assert np.allclose(R[:, :, 0].mean(0), 0, atol=0.1)
    a2, m2 = lc.instances(1024, 8)
    assert np.array_equal(a, a2) and np.array_equal(m, m2)
    sz = ob.batch_sizes(5000, 4096, 4194304)
    assert sz.min() >= 4096 and sz.max() <= 4194304
    assert abs(np.mean(np.log(sz)) - 0.5 * (np.log(4096) + np.log(4194304))) < 0.15  # log-uniform
    assert np.array_equal(sz, lc.batch_sizes(5000, 4096, 4194304))
