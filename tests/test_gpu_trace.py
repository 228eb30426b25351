"""GPU parity: HIP BVH build + traversal against the CPU oracle on the same seeded inputs."""
import numpy as np
import pytest

import luisacompute_amd as lc
from luisacompute_amd import host
from oracle import binding as ob
from parity import compare_any, compare_closest

pytestmark = pytest.mark.gpu


def _oracle_scene(verts, tris):
    return ob.single_mesh_scene(verts, tris)


def _check_mesh(device, stream, verts, tris, rays, label):
    print(f"[{label}] tris={len(tris)} rays={len(rays)}", flush=True)
    accel = host.single_mesh_accel(device, stream, verts, tris)
    g = host.trace_closest(device, stream, accel, rays)
    ga = host.trace_any(device, stream, accel, rays)
    sc = _oracle_scene(verts, tris)
    r2 = sc.trace_closest2(lc.rays_as_floats(host.as_rays(rays)))
    ra = sc.trace_any(lc.rays_as_floats(host.as_rays(rays)))
    rep = compare_closest(g, r2, rays)
    compare_any(ga, ra)
    # any-hit <=> closest-hit is not a miss
    assert np.array_equal(ga.astype(bool), g["prim"] != 0xFFFFFFFF)
    print(f"[{label}] {rep}", flush=True)
    return accel, rep


def test_icosphere_l6_full(device, stream):
    """81,920-triangle icosphere, 1,048,576 rays from r=3 toward the unit ball."""
    verts, tris = lc.icosphere(6)
    rays = lc.rays_orbit_to_ball(1 << 20)
    accel, rep = _check_mesh(device, stream, verts, tris, rays, "icosphere-L6")
    assert rep["hits"] > 0.99 * rep["rays"]
    st = accel._keep[0].stats()
    assert st.leaf_primitive_count == len(tris)
    assert 1 <= st.depth <= 96


def test_soup_scaled(device, stream):
    """Soup recipe scaled to 262,144 triangles and 131,072 rays (same generator as the bench)."""
    n = 1 << 18
    verts, tris = lc.soup(n)
    rays = lc.rays_box_sphere(1 << 17, skip=13 * n)
    _, rep = _check_mesh(device, stream, verts, tris, rays, "soup-262k")
    assert rep["hits"] > 0


@pytest.mark.parametrize("ntris", [1, 2, 3, 4, 7, 9, 25, 100, 1000])
def test_small_meshes(device, stream, ntris):
    verts, tris = lc.soup(ntris, initseq=100 + ntris, extent=4.0, sigma=0.7)
    rays = lc.rays_box_sphere(4096, initseq=200 + ntris, lo=-1.0, hi=5.0)
    _check_mesh(device, stream, verts, tris, rays, f"small-{ntris}")


def test_empty_mesh_misses(device, stream):
    verts = np.zeros((0, 4), np.float32)
    tris = np.zeros((0, 3), np.uint32)
    rays = lc.rays_box_sphere(1000)
    accel = host.single_mesh_accel(device, stream, verts, tris)
    g = host.trace_closest(device, stream, accel, rays)
    assert np.all(g["inst"] == 0xFFFFFFFF) and np.all(g["prim"] == 0xFFFFFFFF)
    assert np.all(g["committed_ray_t"] == rays["t_max"])
    assert np.all(g["bary"] == 0)
    assert not host.trace_any(device, stream, accel, rays).any()


def test_degenerate_and_duplicate_triangles(device, stream):
    """Zero-area triangles, duplicated triangles (exact ties) and coincident centroids."""
    verts, tris = lc.soup(500, initseq=7, extent=3.0, sigma=0.5)
    v = verts.copy()
    v[0:3, :3] = v[0, :3]                   # triangle 0: all three vertices equal
    v[3:6, :3] = [[0, 0, 0], [1, 1, 1], [2, 2, 2]]  # triangle 1: collinear
    tris2 = np.concatenate([tris, tris[10:60]])      # 50 duplicated triangles
    rays = lc.rays_box_sphere(20000, initseq=8, lo=-1.0, hi=4.0)
    _check_mesh(device, stream, v, tris2, rays, "degenerate")


def test_axis_aligned_and_clipped_rays(device, stream):
    verts, tris = lc.icosphere(3)
    rng = np.random.default_rng(5)
    n = 4096
    r = np.zeros(n, lc.RAY_DTYPE)
    r["origin"] = rng.uniform(-0.9, 0.9, (n, 3)).astype(np.float32)
    r["origin"][:, 2] = 3.0
    r["direction"] = np.array([0, 0, -1], np.float32)  # zero x/y components
    r["t_min"] = 0.0
    r["t_max"] = np.where(np.arange(n) % 3 == 0, 2.5, 1e30).astype(np.float32)  # some clipped short
    _check_mesh(device, stream, verts, tris, r, "axis-aligned")


def test_tmax_clipping_excludes_hit(device, stream):
    verts, tris = lc.icosphere(4)
    rays = lc.rays_orbit_to_ball(8192, initseq=11)
    accel = host.single_mesh_accel(device, stream, verts, tris)
    g = host.trace_closest(device, stream, accel, rays)
    hit = g["prim"] != 0xFFFFFFFF
    r2 = rays.copy()
    r2["t_max"] = np.where(hit, g["committed_ray_t"] * np.float32(1 - 1e-6), r2["t_max"])
    g2 = host.trace_closest(device, stream, accel, r2)
    assert hit.sum() > 8000 * 0.9
    assert not np.any((g2["prim"] == g["prim"]) & hit)
    # tmin beyond the hit also excludes it
    r3 = rays.copy()
    r3["t_min"] = np.where(hit, g["committed_ray_t"] * np.float32(1 + 1e-6), 0)
    g3 = host.trace_closest(device, stream, accel, r3)
    assert not np.any((g3["prim"] == g["prim"]) & hit)


def test_deterministic_rebuild(device, stream):
    verts, tris = lc.soup(50000, initseq=21)
    rays = lc.rays_box_sphere(50000, initseq=22)
    a1 = host.single_mesh_accel(device, stream, verts, tris)
    g1 = host.trace_closest(device, stream, a1, rays)
    a2 = host.single_mesh_accel(device, stream, verts, tris)
    g2 = host.trace_closest(device, stream, a2, rays)
    assert np.array_equal(g1["prim"], g2["prim"])
    assert np.array_equal(g1["committed_ray_t"], g2["committed_ray_t"])


def test_trace_count_mode(device, stream):
    """Counter-mode traversal (no output buffers): node/triangle counts, and the same accel
    still traces correctly afterwards."""
    verts, tris = lc.icosphere(5)
    rays = lc.rays_orbit_to_ball(65536, initseq=51)
    accel = host.single_mesh_accel(device, stream, verts, tris)
    rb = host.upload(device, stream, rays)
    n_nodes, n_tris = accel.trace_count(stream, rb, len(rays))
    assert n_nodes >= len(rays)          # every ray visits at least the root
    assert n_tris >= len(rays) * 0.99    # nearly every ray tests its hit triangle
    a_nodes, a_tris = accel.trace_count(stream, rb, len(rays), any_hit=True)
    assert a_nodes <= n_nodes and a_tris <= n_tris
    g = host.trace_closest(device, stream, accel, rays)
    assert np.mean(g["prim"] != 0xFFFFFFFF) > 0.99


@pytest.mark.parametrize("kind,n", [("soup", 1), ("soup", 2), ("soup", 37), ("soup", 5000), ("soup", 300000),
                                    ("ico", 6), ("dup", 20000)])
def test_bvh_invariants(device, stream, kind, n):
    """Every primitive in exactly one leaf, leaf triangles inside their quantised slot boxes,
    child nodes inside their parents' slot boxes, indices in range (device-side check)."""
    if kind == "ico":
        verts, tris = lc.icosphere(n)
    else:
        verts, tris = lc.soup(n, initseq=60 + n)
        if kind == "dup":  # many identical centroids: equal Morton codes everywhere
            verts[:, :3] = np.round(verts[:, :3] / 10) * 10
    print(f"[bvh {kind} {n}]", flush=True)
    mesh = host.mesh_from_arrays(device, stream, verts, tris)
    errs = mesh.validate(stream)
    assert all(v == 0 for v in errs.values()), errs
    st = mesh.stats()
    assert st.leaf_primitive_count == len(tris) and st.depth <= 63


@pytest.mark.parametrize("kind", ["closest", "any"])
def test_sorted_queue_matches_unsorted(device, stream, kind):
    """Sorting the ray queue changes only the traversal order, never a result."""
    n = 1 << 17
    verts, tris = lc.soup(n, initseq=71)
    rays = lc.rays_box_sphere(1 << 16, initseq=72)
    accel = host.single_mesh_accel(device, stream, verts, tris)
    fn = host.trace_closest if kind == "closest" else host.trace_any
    try:
        device.set_ray_sort(kind, False)
        a = fn(device, stream, accel, rays)
        device.set_ray_sort(kind, True)
        b = fn(device, stream, accel, rays)
    finally:
        device.set_ray_sort(kind, False)
    assert np.array_equal(a, b)
    sc = ob.single_mesh_scene(verts, tris)
    if kind == "closest":
        compare_closest(b, sc.trace_closest2(lc.rays_as_floats(rays)), rays)
    else:
        compare_any(b, sc.trace_any(lc.rays_as_floats(rays)))


@pytest.mark.parametrize("hint", ["fast_trace", "fast_build"])
def test_leaf_size_hints_match_oracle(device, stream, hint):
    """FAST_TRACE builds single-triangle leaf slots, FAST_BUILD up to three: both trace like the
    oracle, and the trees satisfy the containment invariants."""
    h = lc.AccelOption.FAST_TRACE if hint == "fast_trace" else lc.AccelOption.FAST_BUILD
    verts, tris = lc.soup(1 << 16, initseq=91, extent=20.0, sigma=0.4)
    rays = lc.rays_box_sphere(1 << 15, initseq=92, lo=-2.0, hi=22.0)
    print(f"[hint {hint}]", flush=True)
    opt = lc.AccelOption(h, 1, 0)
    accel = host.single_mesh_accel(device, stream, verts, tris, option=opt)
    g = host.trace_closest(device, stream, accel, rays)
    ga = host.trace_any(device, stream, accel, rays)
    sc = _oracle_scene(verts, tris)
    rf = lc.rays_as_floats(host.as_rays(rays))
    rep = compare_closest(g, sc.trace_closest2(rf), rays)
    compare_any(ga, sc.trace_any(rf))
    assert rep["hits"] > 0
    mesh = accel._keep[0]
    errs = mesh.validate(stream)
    assert all(v == 0 for v in errs.values()), errs
    st = mesh.stats()
    assert st.leaf_primitive_count == len(tris)
    # single-triangle leaves need more wide nodes than three-triangle leaves
    if hint == "fast_trace":
        assert st.node_count > len(tris) // 8
