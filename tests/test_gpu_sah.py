"""GPU: the SAH refinement of FAST_TRACE builds (radix-tree clusters of <= 64 primitives rebuilt by
binned SAH and reordered). The refined trees pass the device-side invariant check, trace like the
oracle, cover every primitive once, and visit no more nodes per ray than the plain radix tree."""
import numpy as np
import pytest

import luisacompute_amd as lc
from luisacompute_amd import host
from oracle import binding as ob
from parity import compare_any, compare_closest

pytestmark = pytest.mark.gpu

TRACE = lc.AccelOption(lc.AccelOption.FAST_TRACE, 1, 0)


def _build(device, stream, verts, tris, sah):
    device.set_build_sah(sah)
    try:
        mesh = host.mesh_from_arrays(device, stream, verts, tris, option=TRACE)
        accel = device.create_accel(TRACE)
        accel.emplace_back(mesh, lc.IDENTITY)
        stream << accel.build()
        stream.synchronize()
    finally:
        device.set_build_sah(True)
    return mesh, accel


@pytest.mark.parametrize("ntris", [3, 5, 17, 63, 64, 65, 129, 1000, 40000])
def test_sah_tree_sound_and_exact(device, stream, ntris):
    verts, tris = lc.soup(ntris, initseq=300 + ntris, extent=6.0 if ntris < 1000 else 30.0, sigma=0.6)
    rays = lc.rays_box_sphere(8192, initseq=400 + ntris, lo=-1.0, hi=7.0 if ntris < 1000 else 31.0)
    print(f"[sah] tris={ntris}", flush=True)
    mesh, accel = _build(device, stream, verts, tris, True)
    errs = mesh.validate(stream)
    assert not any(errs.values()), errs
    assert mesh.stats().leaf_primitive_count == ntris
    sc = ob.single_mesh_scene(verts, tris)
    g = host.trace_closest(device, stream, accel, rays)
    compare_closest(g, sc.trace_closest2(lc.rays_as_floats(rays)), rays)
    compare_any(host.trace_any(device, stream, accel, rays), sc.trace_any(lc.rays_as_floats(rays)))


def test_sah_coincident_centroids(device, stream):
    """Triangles sharing one centroid (no binned split exists): split by position, still exact."""
    rng = np.random.default_rng(3)
    n = 200
    verts = np.zeros((3 * n, 4), np.float32)
    d = rng.normal(0.0, 1.0, (n, 3)).astype(np.float32)
    verts[0::3, :3] = d
    verts[1::3, :3] = -d
    verts[2::3, :3] = 0.0
    tris = np.arange(3 * n, dtype=np.uint32).reshape(n, 3)
    rays = lc.rays_orbit_to_ball(4096, orbit=4.0)
    mesh, accel = _build(device, stream, verts, tris, True)
    errs = mesh.validate(stream)
    assert not any(errs.values()), errs
    sc = ob.single_mesh_scene(verts, tris)
    compare_closest(host.trace_closest(device, stream, accel, rays), sc.trace_closest2(lc.rays_as_floats(rays)), rays)


def test_sah_visits_fewer_nodes_on_soup(device, stream):
    """262,144-triangle soup (the benchmark's recipe): same hits as the unrefined tree, fewer node
    visits per ray."""
    n = 1 << 18
    verts, tris = lc.soup(n)
    rays = lc.rays_box_sphere(1 << 16, skip=13 * n)
    res = {}
    for sah in (False, True):
        mesh, accel = _build(device, stream, verts, tris, sah)
        errs = mesh.validate(stream)
        assert not any(errs.values()), errs
        rb = host.upload(device, stream, rays)
        nn, nt = accel.trace_count(stream, rb, len(rays))
        res[sah] = (host.trace_closest(device, stream, accel, rays), nn / len(rays), nt / len(rays))
        print(f"[sah={sah}] nodes/ray {res[sah][1]:.2f} tris/ray {res[sah][2]:.2f}", flush=True)
    sc = ob.single_mesh_scene(verts, tris)
    compare_closest(res[True][0], sc.trace_closest2(lc.rays_as_floats(rays)), rays)
    assert res[True][1] < res[False][1]


def test_sah_tlas_instances(device, stream):
    """A FAST_TRACE TLAS over 300 instances of varied size (box build with SAH) traces like the oracle."""
    verts, tris = lc.icosphere(2)
    mesh = host.mesh_from_arrays(device, stream, verts, tris, option=TRACE)
    accel = device.create_accel(TRACE)
    rng = np.random.default_rng(11)
    sc = ob.Scene()
    om = ob.Mesh(verts, tris)
    for k in range(300):
        aff = np.zeros((3, 4), np.float32)
        aff[:, :3] = np.eye(3, dtype=np.float32) * np.float32(rng.uniform(0.2, 3.0))
        aff[:, 3] = rng.uniform(-30.0, 30.0, 3)
        accel.emplace_back(mesh, aff)
        sc.add(om, aff.reshape(-1))
    stream << accel.build()
    stream.synchronize()
    sc.commit()
    rays = lc.rays_box_sphere(16384, initseq=31, lo=-35.0, hi=35.0)
    compare_closest(host.trace_closest(device, stream, accel, rays), sc.trace_closest2(lc.rays_as_floats(rays)), rays)
