"""GPU tests of the BLAS update path (MeshBuildCommand with PREFER_UPDATE on a mesh that allows
updates, over the same buffer ranges: src/backends/cuda/cuda_mesh.cpp:41-60). The refit keeps the
BVH topology and recomputes every box; hits after it must equal the oracle on the moved geometry,
and the tree must still satisfy the containment invariants."""
import numpy as np
import pytest

import luisacompute_amd as lc
from luisacompute_amd import host
from oracle import binding as ob
from parity import compare_any, compare_closest

pytestmark = pytest.mark.gpu


def _mesh(device, stream, verts, tris, allow_update=True, compact=True):
    vb = device.create_buffer(16, verts.shape[0])
    tb = device.create_buffer(12, tris.shape[0])
    stream << vb.copy_from(np.ascontiguousarray(verts, np.float32)) << tb.copy_from(np.ascontiguousarray(tris, np.uint32))
    opt = lc.AccelOption(lc.AccelOption.FAST_TRACE, 1 if compact else 0, 1 if allow_update else 0)
    mesh = device.create_mesh(vb, tb, option=opt)
    accel = device.create_accel()
    accel.emplace_back(mesh, lc.IDENTITY)
    accel._keep = [mesh, vb, tb]
    return mesh, accel, vb


def _check(device, stream, accel, verts, tris, rays, label):
    g = host.trace_closest(device, stream, accel, rays)
    ga = host.trace_any(device, stream, accel, rays)
    sc = ob.single_mesh_scene(verts, tris)
    rf = lc.rays_as_floats(host.as_rays(rays))
    rep = compare_closest(g, sc.trace_closest2(rf), rays)
    compare_any(ga, sc.trace_any(rf))
    print(f"[{label}] {rep}", flush=True)
    return rep


def _deform(verts, k):
    """Smooth non-rigid deformation plus a translation (moves every vertex)."""
    v = verts.copy()
    x, y, z = v[:, 0].astype(np.float64), v[:, 1].astype(np.float64), v[:, 2].astype(np.float64)
    s = 1.0 + 0.25 * np.sin(3.0 * x + k) * np.cos(2.0 * y)
    v[:, 0] = (x * s + 0.1 * k).astype(np.float32)
    v[:, 1] = (y * (2.0 - s)).astype(np.float32)
    v[:, 2] = (z * s + 0.05 * np.sin(5.0 * z)).astype(np.float32)
    return v


@pytest.mark.parametrize("compact", [True, False])
def test_refit_matches_oracle_on_moved_geometry(device, stream, compact):
    """Compacting meshes refit from the kept refit state (the build's counters and 32 B per wide
    node; another mesh's build reuses the arena scratch in between), others from their own build
    scratch."""
    verts, tris = lc.icosphere(5)
    rays = lc.rays_orbit_to_ball(1 << 16)
    mesh, accel, vb = _mesh(device, stream, verts, tris, compact=compact)
    device.enable_timing(True)
    device.reset_timing()
    stream << mesh.build() << accel.build()
    _check(device, stream, accel, verts, tris, rays, "refit base")
    m = mesh.memory()
    if compact:
        # refit state only: the build's counters (a slot of a few 256-B units) + 2 float4 per wide node
        slot = m["scratch_bytes"] - 32 * mesh.stats().node_count
        assert 256 <= slot <= 4096 and slot % 256 == 0, m
    # another compacting build overwrites the device's arena scratch before the refits
    ov, ot = lc.soup(30000, initseq=97, extent=3.0, sigma=0.2)
    other, other_accel, _ = _mesh(device, stream, ov, ot)
    stream << other.build()
    stream.synchronize()
    depth0 = mesh.stats().depth
    for k in (1, 2):
        moved = _deform(verts, k)
        stream << vb.copy_from(moved) << mesh.build() << accel.build()
        rep = _check(device, stream, accel, moved, tris, rays, f"refit {k}")
        assert rep["hits"] > 0.9 * rep["rays"]
        errs = mesh.validate(stream)
        assert all(v == 0 for v in errs.values()), errs
        st = mesh.stats()
        assert st.depth == depth0 and st.leaf_primitive_count == len(tris)
    assert device.timing("mesh_build")[1] == 2
    assert device.timing("mesh_refit")[1] == 2
    device.enable_timing(False)


def test_refit_after_scrambling_positions(device, stream):
    """A refit stays correct (only slower) when vertices move far from their build positions."""
    verts, tris = lc.soup(20000, initseq=71, extent=10.0, sigma=0.3)
    rays = lc.rays_box_sphere(1 << 15, initseq=72, lo=-2.0, hi=12.0)
    mesh, accel, vb = _mesh(device, stream, verts, tris)
    stream << mesh.build() << accel.build()
    perm = np.random.default_rng(5).permutation(verts.shape[0])
    moved = np.ascontiguousarray(verts[perm])
    stream << vb.copy_from(moved) << mesh.build() << accel.build()
    _check(device, stream, accel, moved, tris, rays, "refit scrambled")
    errs = mesh.validate(stream)
    assert all(v == 0 for v in errs.values()), errs


def test_force_build_and_no_update_option_rebuild(device, stream):
    verts, tris = lc.icosphere(3)
    rays = lc.rays_orbit_to_ball(4096, initseq=9)
    device.enable_timing(True)
    device.reset_timing()
    mesh, accel, vb = _mesh(device, stream, verts, tris)
    stream << mesh.build() << accel.build()
    moved = _deform(verts, 3)
    stream << vb.copy_from(moved) << mesh.build(lc.BUILD_FORCE_BUILD) << accel.build()
    _check(device, stream, accel, moved, tris, rays, "force build")
    assert device.timing("mesh_refit")[1] == 0 and device.timing("mesh_build")[1] == 2
    # a mesh created without allow_update always rebuilds
    mesh2, accel2, vb2 = _mesh(device, stream, verts, tris, allow_update=False)
    stream << mesh2.build() << accel2.build()
    stream << vb2.copy_from(moved) << mesh2.build() << accel2.build()
    _check(device, stream, accel2, moved, tris, rays, "no-update option")
    assert device.timing("mesh_refit")[1] == 0 and device.timing("mesh_build")[1] == 4
    device.enable_timing(False)
