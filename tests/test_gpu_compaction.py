"""GPU tests of compacting BLAS builds (AccelOption::allow_compaction, resource.h:114; the OptiX
build compacts into a buffer that only grows, src/backends/cuda/cuda_mesh.cpp:42-89).

A compacting build runs the same kernels into the device's build arena and copies the used wide
nodes out, so its tree holds the boxes and leaves of a non-compacting build (node indices come
from atomic reservations and may differ): traces must match the oracle and each other, while the
primitive holds node_count nodes and no build scratch (unless allow_update)."""
import numpy as np
import pytest

import luisacompute_amd as lc
from luisacompute_amd import host
from oracle import binding as ob
from parity import compare_closest, compare_query, random_boxes

pytestmark = pytest.mark.gpu

_NODE_BYTES = 80


def _same(a, b, min_frac=0.999):
    """Fraction of rows equal in every field is at least min_frac (ties may resolve differently
    when node order differs)."""
    eq = np.ones(len(a), bool)
    for f in a.dtype.names:
        x, y = a[f], b[f]
        eq &= (x == y).reshape(len(a), -1).all(axis=1)
    assert eq.mean() >= min_frac, eq.mean()


def _build(device, stream, verts, tris, compact, update):
    vb = device.create_buffer(16, verts.shape[0])
    tb = device.create_buffer(12, tris.shape[0])
    stream << vb.copy_from(np.ascontiguousarray(verts, np.float32)) << tb.copy_from(np.ascontiguousarray(tris, np.uint32))
    mesh = device.create_mesh(vb, tb, option=lc.AccelOption(lc.AccelOption.FAST_TRACE, int(compact), int(update)))
    accel = device.create_accel()
    accel.emplace_back(mesh, lc.IDENTITY)
    stream << mesh.build() << accel.build()
    accel._keep = [mesh, vb, tb]
    return mesh, accel, vb, tb


def test_compacted_mesh_traces_like_the_full_one(device, stream):
    verts, tris = lc.soup(30000, initseq=91, extent=10.0, sigma=0.3)
    rays = lc.rays_box_sphere(1 << 15, initseq=92, lo=-2.0, hi=12.0)
    hits, mem = {}, {}
    for compact, update in ((0, 0), (1, 0), (1, 1)):
        mesh, accel, _, _ = _build(device, stream, verts, tris, compact, update)
        h = host.trace_closest(device, stream, accel, rays)
        hits[compact, update] = h
        st = mesh.stats()
        m = mesh.memory()
        mem[compact, update] = m
        print(f"[compaction {compact}{update}] nodes {st.node_count} memory {m}", flush=True)
        errs = mesh.validate(stream)
        assert all(v == 0 for v in errs.values()), errs
        if compact:
            assert m["node_bytes"] == st.node_count * _NODE_BYTES
            assert (m["scratch_bytes"] > 0) == bool(update)
        else:
            assert m["node_bytes"] >= len(tris) * _NODE_BYTES and m["scratch_bytes"] > 0
    assert mem[1, 0]["total_bytes"] < mem[1, 1]["total_bytes"] < mem[0, 0]["total_bytes"]
    ref = ob.single_mesh_scene(verts, tris).trace_closest2(lc.rays_as_floats(host.as_rays(rays)))
    for key, h in hits.items():
        rep = compare_closest(h, ref, rays)
        print(f"[compaction oracle {key}] {rep}", flush=True)
        _same(h, hits[0, 0])
    assert device.build_arena_bytes() >= len(tris) * _NODE_BYTES


def test_compacted_mesh_after_arena_release(device, stream):
    """A compacting mesh of a new size after the arena is released: the arena is allocated again
    and the accel rebuilt over the new mesh traces its geometry; a forced rebuild of the first
    mesh keeps its node array (same node count) and its answers."""
    small_v, small_t = lc.soup(5000, initseq=93, extent=10.0, sigma=0.3)
    big_v, big_t = lc.soup(40000, initseq=94, extent=10.0, sigma=0.3)
    rays = lc.rays_box_sphere(1 << 14, initseq=95, lo=-2.0, hi=12.0)
    rf = lc.rays_as_floats(host.as_rays(rays))
    mesh, accel, vb, tb = _build(device, stream, small_v, small_t, 1, 0)
    before = host.trace_closest(device, stream, accel, rays)
    device.release_build_arena()
    assert device.build_arena_bytes() == 0
    vb2 = device.create_buffer(16, big_v.shape[0])
    tb2 = device.create_buffer(12, big_t.shape[0])
    stream << vb2.copy_from(big_v) << tb2.copy_from(big_t)
    mesh2 = device.create_mesh(vb2, tb2, option=lc.AccelOption(lc.AccelOption.FAST_TRACE, 1, 0))
    stream << mesh2.build()
    assert device.build_arena_bytes() >= len(big_t) * _NODE_BYTES
    assert mesh2.memory()["node_bytes"] == mesh2.stats().node_count * _NODE_BYTES
    accel.set(0, mesh2, lc.IDENTITY)
    stream << accel.build()
    accel._keep += [mesh2, vb2, tb2]
    rep = compare_closest(host.trace_closest(device, stream, accel, rays),
                          ob.single_mesh_scene(big_v, big_t).trace_closest2(rf), rays)
    print(f"[compaction new mesh] {rep}", flush=True)
    node_bytes = mesh.memory()["node_bytes"]
    stream << mesh.build(lc.BUILD_FORCE_BUILD)
    assert mesh.memory()["node_bytes"] == node_bytes
    accel.set(0, mesh, lc.IDENTITY)
    stream << accel.build()
    after = host.trace_closest(device, stream, accel, rays)
    compare_closest(after, ob.single_mesh_scene(small_v, small_t).trace_closest2(rf), rays)
    _same(after, before)


def test_compacted_refit_matches_oracle(device, stream):
    verts, tris = lc.icosphere(5)
    rays = lc.rays_orbit_to_ball(1 << 15)
    mesh, accel, vb, _ = _build(device, stream, verts, tris, 1, 1)
    moved = verts.copy()
    moved[:, :3] *= np.float32(1.5)
    moved[:, 0] += np.float32(0.25)
    device.enable_timing(True)
    device.reset_timing()
    stream << vb.copy_from(moved) << mesh.build() << accel.build()
    assert device.timing("mesh_refit")[1] == 1
    device.enable_timing(False)
    sc = ob.single_mesh_scene(moved, tris)
    rep = compare_closest(host.trace_closest(device, stream, accel, rays),
                          sc.trace_closest2(lc.rays_as_floats(host.as_rays(rays))), rays)
    assert rep["hits"] > 0.5 * rep["rays"]
    assert mesh.memory()["node_bytes"] == mesh.stats().node_count * _NODE_BYTES


def test_compacted_procedural_queries(device, stream):
    rng = np.random.default_rng(96)
    boxes = random_boxes(rng, 2000, -3.0, 3.0, 0.05, 0.3)
    rays = lc.rays_box_sphere(1 << 14, initseq=97, lo=-4.0, hi=4.0)
    out = {}
    for compact in (0, 1):
        prim = host.procedural_from_aabbs(device, stream, boxes,
                                          option=lc.AccelOption(lc.AccelOption.FAST_TRACE, compact, 0))
        accel = device.create_accel()
        accel.emplace_back(prim, lc.IDENTITY)
        stream << accel.build()
        out[compact] = host.query(device, stream, accel, rays)
        m = prim.memory()
        if compact:
            assert m["node_bytes"] == prim.stats().node_count * _NODE_BYTES and m["scratch_bytes"] == 0
        accel._keep = [prim]
    sc = ob.Scene()
    sc.add(ob.Procedural(boxes), lc.IDENTITY, 0xFF, opaque=True)
    sc.commit()
    rf = lc.rays_as_floats(rays)
    ref = sc.query(rf)
    for compact in (0, 1):
        rep = compare_query(out[compact], ref, sc, rf)
        print(f"[compaction procedural {compact}] {rep}", flush=True)
    _same(out[1], out[0])
