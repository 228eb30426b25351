"""GPU parity of ray queries (accel.traverse / traverse_any with surface and procedural candidate
callbacks, include/luisa/dsl/rtx/ray_query.h:11-243) and procedural primitives
(ProceduralPrimitiveBuildCommand, command.h:489-510) against the oracle's restatement of
AccelImpl::ray_query (cpu/accel.rs:582-800) with the callbacks of src/tests/test_procedural.cpp."""
import numpy as np
import pytest

import luisacompute_amd as lc
from luisacompute_amd import host
from oracle import binding as ob
from parity import compare_any, compare_closest, compare_query, random_boxes, random_rigid

pytestmark = pytest.mark.gpu

MISS = 0xFFFFFFFF


def _scene(device, stream, seed, masks=None):
    """Two meshes (a non-opaque icosphere, an opaque soup) and two procedural primitives, five
    instances under random rigid transforms; returns the accel, the oracle scene and references."""
    rng = np.random.default_rng(seed)
    meshes = [lc.icosphere(3), lc.soup(3000, initseq=9, extent=2.0, sigma=0.2)]
    boxes = [random_boxes(rng, 300, -2.0, 2.0, 0.05, 0.4), random_boxes(rng, 40, -1.5, 1.5, 0.3, 1.0)]
    gm = [host.mesh_from_arrays(device, stream, v, t) for v, t in meshes]
    gp = [host.procedural_from_aabbs(device, stream, b) for b in boxes]
    om = [ob.Mesh(v, t) for v, t in meshes]
    op = [ob.Procedural(b) for b in boxes]
    layout = [(gm[0], om[0], False), (gp[0], op[0], True), (gm[1], om[1], True), (gp[1], op[1], False),
              (gm[0], om[0], True)]
    masks = masks or [0xFF] * len(layout)
    accel = device.create_accel()
    sc = ob.Scene()
    for (g, o, opaque), m in zip(layout, masks):
        a = random_rigid(rng, -3, 3, 0.5, 1.5)
        accel.emplace_back(g, a, visibility_mask=m, opaque=opaque)
        sc.add(o, a, m, opaque=opaque)
    stream << accel.build()
    sc.commit()
    return accel, sc, (gm, gp, om, op)


def _rays(n, seed):
    return lc.rays_as_floats(lc.rays_box_sphere(n, initseq=seed, lo=-4.0, hi=4.0))


POLICIES = [
    dict(surface=lc.QUERY_SURFACE_COMMIT, procedural=lc.QUERY_PROCEDURAL_SPHERE),
    dict(surface=lc.QUERY_SURFACE_CUTOUT, procedural=lc.QUERY_PROCEDURAL_SPHERE),
    dict(surface=lc.QUERY_SURFACE_IGNORE, procedural=lc.QUERY_PROCEDURAL_SPHERE),
    dict(surface=lc.QUERY_SURFACE_CUTOUT, procedural=lc.QUERY_PROCEDURAL_IGNORE),
]


@pytest.mark.parametrize("pol", POLICIES, ids=["commit-sphere", "cutout-sphere", "ignore-sphere", "cutout-ignore"])
def test_query_matches_oracle(device, stream, pol):
    accel, sc, keep = _scene(device, stream, seed=21)
    rays = _rays(30000, 31)
    print(f"[query] policy {pol}", flush=True)
    g = host.query(device, stream, accel, rays, cutout=0.8, **pol)
    r = sc.query(rays, cutout=0.8, **pol)
    rep = compare_query(g, r, sc, rays, cutout=0.8, **pol)
    print(f"[query] {rep}", flush=True)
    if pol["procedural"] == lc.QUERY_PROCEDURAL_SPHERE:
        assert rep["procedural"] > 500
    if pol["surface"] != lc.QUERY_SURFACE_IGNORE:
        assert rep["surface"] > 500


def test_query_any_reports_a_committed_candidate(device, stream):
    accel, sc, keep = _scene(device, stream, seed=22)
    rays = _rays(20000, 32)
    pol = dict(surface=lc.QUERY_SURFACE_CUTOUT, cutout=0.8, procedural=lc.QUERY_PROCEDURAL_SPHERE)
    g = host.query(device, stream, accel, rays, terminate_on_first=True, **pol)
    r = sc.query(rays, **pol)
    # a committed candidate exists iff the nearest-commit query hits
    np.testing.assert_array_equal(g["hit_type"] != 0, r["hit_type"] != 0)
    miss = g["hit_type"] == 0
    assert (g["inst"][miss] == MISS).all() and (g["committed_ray_t"][miss] == rays[miss, 7]).all()
    hits = np.flatnonzero(~miss)
    assert len(hits) > 1000
    for i in hits[:3000]:
        c = sc.query_candidate(rays[i], g["inst"][i], g["prim"][i], g["hit_type"][i], **pol)
        assert c is not None, f"ray {i}: {g[i]} is not a committed candidate"
        assert np.float32(c[0]) == g["committed_ray_t"][i]
        assert np.float32(c[1]) == g["bary"][i, 0] and np.float32(c[2]) == g["bary"][i, 1]
    # never nearer than the nearest commit
    assert (g["committed_ray_t"][hits] >= r["t"][hits]).all()


def test_query_visibility_masks(device, stream):
    masks = [0x01, 0x02, 0x01, 0x04, 0x02]
    accel, sc, keep = _scene(device, stream, seed=23, masks=masks)
    rays = _rays(15000, 33)
    pol = dict(surface=lc.QUERY_SURFACE_COMMIT, procedural=lc.QUERY_PROCEDURAL_SPHERE)
    for m in (0x01, 0x02, 0x06, 0x00):
        g = host.query(device, stream, accel, rays, vis_mask=m, **pol)
        r = sc.query(rays, mask=m, **pol)
        rep = compare_query(g, r, sc, rays, mask=m, **pol)
        print(f"[query mask {m:#x}] {rep}", flush=True)
        allowed = {i for i, im in enumerate(masks) if im & m} | {MISS}
        assert set(np.unique(g["inst"]).tolist()) <= allowed


def test_plain_traces_ignore_procedural_instances(device, stream):
    accel, sc, keep = _scene(device, stream, seed=24)
    rays = _rays(20000, 34)
    g = host.trace_closest(device, stream, accel, rays)
    rep = compare_closest(g, sc.trace_closest2(rays), rays)
    assert not np.isin(g["inst"], [1, 3]).any()
    compare_any(host.trace_any(device, stream, accel, rays), sc.trace_any(rays))
    print(f"[plain traces with procedural instances] {rep}", flush=True)


def test_single_procedural_instance(device, stream):
    """One instance: the walk starts in its BLAS (no TLAS); plain traces see nothing."""
    rng = np.random.default_rng(25)
    boxes = random_boxes(rng, 2000, -3.0, 3.0, 0.05, 0.5)
    prim = host.procedural_from_aabbs(device, stream, boxes)
    a = random_rigid(rng, -1, 1, 0.7, 1.3)
    accel = device.create_accel()
    accel.emplace_back(prim, a)
    stream << accel.build()
    op = ob.Procedural(boxes)
    sc = ob.Scene().add(op, a).commit()
    rays = _rays(20000, 35)
    pol = dict(surface=lc.QUERY_SURFACE_COMMIT, procedural=lc.QUERY_PROCEDURAL_SPHERE)
    g = host.query(device, stream, accel, rays, **pol)
    rep = compare_query(g, sc.query(rays, **pol), sc, rays, **pol)
    assert rep["procedural"] > 2000
    print(f"[single procedural] {rep}", flush=True)
    assert (host.trace_closest(device, stream, accel, rays)["prim"] == MISS).all()
    assert not host.trace_any(device, stream, accel, rays).any()
    st = prim.stats()
    assert st.primitive_count == 2000


def test_procedural_rebuild_and_empty(device, stream):
    rng = np.random.default_rng(26)
    b1 = random_boxes(rng, 500, -2.0, 2.0, 0.1, 0.5)
    b2 = random_boxes(rng, 700, -2.0, 2.0, 0.1, 0.5)
    ab = device.create_buffer(24, 700)
    stream << ab.copy_from(np.ascontiguousarray(b1))
    prim = device.create_procedural_primitive(ab, size=24 * 500)
    stream << prim.build()
    accel = device.create_accel()
    accel.emplace_back(prim)
    v, t = lc.icosphere(2)
    mesh = host.mesh_from_arrays(device, stream, v, t)
    accel.emplace_back(mesh, random_rigid(rng, -2, 2))
    stream << accel.build()
    rays = _rays(10000, 36)
    pol = dict(surface=lc.QUERY_SURFACE_COMMIT, procedural=lc.QUERY_PROCEDURAL_SPHERE)
    sc1 = ob.Scene().add(ob.Procedural(b1)).add(ob.Mesh(v, t), accel.instance_info(1)[0]).commit()
    compare_query(host.query(device, stream, accel, rays, **pol), sc1.query(rays, **pol), sc1, rays, **pol)
    # new boxes in the buffer, a larger range: rebuild the primitive, then the accel
    stream << ab.copy_from(np.ascontiguousarray(b2))
    prim.size = 24 * 700
    stream << prim.build() << accel.build()
    sc2 = ob.Scene().add(ob.Procedural(b2)).add(ob.Mesh(v, t), accel.instance_info(1)[0]).commit()
    rep = compare_query(host.query(device, stream, accel, rays, **pol), sc2.query(rays, **pol), sc2, rays, **pol)
    print(f"[procedural rebuild] {rep}", flush=True)
    # an empty primitive: its instance never reports
    prim.size = 0
    stream << prim.build() << accel.build()
    g = host.query(device, stream, accel, rays, **pol)
    assert not (g["hit_type"] == lc.HIT_PROCEDURAL).any()


def test_procedural_build_errors(device, stream):
    ab = device.create_buffer(4, 30)
    stream << ab.copy_from(np.ascontiguousarray(random_boxes(np.random.default_rng(27), 5, -1.0, 1.0, 0.1, 0.2)))
    prim = device.create_procedural_primitive(ab, size=100)  # not a multiple of 24
    with pytest.raises(lc.LCError):
        stream << prim.build()
    prim.size = 240  # beyond the buffer
    with pytest.raises(lc.LCError):
        stream << prim.build()
    prim.size = 96
    stream << prim.build()
    stream.synchronize()
