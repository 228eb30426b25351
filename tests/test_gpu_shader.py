"""GPU tests of run-time shaders (create_shader + ShaderDispatchCommand over HIP source compiled by
hiprtc against the device library): the kernels a codegen emits for intersect / intersect_any,
ray queries with the shader's own callbacks, instance ops and the ray helpers, against the oracle
and the library's own entry points."""
import numpy as np
import pytest

import luisacompute_amd as lc
from luisacompute_amd import host
from oracle import binding as ob
from parity import compare_any, compare_closest, compare_query, random_boxes, random_rigid
import shaders

pytestmark = pytest.mark.gpu

MISS = 0xFFFFFFFF


def _mesh_accel(device, stream, seed, n_inst=6, masks=None):
    rng = np.random.default_rng(seed)
    meshes = [lc.icosphere(3), lc.soup(2000, initseq=13, extent=2.0, sigma=0.2)]
    gm = [host.mesh_from_arrays(device, stream, v, t) for v, t in meshes]
    om = [ob.Mesh(v, t) for v, t in meshes]
    accel = device.create_accel()
    sc = ob.Scene()
    masks = masks or [0xFF] * n_inst
    for i in range(n_inst):
        a = random_rigid(rng, -4, 4, 0.5, 1.5)
        accel.emplace_back(gm[i % 2], a, visibility_mask=masks[i], user_id=1000 + i)
        sc.add(om[i % 2], a, masks[i])
    stream << accel.build()
    return accel, sc.commit(), (gm, om)


def _trace_shader(device, stream, accel, rays, mask=0xFF):
    sh = device.create_shader(shaders.TRACE, "trace_kernel")
    n = rays.shape[0]
    rb = host.upload(device, stream, host.as_rays(rays))
    hb, ab = device.create_buffer(24, n), device.create_buffer(1, n)
    stream << sh(rb, hb, ab, accel, np.uint32(mask)).dispatch(n)
    hits, anys = np.zeros(n, lc.HIT_DTYPE), np.zeros(n, np.uint8)
    stream << hb.copy_to(hits) << ab.copy_to(anys) << lc.synchronize()
    return hits, anys


@pytest.mark.parametrize("n_inst,mask", [(1, 0xFF), (6, 0xFF), (6, 0x02)])
def test_trace_shader_matches_oracle_and_entry_points(device, stream, n_inst, mask):
    masks = [0x01, 0x02, 0x03, 0x02, 0x01, 0x02][:n_inst]
    accel, sc, keep = _mesh_accel(device, stream, seed=40 + n_inst, n_inst=n_inst, masks=masks)
    rays = lc.rays_as_floats(lc.rays_box_sphere(40000, initseq=41, lo=-5.0, hi=5.0))
    print(f"[shader trace] instances {n_inst} mask {mask:#x}", flush=True)
    hits, anys = _trace_shader(device, stream, accel, rays, mask)
    rep = compare_closest(hits, sc.trace_closest2(rays, mask=mask), rays)
    compare_any(anys, sc.trace_any(rays, mask=mask))
    # the same answers as the library's persistent trace kernels
    g = host.trace_closest(device, stream, accel, rays, vis_mask=mask)
    np.testing.assert_array_equal(hits.view(np.uint32).reshape(-1, 6)[:, :5], g.view(np.uint32).reshape(-1, 6)[:, :5])
    np.testing.assert_array_equal(anys, host.trace_any(device, stream, accel, rays, vis_mask=mask))
    print(f"[shader trace] {rep}", flush=True)


def _query_scene(device, stream, seed):
    rng = np.random.default_rng(seed)
    v, t = lc.icosphere(3)
    boxes = random_boxes(rng, 200, -2.0, 2.0, 0.1, 0.5)
    gm, gp = host.mesh_from_arrays(device, stream, v, t), host.procedural_from_aabbs(device, stream, boxes)
    om, op = ob.Mesh(v, t), ob.Procedural(boxes)
    accel, sc = device.create_accel(), ob.Scene()
    for g, o, opaque in ((gm, om, False), (gp, op, True), (gm, om, True)):
        a = random_rigid(rng, -3, 3, 0.5, 1.5)
        accel.emplace_back(g, a, opaque=opaque)
        sc.add(o, a, opaque=opaque)
    stream << accel.build()
    return accel, sc.commit(), (gm, gp, om, op)


@pytest.mark.parametrize("any_hit", [False, True])
def test_query_shader_with_its_own_callbacks(device, stream, any_hit):
    """The reference test's callbacks written inside the shader give the built-in query's answers."""
    accel, sc, keep = _query_scene(device, stream, seed=42)
    rays = lc.rays_as_floats(lc.rays_box_sphere(30000, initseq=43, lo=-4.0, hi=4.0))
    sh = device.create_shader(shaders.QUERY, "query_kernel")
    n = rays.shape[0]
    rb = host.upload(device, stream, host.as_rays(rays))
    ob_ = device.create_buffer(24, n)
    stream << sh(rb, ob_, accel, np.float32(0.8), np.uint32(1 if any_hit else 0)).dispatch(n)
    out = np.zeros(n, lc.COMMITTED_HIT_DTYPE)
    stream << ob_.copy_to(out) << lc.synchronize()
    pol = dict(surface=lc.QUERY_SURFACE_CUTOUT, cutout=0.8, procedural=lc.QUERY_PROCEDURAL_SPHERE)
    ref = sc.query(rays, **pol)
    if not any_hit:
        rep = compare_query(out, ref, sc, rays, **pol)
        builtin = host.query(device, stream, accel, rays, **pol)
        np.testing.assert_array_equal(out.view(np.uint32), builtin.view(np.uint32))
        assert rep["surface"] > 300 and rep["procedural"] > 300
        print(f"[shader query] {rep}", flush=True)
    else:
        np.testing.assert_array_equal(out["hit_type"] != 0, ref["hit_type"] != 0)
        for i in np.flatnonzero(out["hit_type"] != 0)[:2000]:
            c = sc.query_candidate(rays[i], out["inst"][i], out["prim"][i], out["hit_type"][i], **pol)
            assert c is not None and np.float32(c[0]) == out["committed_ray_t"][i]


def test_query_shader_even_primitive_callback(device, stream):
    """A callback of the shader's own (commit even primitive ids only) equals the closest hit of a
    mesh whose odd triangles are degenerate (never hit; ids unchanged)."""
    v, t = lc.icosphere(4)
    mesh = host.mesh_from_arrays(device, stream, v, t)
    accel = device.create_accel()
    accel.emplace_back(mesh, lc.IDENTITY, opaque=False)
    stream << accel.build()
    t_even = t.copy()
    t_even[1::2] = t_even[1::2, :1]  # odd triangles collapse to a point
    sc = ob.Scene().add(ob.Mesh(v, t_even)).commit()
    rays = lc.rays_as_floats(lc.rays_orbit_to_ball(50000, initseq=44))
    sh = device.create_shader(shaders.QUERY_EVEN, "even_kernel")
    n = rays.shape[0]
    rb, ob_ = host.upload(device, stream, host.as_rays(rays)), device.create_buffer(24, n)
    stream << sh(rb, ob_, accel).dispatch(n)
    out = np.zeros(n, lc.COMMITTED_HIT_DTYPE)
    stream << ob_.copy_to(out) << lc.synchronize()
    assert (out["prim"][out["hit_type"] != 0] % 2 == 0).all()
    as_hits = np.zeros(n, lc.HIT_DTYPE)
    for f in ("inst", "prim", "bary", "committed_ray_t"):
        as_hits[f] = out[f]
    rep = compare_closest(as_hits, sc.trace_closest2(rays), rays)
    print(f"[shader even-primitive query] {rep}", flush=True)


def test_instance_ops_shader(device, stream):
    accel, sc, keep = _mesh_accel(device, stream, seed=45, n_inst=7)
    sh = device.create_shader(shaders.INSTANCES, "instance_kernel", block=(64, 1, 1))
    tb, ib = device.create_buffer(4, 12 * 7), device.create_buffer(4, 2 * 7)
    stream << sh(tb, ib, accel, np.uint32(5000), np.uint32(0x0F)).dispatch(7)
    tr, ids = np.zeros(12 * 7, np.float32), np.zeros(14, np.uint32)
    stream << tb.copy_to(tr) << ib.copy_to(ids) << lc.synchronize()
    for i in range(7):
        a, vis, uid = (accel.instance_info(i))
        if i % 3 == 0:
            assert uid == 5000 + i and vis == 0x0F  # the shader's writes are visible to reads
        else:
            assert uid == 1000 + i and vis == 0xFF
        np.testing.assert_array_equal(tr[12 * i:12 * i + 12], a)
        assert ids[2 * i] == 1000 + i and ids[2 * i + 1] == 0xFF  # read before this item's write
    # the writes reach traversal at the next build
    stream << accel.build()
    assert accel.instance_info(3)[1:] == (0x0F, 5003)
    rays = lc.rays_as_floats(lc.rays_box_sphere(20000, initseq=46, lo=-5.0, hi=5.0))
    g = host.trace_closest(device, stream, accel, rays, vis_mask=0xF0)
    assert not np.isin(g["inst"], [0, 3, 6]).any()


def test_ray_helpers_shader(device, stream):
    rng = np.random.default_rng(47)
    n = 4096
    x = np.zeros((n, 20), np.float32)
    scale = 10.0 ** rng.uniform(-4, 3, (n, 1))
    x[:, 0:3] = rng.normal(size=(n, 3)) * scale          # p: near 0 and far from it
    nrm = rng.normal(size=(n, 3))
    x[:, 3:6] = nrm / np.linalg.norm(nrm, axis=1, keepdims=True)
    x[:, 6:9] = rng.normal(size=(n, 3))                   # w
    x[:, 9:18] = rng.normal(size=(n, 9)) * 5
    uv = rng.random((n, 2))
    uv[uv.sum(1) > 1] = 1 - uv[uv.sum(1) > 1]
    x[:, 18:20] = uv
    sh = device.create_shader(shaders.HELPERS, "helper_kernel", block=(64, 1, 1))
    ib, obuf = host.upload(device, stream, x), device.create_buffer(4, 17 * n)
    stream << sh(ib, obuf).dispatch(n)
    y = np.zeros((n, 17), np.float32)
    stream << obuf.copy_to(y) << lc.synchronize()
    p, nn, w = x[:, 0:3], x[:, 3:6], x[:, 6:9]
    np.testing.assert_array_equal(y[:, 0:3], ob.offset_ray_origin(p, nn))
    flip = (nn[:, 0] * w[:, 0] + nn[:, 1] * w[:, 1] + nn[:, 2] * w[:, 2]) < 0
    np.testing.assert_array_equal(y[:, 3:6], ob.offset_ray_origin(p, np.where(flip[:, None], -nn, nn)))
    u, v = x[:, 18:19], x[:, 19:20]
    wq = (np.float32(1) - u) - v
    interp = (wq * x[:, 9:12] + u * x[:, 12:15]) + v * x[:, 15:18]
    np.testing.assert_array_equal(y[:, 6:9], interp.astype(np.float32))
    np.testing.assert_array_equal(y[:, 9:12], y[:, 0:3])
    assert (y[:, 12] == 0).all() and (y[:, 16] == np.finfo(np.float32).max).all()
    np.testing.assert_array_equal(y[:, 13:16], nn)


def test_dispatch_argument_errors(device, stream):
    sh = device.create_shader(shaders.HELPERS, "helper_kernel", block=(64, 1, 1))
    b = device.create_buffer(4, 16)
    with pytest.raises(lc.LCError):
        stream << sh((b, 128), b).dispatch(1)  # offset beyond the buffer
    with pytest.raises(lc.LCError):
        device.create_shader(shaders.HELPERS, "no_such_kernel")
    # the failed lookup leaves no error behind for the next launches
    v, t = lc.icosphere(2)
    host.single_mesh_accel(device, stream, v, t)
    unbuilt = device.create_accel()
    sh2 = device.create_shader(shaders.INSTANCES, "instance_kernel", block=(64, 1, 1))
    with pytest.raises(lc.LCError):
        stream << sh2(b, b, unbuilt, np.uint32(0), np.uint32(0)).dispatch(1)


def test_bindless_heap_shader(device, stream):
    """A path tracer's heap access through a bindless array of buffer slots (vertex and triangle
    buffers per instance): hit points interpolated from the slot data equal the host's."""
    rng = np.random.default_rng(48)
    meshes = [lc.icosphere(2), lc.soup(500, initseq=19, extent=2.0, sigma=0.2)]
    gm = [host.mesh_from_arrays(device, stream, v, t) for v, t in meshes]
    accel = device.create_accel()
    heap = device.create_bindless_array(8)
    kinds = [0, 1, 0]
    for i, k in enumerate(kinds):
        accel.emplace_back(gm[k], random_rigid(rng, -3, 3, 0.5, 1.5))
        heap.emplace(2 * i, gm[k].vb)
        heap.emplace(2 * i + 1, gm[k].tb)
    stream << accel.build() << heap.update()
    rays = lc.rays_as_floats(lc.rays_box_sphere(20000, initseq=49, lo=-4.0, hi=4.0))
    sh = device.create_shader(shaders.BINDLESS, "bindless_kernel")
    n = rays.shape[0]
    rb, pb = host.upload(device, stream, host.as_rays(rays)), device.create_buffer(16, n)

    def run():
        stream << sh(rb, pb, accel, heap).dispatch(n)
        out = np.zeros((n, 4), np.float32)
        stream << pb.copy_to(out) << lc.synchronize()
        return out

    out = run()
    h = host.trace_closest(device, stream, accel, rays)
    hit = h["inst"] != 0xFFFFFFFF
    assert hit.sum() > 2000
    assert (out[~hit] == 0).all()
    for i in np.flatnonzero(hit)[:4000]:
        v, t = meshes[kinds[h["inst"][i]]]
        a, b, c = (v[t[h["prim"][i]][j], :3] for j in range(3))
        u, w = np.float32(h["bary"][i, 0]), np.float32(h["bary"][i, 1])
        k = (np.float32(1) - u) - w
        np.testing.assert_array_equal(out[i, :3], (k * a + u * b) + w * c)
        assert out[i, 3] == len(t)
    # re-bind instance 0's triangle slot at an offset of 10 triangles: the slot size follows
    heap.emplace(1, gm[0].tb, 12 * 10)
    heap.remove(7)
    stream << heap.update()
    out = run()
    ntri = len(meshes[0][1]) - 10
    inside = np.flatnonzero(hit & (h["inst"] == 0) & (h["prim"] < ntri))
    beyond = np.flatnonzero(hit & (h["inst"] == 0) & (h["prim"] >= ntri))
    assert len(inside) and (out[inside, 3] == ntri).all() and (out[beyond, 3] == -1).all()


def test_multiple_and_indirect_dispatch(device, stream):
    prod = device.create_shader(shaders.INDIRECT_PRODUCER, "producer", block=(64, 1, 1))
    cons = device.create_shader(shaders.INDIRECT_CONSUMER, "consumer", block=(64, 1, 1))
    out = device.create_buffer(4, 64 * 16)
    zeros = np.zeros(64 * 16, np.uint32)
    # multiple dispatch: launch k has kernel id k
    stream << out.copy_from(zeros) << cons(out).dispatch_multiple([(5,), (2,), (9,)])
    got = np.zeros(64 * 16, np.uint32)
    stream << out.copy_to(got) << lc.synchronize()
    want = zeros.copy()
    for k, n in enumerate((5, 2, 9)):
        want[64 * k:64 * k + n] = 1000 * k + np.arange(n) + 1
    np.testing.assert_array_equal(got, want)
    # indirect: 6 entries written on the device, entries [1, 1 + min(6, 4)) launched
    ind = device.create_buffer(4, 4 + 8 * 8)
    stream << out.copy_from(zeros) << prod(ind, np.uint32(6)).dispatch(6) << cons(out).dispatch_indirect(ind, 1, 4)
    stream << out.copy_to(got) << lc.synchronize()
    want = zeros.copy()
    for i in range(1, 5):
        k, n = 10 + i, 3 * i + 1
        want[64 * (k % 16):64 * (k % 16) + n] = 1000 * k + np.arange(n) + 1
    np.testing.assert_array_equal(got, want)
