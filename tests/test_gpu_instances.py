"""GPU parity of the two-level (TLAS + BLAS) path: instance transforms, ids and visibility masks."""
import numpy as np
import pytest

import luisacompute_amd as lc
from luisacompute_amd import host
from oracle import binding as ob
from parity import compare_any, compare_closest, random_rigid

pytestmark = pytest.mark.gpu


def rot_z(a):
    c, s = np.cos(a), np.sin(a)
    return np.array([[c, -s, 0, 0], [s, c, 0, 0], [0, 0, 1, 0], [0, 0, 0, 1]], np.float64)


def translate(t):
    m = np.eye(4)
    m[:3, 3] = t
    return m


def test_rtx_scene_closed_form(device, stream):
    """src/tests/test_rtx.cpp:24-28,101-107: one triangle, instance 0 scaled 1.5, instance 1
    translated (-0.25, 0, 0.1) after a 0.5 rad rotation about z. Rays from z = 1 along -z hit
    instance 1 at t = 0.9 where the two overlap, otherwise instance 0 at t = 1."""
    verts = np.array([[-0.5, -0.5, 0, 0], [0.5, -0.5, 0, 0], [0, 0.5, 0, 0]], np.float32)
    tris = np.array([[0, 1, 2]], np.uint32)
    m0 = np.diag([1.5, 1.5, 1.5, 1.0])
    m1 = translate([-0.25, 0.0, 0.1]) @ rot_z(0.5)
    a0 = m0[:3, :4].astype(np.float32).reshape(12)
    a1 = m1[:3, :4].astype(np.float32).reshape(12)
    mesh = host.mesh_from_arrays(device, stream, verts, tris)
    accel = device.create_accel()
    accel.emplace_back(mesh, a0)
    accel.emplace_back(mesh, a1)
    stream << accel.build()
    n = 256
    xs = (np.arange(n) + 0.5) / n * 2 - 1
    X, Y = np.meshgrid(xs, xs)
    rays = np.zeros(n * n, lc.RAY_DTYPE)
    rays["origin"][:, 0] = X.ravel()
    rays["origin"][:, 1] = Y.ravel()
    rays["origin"][:, 2] = 1.0
    rays["direction"] = [0, 0, -1]
    rays["t_max"] = np.finfo(np.float32).max
    g = host.trace_closest(device, stream, accel, rays)

    def inside(m, p, margin):
        w = [(m @ np.array([*v[:3], 1.0]))[:2] for v in verts.astype(np.float64)]
        d = []
        for i in range(3):
            a, b = w[i], w[(i + 1) % 3]
            e = b - a
            d.append((e[0] * (p[:, 1] - a[1]) - e[1] * (p[:, 0] - a[0])) / np.hypot(*e))
        d = np.stack(d, 1)
        return np.all(d > margin, 1), np.all(d > -margin, 1)

    p = np.stack([X.ravel(), Y.ravel()], 1)
    in1, near1 = inside(m1, p, 1e-4)
    in0, near0 = inside(m0, p, 1e-4)
    sure1 = in1
    sure0 = in0 & ~near1
    sure_miss = ~near0 & ~near1
    assert np.all(g["inst"][sure1] == 1) and np.allclose(g["committed_ray_t"][sure1], 0.9, rtol=1e-6)
    assert np.all(g["inst"][sure0] == 0) and np.allclose(g["committed_ray_t"][sure0], 1.0, rtol=1e-6)
    assert np.all(g["inst"][sure_miss] == 0xFFFFFFFF)
    assert sure1.sum() > 1000 and sure0.sum() > 1000 and sure_miss.sum() > 1000
    # and the oracle agrees ray by ray
    om = ob.Mesh(verts, tris)
    sc = ob.Scene().add(om, a0).add(om, a1).commit()
    compare_closest(g, sc.trace_closest2(lc.rays_as_floats(rays)), rays)


def _random_scene(device, stream, ninst, seed):
    rng = np.random.default_rng(seed)
    meshes = [lc.icosphere(1), lc.icosphere(2), lc.icosphere(3), lc.soup(2000, initseq=31, extent=2.0, sigma=0.2)]
    gm = [host.mesh_from_arrays(device, stream, v, t) for v, t in meshes]
    om = [ob.Mesh(v, t) for v, t in meshes]
    accel = device.create_accel()
    sc = ob.Scene()
    masks = []
    for i in range(ninst):
        k = int(rng.integers(len(meshes)))
        a = random_rigid(rng, -10, 10)
        mask = 0xFF if i % 5 else 0x01
        accel.emplace_back(gm[k], a, visibility_mask=mask)
        sc.add(om[k], a, mask)
        masks.append(mask)
    stream << accel.build()
    sc.commit()
    return accel, sc, (gm, om)


@pytest.mark.parametrize("ninst", [2, 3, 17, 200])
def test_random_instances(device, stream, ninst):
    accel, sc, keep = _random_scene(device, stream, ninst, seed=ninst)
    rays = lc.rays_box_sphere(50000, initseq=300 + ninst, lo=-12.0, hi=12.0)
    rf = lc.rays_as_floats(rays)
    for mask in (0xFF, 0x02):
        g = host.trace_closest(device, stream, accel, rays, vis_mask=mask)
        rep = compare_closest(g, sc.trace_closest2(rf, mask=mask), rays)
        ga = host.trace_any(device, stream, accel, rays, vis_mask=mask)
        compare_any(ga, sc.trace_any(rf, mask=mask))
        print(f"[instances {ninst} mask {mask:#x}] {rep}", flush=True)
    st = accel.stats()
    assert st.leaf_primitive_count == ninst


def test_instance_transform_invariance(device, stream):
    """An instanced mesh under affine M gives the same prims as the mesh pre-transformed by M."""
    v, t = lc.icosphere(3)
    rng = np.random.default_rng(3)
    a = random_rigid(rng, -2, 2)
    M = np.vstack([a.reshape(3, 4).astype(np.float64), [0, 0, 0, 1]])
    vw = (M @ np.c_[v[:, :3].astype(np.float64), np.ones(len(v))].T).T
    vw4 = np.c_[vw[:, :3], np.zeros(len(v))].astype(np.float32)
    m = host.mesh_from_arrays(device, stream, v, t)
    mw = host.mesh_from_arrays(device, stream, vw4, t)
    a_inst = device.create_accel()
    a_inst.emplace_back(m, a)
    a_inst.emplace_back(m, a)  # second copy: forces the two-level path; same prims, either inst
    stream << a_inst.build()
    a_world = host.single_mesh_accel(device, stream, vw4, t)
    # object-space rays aimed at the unit ball, carried to world space by M
    ro = lc.rays_orbit_to_ball(20000, initseq=41)
    rays = ro.copy()
    o = np.c_[ro["origin"].astype(np.float64), np.ones(len(ro))]
    rays["origin"] = (M @ o.T).T[:, :3].astype(np.float32)
    dw = (M[:3, :3] @ ro["direction"].astype(np.float64).T).T
    rays["direction"] = (dw / np.linalg.norm(dw, axis=1, keepdims=True)).astype(np.float32)
    g1 = host.trace_closest(device, stream, a_inst, rays)
    g2 = host.trace_closest(device, stream, a_world, rays)
    both = (g1["prim"] != 0xFFFFFFFF) & (g2["prim"] != 0xFFFFFFFF)
    assert both.sum() > 1000
    agree = np.mean(g1["prim"][both] == g2["prim"][both])
    assert agree > 0.999
    np.testing.assert_allclose(g1["committed_ray_t"][both], g2["committed_ray_t"][both], rtol=1e-4)


def test_update_modifications(device, stream):
    """Visibility and transform modifications applied by a later build (accel.rs:324-447)."""
    v, t = lc.icosphere(2)
    m = host.mesh_from_arrays(device, stream, v, t)
    accel = device.create_accel()
    accel.emplace_back(m, np.array([1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0], np.float32))
    accel.emplace_back(m, np.array([1, 0, 0, 5, 0, 1, 0, 0, 0, 0, 1, 0], np.float32))
    stream << accel.build()
    rays = np.zeros(2, lc.RAY_DTYPE)
    rays["origin"] = [[0, 0, 5], [5, 0, 5]]
    rays["direction"] = [0, 0, -1]
    rays["t_max"] = 1e30
    g = host.trace_closest(device, stream, accel, rays)
    assert list(g["inst"]) == [0, 1]
    accel.set_visibility_on_update(1, 0)
    accel.set_transform_on_update(0, np.array([1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, -1], np.float32))
    stream << accel.build()
    g = host.trace_closest(device, stream, accel, rays)
    assert g["inst"][1] == 0xFFFFFFFF
    assert g["inst"][0] == 0 and abs(g["committed_ray_t"][0] - 5.0) < 0.05
    aff, vis, uid = accel.instance_info(0)
    assert aff[11] == -1 and vis == 0xFF


def test_instanced_batches_one_launch(device, stream):
    """The instanced workload scaled down (8 small meshes, 1,024 instances, 64 batches of 16-4096
    rays): one batched launch == one launch per batch == the oracle."""
    from luisacompute_amd import workloads
    w = workloads.Instanced(small=True)
    gm = [host.mesh_from_arrays(device, stream, v, t) for v, t in w.meshes]
    om = [ob.Mesh(v, t) for v, t in w.meshes]
    accel = device.create_accel()
    sc = ob.Scene()
    for a, k in zip(w.affines, w.mesh_ids):
        accel.emplace_back(gm[int(k)], a)
        sc.add(om[int(k)], a)
    stream << accel.build()
    sc.commit()
    rays = [w.batch_rays(j) for j in range(len(w.batch_sizes))]
    rbufs = [host.upload(device, stream, r) for r in rays]
    hbufs = [device.create_buffer(24, len(r)) for r in rays]
    accel.trace_batched(stream, [(rb, hb, len(r)) for rb, hb, r in zip(rbufs, hbufs, rays)])
    got = []
    for hb, r in zip(hbufs, rays):
        out = np.zeros(len(r), lc.HIT_DTYPE)
        stream << hb.copy_to(out)
        got.append(out)
    stream.synchronize()
    allr = np.concatenate(rays)
    allg = np.concatenate(got)
    single = np.concatenate([host.trace_closest(device, stream, accel, r) for r in rays])
    assert np.array_equal(allg, single)
    rep = compare_closest(allg, sc.trace_closest2(lc.rays_as_floats(allr)), allr)
    print(f"[instanced small] {rep}", flush=True)
    assert rep["hits"] > 0.05 * len(allr)
    # any-hit, batched
    abufs = [device.create_buffer(1, len(r)) for r in rays]
    accel.trace_batched(stream, [(rb, ab, len(r)) for rb, ab, r in zip(rbufs, abufs, rays)], any_hit=True)
    ga = []
    for ab, r in zip(abufs, rays):
        out = np.zeros(len(r), np.uint8)
        stream << ab.copy_to(out)
        ga.append(out)
    stream.synchronize()
    compare_any(np.concatenate(ga), sc.trace_any(lc.rays_as_floats(allr)))
