"""GPU tests of Python-DSL kernels (AST -> HIP source -> hiprtc): answers equal the library's
entry points and a float32 restatement of the same program."""
import numpy as np
import pytest

import luisacompute_amd as lc
from luisacompute_amd import host
from oracle import binding as ob
from parity import random_rigid
import dsl_kernels

pytestmark = pytest.mark.gpu
f32 = np.float32


def _scene(device, stream, seed):
    rng = np.random.default_rng(seed)
    meshes = [lc.icosphere(3), lc.soup(1500, initseq=21, extent=2.0, sigma=0.2)]
    gm = [host.mesh_from_arrays(device, stream, v, t) for v, t in meshes]
    accel = device.create_accel()
    for i in range(5):
        accel.emplace_back(gm[i % 2], random_rigid(rng, -3, 3, 0.5, 1.5), visibility_mask=1 << (i % 3))
    stream << accel.build()
    return accel, gm


def test_dsl_trace_kernel(device, stream):
    accel, keep = _scene(device, stream, 70)
    rays = lc.rays_as_floats(lc.rays_box_sphere(30000, initseq=71, lo=-4.0, hi=4.0))
    n = rays.shape[0]
    k = dsl_kernels.trace_kernel.compile(device)
    rb = host.upload(device, stream, host.as_rays(rays))
    hb, ob_ = device.create_buffer(24, n), device.create_buffer(4, n)
    for mask in (0xFF, 0x02):
        stream << k(rb, hb, ob_, accel, mask).dispatch(n)
        hits, occ = np.zeros(n, lc.HIT_DTYPE), np.zeros(n, np.uint32)
        stream << hb.copy_to(hits) << ob_.copy_to(occ) << lc.synchronize()
        g = host.trace_closest(device, stream, accel, rays, vis_mask=mask)
        np.testing.assert_array_equal(hits.view(np.uint32).reshape(-1, 6)[:, :5], g.view(np.uint32).reshape(-1, 6)[:, :5])
        np.testing.assert_array_equal(occ, host.trace_any(device, stream, accel, rays, vis_mask=mask))
        assert (g["inst"] != 0xFFFFFFFF).sum() > 300


def test_dsl_control_flow_kernel(device, stream):
    rng = np.random.default_rng(72)
    xs = rng.uniform(-1.0, 2.0, 5000).astype(f32)
    n_terms, scale = 11, f32(0.75)
    k = dsl_kernels.control_kernel.compile(device)
    xb, outb = host.upload(device, stream, xs), device.create_buffer(4, len(xs))
    stream << k(xb, outb, n_terms, scale).dispatch(len(xs))
    out = np.zeros(len(xs), f32)
    stream << outb.copy_to(out) << lc.synchronize()
    want = np.zeros(len(xs), f32)
    for i, x in enumerate(xs):
        acc = f32(0.0)
        for kk in range(n_terms):
            if kk == 3:
                continue
            term = f32(x * f32(kk)) if x > 0.5 else f32(x - f32(kk))
            if kk > 6 and term < f32(-5.0):
                break
            acc = f32(acc + term)
        for _ in range(2):
            acc = f32(f32(acc * f32(0.5)) + abs(x))
        want[i] = f32(scale * acc)
    np.testing.assert_array_equal(out, want)


def test_dsl_shade_kernel(device, stream):
    accel, keep = _scene(device, stream, 73)
    rays = lc.rays_as_floats(lc.rays_box_sphere(20000, initseq=74, lo=-4.0, hi=4.0))
    n = rays.shape[0]
    k = dsl_kernels.shade_kernel.compile(device)
    rb, pb = host.upload(device, stream, host.as_rays(rays)), device.create_buffer(16, n)
    stream << k(rb, pb, accel).dispatch(n)
    pts = np.zeros((n, 4), f32)
    stream << pb.copy_to(pts) << lc.synchronize()
    g = host.trace_closest(device, stream, accel, rays)
    hit = g["inst"] != 0xFFFFFFFF
    assert (pts[~hit, 3] == -1).all() and hit.sum() > 2000
    o, d, t = rays[hit, :3], rays[hit, 4:7], g["committed_ray_t"][hit]
    p = o + d * t[:, None]
    l = np.sqrt((d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2])
    nrm = (d / l[:, None]) * f32(-1.0)
    q = ob.offset_ray_origin(p, nrm)
    np.testing.assert_array_equal(pts[hit, :3], q)
    shadow = np.zeros((hit.sum(), 8), f32)
    shadow[:, :3], shadow[:, 6], shadow[:, 7] = q, 1.0, 100.0
    np.testing.assert_array_equal(pts[hit, 3], host.trace_any(device, stream, accel, shadow).astype(f32))


@pytest.mark.parametrize("any_hit", [False, True])
def test_dsl_query_kernel(device, stream, any_hit):
    """Ray queries with callbacks written in the DSL equal the built-in query (nearest commit) or
    report committed candidates (first commit)."""
    from test_gpu_shader import _query_scene
    accel, sc, keep = _query_scene(device, stream, seed=75)
    rays = lc.rays_as_floats(lc.rays_box_sphere(30000, initseq=76, lo=-4.0, hi=4.0))
    n = rays.shape[0]
    k = dsl_kernels.query_kernel.compile(device)
    rb, qb = host.upload(device, stream, host.as_rays(rays)), device.create_buffer(24, n)
    stream << k(rb, qb, accel, 0.8, 1 if any_hit else 0).dispatch(n)
    out = np.zeros(n, lc.COMMITTED_HIT_DTYPE)
    stream << qb.copy_to(out) << lc.synchronize()
    pol = dict(surface=lc.QUERY_SURFACE_CUTOUT, cutout=0.8, procedural=lc.QUERY_PROCEDURAL_SPHERE)
    ref = host.query(device, stream, accel, rays, **pol)
    assert (ref["hit_type"] == 1).sum() > 300 and (ref["hit_type"] == 2).sum() > 300
    if not any_hit:
        np.testing.assert_array_equal(out.view(np.uint32), ref.view(np.uint32))
    else:
        np.testing.assert_array_equal(out["hit_type"] != 0, ref["hit_type"] != 0)
        for i in np.flatnonzero(out["hit_type"] != 0)[:2000]:
            c = sc.query_candidate(rays[i], out["inst"][i], out["prim"][i], out["hit_type"][i], **pol)
            assert c is not None and f32(c[0]) == out["committed_ray_t"][i]


def test_dsl_path_tracer_equals_the_library_path_tracer(device, stream):
    """test_path_tracing.cpp's program written in the DSL (bindless heap of vertex / triangle
    buffers, AST -> HIP -> hiprtc) renders the same image bit for bit as the library's megakernel."""
    from luisacompute_amd import workloads
    from test_gpu_pathtracer import Render, options
    room = workloads.BoxRoom(wall_quads=16, block_quads=4)
    R = Render(device, stream, room)
    w, h, spp = 64, 48, 4
    want, n_closest, n_shadow = R(options(w, h, spp, lc.PT_MEGAKERNEL))
    n = len(R.meshes)
    heap = device.create_bindless_array(2 * n)
    for i, m in enumerate(R.meshes):
        heap.emplace(2 * i, m.vb)
        heap.emplace(2 * i + 1, m.tb)
    stream << heap.update()
    xf = host.upload(device, stream, np.concatenate([R.accel.instance_info(i)[0] for i in range(n)]).astype(f32))
    mats = np.array([list(room.albedo[i]) + [1.0 if room.emissive[i] else 0.0] for i in range(n)], f32)
    mb = host.upload(device, stream, mats)
    sc = room.scene()
    vec = lambda name: np.array(getattr(sc, name)[:], f32)
    k = dsl_kernels.path_tracer.compile(device)
    img = device.create_buffer(16, w * h)
    stream << k(img, R.accel, heap, xf, mb, vec("light_position"), vec("light_u"), vec("light_v"),
                vec("light_emission"), vec("camera_origin"), vec("camera_right"), vec("camera_up"),
                vec("camera_forward"), spp, 0, 8, 3, 42, 3 << 32).dispatch(w, h)
    got = np.zeros((h, w, 4), f32)
    stream << img.copy_to(got) << lc.synchronize()
    print(f"[dsl path tracer] {n_closest} closest + {n_shadow} shadow rays; mean {got[..., :3].mean():.4f}", flush=True)
    assert got[..., :3].mean() > 0.01
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))


def test_dsl_integer_division(device, stream):
    """'/' of two unsigned integers is float division, '//' truncating integer division."""
    n, k = 1000, 7
    kern = dsl_kernels.division_kernel.compile(device)
    a, b = device.create_buffer(4, n), device.create_buffer(4, n)
    stream << kern(a, b, k).dispatch(n)
    ft, fl = np.zeros(n, np.float32), np.zeros(n, np.uint32)
    stream << a.copy_to(ft) << b.copy_to(fl) << lc.synchronize()
    i = np.arange(n, dtype=np.uint32)
    np.testing.assert_array_equal(ft, i.astype(np.float32) / np.float32(k))
    np.testing.assert_array_equal(fl, i // k)
    assert ft[3] != 0.0  # 3 / 7 does not truncate to 0
