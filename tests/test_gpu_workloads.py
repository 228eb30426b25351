"""GPU parity on the benchmark configurations' own workloads (the inputs the benchmark times):

  * soup: the full 16,777,216-triangle soup built with the benchmark's FAST_BUILD hint (and with
    FAST_TRACE), a 65,536-ray sample of the benchmark's own rays against the oracle, and all
    8,388,608 rays through size-independent properties (any-hit <=> closest-hit is not a miss;
    the sorted and unsorted queues give the same hits);
  * instanced: the real 1,024-instance TLAS over icospheres L3..L8 and the 200k / 500k soups, all
    64 ray batches (capped at 4,096 rays each) in one batched launch against the oracle;
  * serving: the four resident scenes of the serving workload (scaled down) under a Poisson
    request stream submitted while the server runs: every image equals its render alone bit for
    bit, and a band of an icosphere request matches the oracle's path tracer.
"""
import threading
import time

import numpy as np
import pytest

import luisacompute_amd as lc
from luisacompute_amd import host, workloads
from oracle import binding as ob
from parity import compare_any, compare_closest

pytestmark = pytest.mark.gpu

SOUP_TRIS = 1 << 24
SOUP_RAYS = 1 << 23
SAMPLE = 1 << 16


@pytest.fixture(scope="module")
def soup16m():
    verts, tris = lc.soup(SOUP_TRIS)
    return verts, tris


def test_soup_16m_sample_vs_oracle(device, stream, soup16m):
    verts, tris = soup16m
    print(f"[soup16m] oracle build of {len(tris)} triangles", flush=True)
    sc = ob.single_mesh_scene(verts, tris)
    rays = lc.rays_box_sphere(SAMPLE, skip=13 * SOUP_TRIS)  # the first rays of the benchmark's stream
    rf = lc.rays_as_floats(rays)
    ref2, ref_any = sc.trace_closest2(rf), sc.trace_any(rf)
    for hint in (lc.AccelOption.FAST_BUILD, lc.AccelOption.FAST_TRACE):
        accel = host.single_mesh_accel(device, stream, verts, tris, option=lc.AccelOption(hint, 1, 0))
        st = accel._keep[0].stats()
        assert st.leaf_primitive_count == SOUP_TRIS
        errs = accel._keep[0].validate(stream)
        assert not any(errs.values()), f"BVH invariants violated: {errs}"
        g = host.trace_closest(device, stream, accel, rays)
        rep = compare_closest(g, ref2, rays)
        compare_any(host.trace_any(device, stream, accel, rays), ref_any)
        print(f"[soup16m hint {hint}] {rep}", flush=True)
        assert rep["hits"] > SAMPLE // 4
        del accel


def test_soup_16m_all_rays_properties(device, stream, soup16m):
    """All 8,388,608 rays of the benchmark on its BVH: any-hit agrees with closest-hit, and the
    sorted queue (the benchmark's setting) gives the same hits as the unsorted one."""
    verts, tris = soup16m
    accel = host.single_mesh_accel(device, stream, verts, tris,
                                   option=lc.AccelOption(lc.AccelOption.FAST_BUILD, 1, 0))
    rays = lc.rays_box_sphere(SOUP_RAYS, skip=13 * SOUP_TRIS)
    rb = host.upload(device, stream, rays)
    hb, ab = device.create_buffer(24, SOUP_RAYS), device.create_buffer(1, SOUP_RAYS)
    got = {}
    for sort in (False, True):
        device.set_ray_sort("closest", sort)
        device.set_ray_sort("any", sort)
        accel.trace_closest(stream, rb, hb, SOUP_RAYS)
        accel.trace_any(stream, rb, ab, SOUP_RAYS)
        g, ga = np.zeros(SOUP_RAYS, lc.HIT_DTYPE), np.zeros(SOUP_RAYS, np.uint8)
        stream << hb.copy_to(g) << ab.copy_to(ga) << lc.synchronize()
        got[sort] = (g, ga)
    device.set_ray_sort("closest", False)
    device.set_ray_sort("any", False)
    (g0, a0), (g1, a1) = got[False], got[True]
    assert np.array_equal(g0.view(np.uint32).reshape(-1, 6)[:, :5], g1.view(np.uint32).reshape(-1, 6)[:, :5])
    assert np.array_equal(a0, a1)
    assert np.array_equal(a0.astype(bool), g0["prim"] != 0xFFFFFFFF)
    hit = g0["prim"] != 0xFFFFFFFF
    assert hit.mean() > 0.5
    # hits inside the ray interval, valid barycentrics, and the hit point on the reported triangle
    assert np.all(g0["committed_ray_t"][hit] >= 0.0)
    u, v = g0["bary"][hit, 0].astype(np.float64), g0["bary"][hit, 1].astype(np.float64)
    assert np.all(u >= 0) and np.all(v >= 0) and np.all(u + v <= 1 + 1e-6)
    idx = np.flatnonzero(hit)[::997]
    tri = tris[g0["prim"][idx]]
    p0, p1, p2 = (verts[tri[:, k], :3].astype(np.float64) for k in range(3))
    uu, vv = g0["bary"][idx, 0].astype(np.float64), g0["bary"][idx, 1].astype(np.float64)
    pb = (1 - uu - vv)[:, None] * p0 + uu[:, None] * p1 + vv[:, None] * p2
    o = rays["origin"][idx].astype(np.float64)
    d = rays["direction"][idx].astype(np.float64)
    pt = o + g0["committed_ray_t"][idx].astype(np.float64)[:, None] * d
    assert np.max(np.abs(pb - pt)) < 1e-3


def test_instanced_full_meshes_vs_oracle(device, stream):
    """The instanced workload with its real meshes (icospheres L3..L8, soups of 200k and 500k
    triangles), 1,024 instances, the 64 batches capped at 4,096 rays, one batched launch."""
    w = workloads.Instanced()
    print(f"[instanced] meshes {[len(t) for _, t in w.meshes]}", flush=True)
    gm = [host.mesh_from_arrays(device, stream, v, t) for v, t in w.meshes]
    om = [ob.Mesh(v, t) for v, t in w.meshes]
    accel = device.create_accel()
    sc = ob.Scene()
    for a, k in zip(w.affines, w.mesh_ids):
        accel.emplace_back(gm[int(k)], a)
        sc.add(om[int(k)], a)
    stream << accel.build()
    sc.commit()
    cap = 4096
    rays = []
    for j in range(len(w.batch_sizes)):
        skip = w.ray_skip0 + 5 * int(np.sum(w.batch_sizes[:j], dtype=np.uint64))
        rays.append(lc.rays_box_sphere(min(cap, int(w.batch_sizes[j])), initseq=4, skip=skip, lo=-w.extent,
                                       hi=w.extent))
    assert np.array_equal(rays[0], w.batch_rays(0)[:len(rays[0])])
    rbufs = [host.upload(device, stream, r) for r in rays]
    hbufs = [device.create_buffer(24, len(r)) for r in rays]
    abufs = [device.create_buffer(1, len(r)) for r in rays]
    accel.trace_batched(stream, [(rb, hb, len(r)) for rb, hb, r in zip(rbufs, hbufs, rays)])
    accel.trace_batched(stream, [(rb, ab, len(r)) for rb, ab, r in zip(rbufs, abufs, rays)], any_hit=True)
    got, gota = [], []
    for hb, ab, r in zip(hbufs, abufs, rays):
        out, outa = np.zeros(len(r), lc.HIT_DTYPE), np.zeros(len(r), np.uint8)
        stream << hb.copy_to(out) << ab.copy_to(outa)
        got.append(out)
        gota.append(outa)
    stream.synchronize()
    allr = np.concatenate(rays)
    rf = lc.rays_as_floats(allr)
    rep = compare_closest(np.concatenate(got), sc.trace_closest2(rf), allr)
    compare_any(np.concatenate(gota), sc.trace_any(rf))
    print(f"[instanced full] {rep}", flush=True)
    assert rep["hits"] > 0.02 * len(allr)
    assert accel.stats().leaf_primitive_count == 1024


def _alone(device, stream, acc, inst, scene, res, spp, seq):
    o = lc.PtOptions()
    o.width = o.height = res
    o.spp, o.sample_offset, o.max_depth, o.rr_depth = spp, 0, 8, 3
    o.seed, o.sequence, o.mode = 42, seq, lc.PT_MEGAKERNEL
    img = device.create_buffer(16, res * res)
    acc.pathtrace(stream, inst, scene, o, img)
    out = np.zeros((res, res, 4), np.float32)
    stream << img.copy_to(out) << lc.synchronize()
    return out


def test_serving_four_scenes_poisson(device, stream):
    scenes = workloads.serving_scenes(device, stream, small=True)
    stream.synchronize()
    n, rate = 24, 400.0
    reqs = workloads.serving_requests(n, rate)
    for r in reqs:                       # capped for time: 256^2 at most, 4 spp at most
        r["res"] = min(r["res"], 256)
        r["spp"] = min(r["spp"], 4)
    assert len({r["scene"] for r in reqs}) == 4
    srv = lc.Server(device, hbm_budget=3 * 256 * 256 * 16, wave_paths=1 << 18, max_active=16, fetch=(128, 8))
    ids = [srv.add_scene(acc, inst, sc) for _, acc, inst, sc in scenes]
    bufs = [device.create_buffer(16, r["res"] ** 2) for r in reqs]
    rid = {}
    t0 = time.perf_counter()

    def submitter():
        for r in reqs:
            wait = t0 + r["arrival"] - time.perf_counter()
            if wait > 0:
                time.sleep(wait)
            rid[r["index"]] = srv.submit(ids[r["scene"]], r["res"], r["res"], r["spp"], bufs[r["index"]], seed=42,
                                         sequence=r["sequence"])

    th = threading.Thread(target=submitter)
    th.start()
    th.join()
    srv.drain(timeout=300)
    st = srv.stats()
    assert st.requests_done == n
    assert st.peak_bytes <= 3 * 256 * 256 * 16
    admitted_while_busy = 0
    for r in reqs:
        q = srv.query(rid[r["index"]])
        assert q.state == lc.REQUEST_DONE and q.samples_done == r["spp"]
        assert 0 <= q.t_submit <= q.t_admit <= q.t_done
        got = np.zeros((r["res"], r["res"], 4), np.float32)
        stream << bufs[r["index"]].copy_to(got) << lc.synchronize()
        _, acc, inst, sc = scenes[r["scene"]]
        ref = _alone(device, stream, acc, inst, sc, r["res"], r["spp"], r["sequence"])
        assert np.array_equal(got.view(np.uint32), ref.view(np.uint32)), \
            f"request {r['index']} (scene {r['scene']}, {r['res']}^2, {r['spp']} spp) differs from its render alone"
        admitted_while_busy += any(srv.query(rid[o["index"]]).t_admit < q.t_admit < srv.query(rid[o["index"]]).t_done
                                   for o in reqs if o is not r)
    assert admitted_while_busy > 0  # requests joined while others were on the GPU
    # an icosphere request against the oracle's path tracer (8 rows in the middle of the image)
    r = next(r for r in reqs if r["scene"] == 0)
    res, spp = r["res"], r["spp"]
    v, t = lc.icosphere(3)
    osc = ob.Scene().add(ob.Mesh(v, t)).commit()
    _, _, inst, sc = scenes[0]
    light = (sc.light_position[:], sc.light_u[:], sc.light_v[:], sc.light_emission[:])
    cam = dict(origin=sc.camera_origin[:], right=sc.camera_right[:], up=sc.camera_up[:], forward=sc.camera_forward[:])
    row0 = res // 2 - 4
    ref, _, _ = ob.pathtrace(osc, [inst[0].albedo[:]], [0], light, cam, res, res, spp, seed=42, sequence=r["sequence"],
                             pixel_range=(row0 * res, (row0 + 8) * res))
    got = np.zeros((res, res, 4), np.float32)
    stream << bufs[r["index"]].copy_to(got) << lc.synchronize()
    a, b = got[row0:row0 + 8, :, :3].astype(np.float64), ref[row0:row0 + 8, :, :3].astype(np.float64)
    rmse = np.sqrt(np.mean((a - b) ** 2)) / max(np.sqrt(np.mean(b ** 2)), 1e-30)
    print(f"[serving] icosphere band relative RMSE {rmse:.3e}", flush=True)
    assert rmse <= 1e-3
    assert np.mean(b) > 0.0
    srv.close()
