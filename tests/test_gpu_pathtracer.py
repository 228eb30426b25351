"""GPU parity of the path tracer: megakernel and wavefront forms against the oracle's path
tracer on the same scene and seeds (same image bit for bit; the bar is relative RMSE 1e-3)."""
import numpy as np
import pytest

import luisacompute_amd as lc
from luisacompute_amd import workloads
from oracle import binding as ob

pytestmark = pytest.mark.gpu


def options(w, h, spp, mode, max_depth=8, rr_depth=3, sample_offset=0, chunk=0, sort=1, graph=1, seed=42,
            sequence=3 << 32):
    o = lc.PtOptions()
    o.width, o.height, o.spp, o.sample_offset, o.max_depth, o.rr_depth = w, h, spp, sample_offset, max_depth, rr_depth
    o.seed, o.sequence, o.mode, o.chunk_spp, o.sort, o.graph = seed, sequence, mode, chunk, sort, graph
    return o


class Render:
    def __init__(self, device, stream, room):
        self.device, self.stream, self.room = device, stream, room
        self.accel, self.meshes, self.inst = room.build(device, stream)
        self.scene = room.scene()
        self.counters = device.create_buffer(8, 2)

    def __call__(self, opt):
        w, h = opt.width, opt.height
        img = self.device.create_buffer(16, w * h)
        self.accel.pathtrace(self.stream, self.inst, self.scene, opt, img, counters=self.counters)
        out = np.zeros((h, w, 4), np.float32)
        cnt = np.zeros(2, np.uint64)
        self.stream << img.copy_to(out) << self.counters.copy_to(cnt) << lc.synchronize()
        return out, int(cnt[0]), int(cnt[1])


def oracle_render(room, w, h, spp, **kw):
    meshes = [ob.Mesh(v, t) for v, t in room.meshes]
    sc = ob.Scene()
    for m in meshes:
        sc.add(m)
    sc.commit()
    light = (room.LIGHT_POSITION, room.LIGHT_U, room.LIGHT_V, room.emission)
    return ob.pathtrace(sc, room.albedo, room.emissive, light, room.camera, w, h, spp, **kw)


def rel_rmse(a, b):
    a, b = a[..., :3].astype(np.float64), b[..., :3].astype(np.float64)
    return float(np.sqrt(np.mean((a - b) ** 2)) / max(np.sqrt(np.mean(b ** 2)), 1e-30))


def check(got, want, label):
    diff = np.any(got[..., :3] != want[..., :3], axis=-1)
    r = rel_rmse(got, want)
    print(f"[{label}] pixels differing: {int(diff.sum())}/{diff.size}, relative RMSE {r:.3e}", flush=True)
    assert r <= 1e-3
    return int(diff.sum())


@pytest.fixture(scope="module")
def small(device, stream):
    return Render(device, stream, workloads.BoxRoom(wall_quads=16, block_quads=4))


def test_path_rays_trace_identically(small):
    """Every ray the path tracer traces (offset origins on surfaces, shadow rays ending at the
    light), traced one by one through trace_closest / trace_any: the same results as the oracle."""
    from luisacompute_amd import host
    from parity import compare_any, compare_closest
    room = small.room
    meshes = [ob.Mesh(v, t) for v, t in room.meshes]
    sc = ob.Scene()
    for m in meshes:
        sc.add(m)
    sc.commit()
    light = (room.LIGHT_POSITION, room.LIGHT_U, room.LIGHT_V, room.emission)
    cr, sr = ob.pathtrace_rays(sc, room.albedo, room.emissive, light, room.camera, 64, 36, 8)
    rays = cr.view(lc.RAY_DTYPE).reshape(-1)
    g = host.trace_closest(small.device, small.stream, small.accel, rays)
    ref = sc.trace_closest2(cr)
    exact = (g["inst"] == ref["inst"]) & (g["prim"] == ref["prim"]) & (g["committed_ray_t"] == ref["t"]) & \
        (g["bary"][:, 0] == ref["u"]) & (g["bary"][:, 1] == ref["v"])
    print(f"[path rays] closest: {int((~exact).sum())} of {len(rays)} differ", flush=True)
    for i in np.nonzero(~exact)[0][:8]:
        print("  ray", cr[i].tolist(), "gpu", g[i], "oracle", ref[i], flush=True)
    compare_closest(g, ref, rays)
    srays = sr.view(lc.RAY_DTYPE).reshape(-1)
    ga = host.trace_any(small.device, small.stream, small.accel, srays)
    ra = sc.trace_any(sr)
    print(f"[path rays] shadow: {int((ga != ra).sum())} of {len(srays)} differ", flush=True)
    for i in np.nonzero(ga != ra)[0][:8]:
        print("  shadow ray", sr[i].tolist(), "gpu", ga[i], "oracle", ra[i], flush=True)
    compare_any(ga, ra)


def test_megakernel_matches_oracle(small):
    w, h, spp = 64, 36, 8
    img, nc, ns = small(options(w, h, spp, lc.PT_MEGAKERNEL))
    ref, rc, rs = oracle_render(small.room, w, h, spp)
    assert check(img, ref, "megakernel vs oracle") == 0
    assert (nc, ns) == (rc, rs)
    assert np.all(img[..., 3] == 1.0)


@pytest.mark.parametrize("chunk,sort,graph", [(0, 1, 1), (3, 1, 1), (3, 0, 0), (1, 1, 0), (8, 0, 1)])
def test_wavefront_matches_megakernel(small, chunk, sort, graph):
    w, h, spp = 64, 36, 8
    mega, nc, ns = small(options(w, h, spp, lc.PT_MEGAKERNEL))
    wave, wc, ws = small(options(w, h, spp, lc.PT_WAVEFRONT, chunk=chunk, sort=sort, graph=graph))
    assert check(wave, mega, f"wavefront chunk={chunk} sort={sort} graph={graph}") == 0
    assert (wc, ws) == (nc, ns)


def test_graph_replay_and_sample_offset(small):
    w, h = 40, 24
    a, _, _ = small(options(w, h, 2, lc.PT_WAVEFRONT, chunk=1))
    b, _, _ = small(options(w, h, 2, lc.PT_WAVEFRONT, chunk=1))   # replay of the cached graph
    np.testing.assert_array_equal(a, b)
    s0, _, _ = small(options(w, h, 1, lc.PT_WAVEFRONT, sample_offset=0))
    s1, _, _ = small(options(w, h, 1, lc.PT_WAVEFRONT, sample_offset=1))
    np.testing.assert_array_equal(a[..., :3], ((s0 + s1) / np.float32(2))[..., :3])
    m1, _, _ = small(options(w, h, 1, lc.PT_MEGAKERNEL, sample_offset=1))
    np.testing.assert_array_equal(m1, s1)


def test_rr_and_depth_variants(small):
    w, h, spp = 32, 20, 4
    for max_depth, rr in ((1, 3), (2, 0), (12, 12)):
        wave, wc, ws = small(options(w, h, spp, lc.PT_WAVEFRONT, max_depth=max_depth, rr_depth=rr))
        ref, rc, rs = oracle_render(small.room, w, h, spp, max_depth=max_depth, rr_depth=rr)
        assert check(wave, ref, f"wavefront depth={max_depth} rr={rr} vs oracle") == 0
        assert (wc, ws) == (rc, rs)


def test_bad_options_rejected(small):
    with pytest.raises(lc.LCError):
        small(options(8, 8, 1, lc.PT_MEGAKERNEL, max_depth=13))
    with pytest.raises(lc.LCError):
        small(options(8, 8, 0, lc.PT_WAVEFRONT))


def test_full_room_matches_oracle(device, stream):
    """The full box room (3,538,946 triangles) at a small resolution."""
    room = workloads.BoxRoom()
    r = Render(device, stream, room)
    w, h, spp = 96, 54, 2
    ref, rc, rs = oracle_render(room, w, h, spp)
    for mode in (lc.PT_MEGAKERNEL, lc.PT_WAVEFRONT):
        img, nc, ns = r(options(w, h, spp, mode))
        check(img, ref, f"full room mode={mode} vs oracle")
        assert (nc, ns) == (rc, rs)
