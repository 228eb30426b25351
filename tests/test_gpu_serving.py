"""GPU tests of the render-serving loop: every request's image equals a render of the request alone
(bit for bit), with requests of mixed scenes, resolutions and sample counts batched together,
admission limited by the HBM budget, and waves smaller than one request."""
import numpy as np
import pytest

import luisacompute_amd as lc
from luisacompute_amd import host, workloads

pytestmark = pytest.mark.gpu


def sphere_scene(device, stream):
    v, t = lc.icosphere(3)
    accel = host.single_mesh_accel(device, stream, v, t)
    inst = [lc.PtInstance()]
    inst[0].albedo[:] = [0.7, 0.5, 0.3]
    s = lc.PtScene()
    s.light_position[:] = [-1.0, 3.0, -1.0]
    s.light_u[:] = [2.0, 0.0, 0.0]
    s.light_v[:] = [0.0, 0.0, 2.0]
    s.light_emission[:] = [10.0, 10.0, 10.0]
    s.camera_origin[:] = [0.0, 0.0, 3.5]
    s.camera_right[:] = [0.5, 0.0, 0.0]
    s.camera_up[:] = [0.0, 0.5, 0.0]
    s.camera_forward[:] = [0.0, 0.0, -1.0]
    return accel, inst, s


def alone(device, stream, accel, inst, scene, w, h, spp, seed, sequence, max_depth=8, rr_depth=3):
    o = lc.PtOptions()
    o.width, o.height, o.spp, o.sample_offset, o.max_depth, o.rr_depth = w, h, spp, 0, max_depth, rr_depth
    o.seed, o.sequence, o.mode = seed, sequence, lc.PT_MEGAKERNEL
    img = device.create_buffer(16, w * h)
    accel.pathtrace(stream, inst, scene, o, img)
    out = np.zeros((h, w, 4), np.float32)
    stream << img.copy_to(out) << lc.synchronize()
    return out


@pytest.fixture(scope="module")
def scenes(device, stream):
    room = workloads.BoxRoom(wall_quads=16, block_quads=4)
    r_accel, _, r_inst = room.build(device, stream)
    s_accel, s_inst, s_scene = sphere_scene(device, stream)
    stream.synchronize()
    return [(r_accel, r_inst, room.scene()), (s_accel, s_inst, s_scene)]


def test_requests_in_the_mix_equal_alone(device, stream, scenes):
    # budget for about three 64x48 accumulators at once; waves of 20,000 paths (smaller than
    # some requests' sample counts, so those span several waves)
    srv = lc.Server(device, hbm_budget=3 * 64 * 48 * 16 + 4096, wave_paths=20000, max_active=8)
    ids = [srv.add_scene(a, i, s) for a, i, s in scenes]
    rng = np.random.default_rng(7)
    reqs = []
    for k in range(12):
        sc = ids[k % 2]
        w, h = [(32, 24), (64, 48), (40, 40), (17, 9)][k % 4]
        spp = int(rng.choice([1, 3, 8]))
        buf = device.create_buffer(16, w * h)
        rid = srv.submit(sc, w, h, spp, buf, seed=42 + k, sequence=(5 << 32) + 1000 * k)
        reqs.append((rid, sc, w, h, spp, 42 + k, (5 << 32) + 1000 * k, buf))
    srv.drain(timeout=300)
    st = srv.stats()
    assert st.requests_done == len(reqs)
    # all 12 accumulators (about 360 KB) never fit the budget at once; small requests may stay
    # admitted alongside larger ones up to max_active
    assert 1 <= st.peak_active <= 8
    assert st.peak_bytes <= 3 * 64 * 48 * 16 + 4096
    assert st.waves >= 2
    for rid, sc, w, h, spp, seed, seq, buf in reqs:
        q = srv.query(rid)
        assert q.state == lc.REQUEST_DONE and q.samples_done == spp
        assert 0 <= q.t_submit <= q.t_admit <= q.t_done
        got = np.zeros((h, w, 4), np.float32)
        stream << buf.copy_to(got) << lc.synchronize()
        a, i, s = scenes[sc]
        ref = alone(device, stream, a, i, s, w, h, spp, seed, seq)
        assert np.array_equal(got, ref), f"request {rid} ({w}x{h}, {spp} spp, scene {sc}) differs from alone"
    srv.close()


def test_late_arrivals_join_running_waves(device, stream, scenes):
    srv = lc.Server(device, hbm_budget=1 << 24, wave_paths=1 << 16, max_active=16)
    sid = srv.add_scene(*scenes[0])
    big = device.create_buffer(16, 128 * 96)
    r0 = srv.submit(sid, 128, 96, 16, big, seed=1, sequence=3 << 32)
    late = []
    for k in range(4):
        b = device.create_buffer(16, 24 * 16)
        late.append((srv.submit(sid, 24, 16, 2, b, seed=9 + k, sequence=7 << 32), b))
    srv.drain(timeout=300)
    q0 = srv.query(r0)
    # the small requests finish while the big one is still running (continuous batching)
    assert all(srv.query(r).t_done <= q0.t_done for r, _ in late)
    got = np.zeros((96, 128, 4), np.float32)
    stream << big.copy_to(got) << lc.synchronize()
    ref = alone(device, stream, *scenes[0], 128, 96, 16, 1, 3 << 32)
    assert np.array_equal(got, ref)
    srv.close()


def test_rejects_requests_that_cannot_fit(device, stream, scenes):
    srv = lc.Server(device, hbm_budget=1 << 16, wave_paths=4096, max_active=4)
    sid = srv.add_scene(*scenes[1])
    buf = device.create_buffer(16, 128 * 128)
    with pytest.raises(lc.LCError):
        srv.submit(sid, 128, 128, 1, buf)      # one sample exceeds the wave capacity
    with pytest.raises(lc.LCError):
        srv.submit(sid + 5, 8, 8, 1, buf)      # unknown scene
    srv.close()
