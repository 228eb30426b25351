"""Traces over device-counted queues, stream capture into graphs, and the path tracer's wavefront
form written in the DSL (camera / shade / resolve kernels over queues, traversal by the persistent
trace): the image equals the DSL megakernel's bit for bit, run eagerly and as graph replays."""
import numpy as np
import pytest

import luisacompute_amd as lc
from luisacompute_amd import host, programs

pytestmark = pytest.mark.gpu
f32 = np.float32


def _soup_accel(device, stream):
    v, t = lc.soup(20000, initseq=31, extent=4.0, sigma=0.3)
    m = host.mesh_from_arrays(device, stream, v, t)
    accel = device.create_accel()
    accel.emplace_back(m)
    stream << accel.build()
    return accel, m


def _hits(stream, hb, n):
    out = np.zeros(n, lc.HIT_DTYPE)
    stream << hb.copy_to(out) << lc.synchronize()
    return out


def test_trace_indirect_reads_the_count_on_the_device(device, stream):
    """trace_closest_indirect / trace_any_indirect trace the first min(count, capacity) rays; the
    other hit slots keep what they held."""
    accel, keep = _soup_accel(device, stream)
    n = 50000
    rays = lc.rays_as_floats(lc.rays_box_sphere(n, initseq=33, lo=-5.0, hi=5.0))
    rb = host.upload(device, stream, host.as_rays(rays))
    want = host.trace_closest(device, stream, accel, rays)
    want_any = host.trace_any(device, stream, accel, rays)
    assert (want["inst"] != 0xFFFFFFFF).sum() > 1000
    sentinel = np.full(n * 6, 0xA5A5A5A5, np.uint32)
    cb = device.create_buffer(4, 4)
    hb, ab = device.create_buffer(24, n), device.create_buffer(1, n)
    for count in (0, 1, 777, 31000, n, n + 5000):
        print(f"[indirect] count={count}", flush=True)
        stream << cb.copy_from(np.array([0, count, 0, 0], np.uint32))
        stream << hb.copy_from(sentinel) << ab.copy_from(np.full(n, 7, np.uint8))
        accel.trace_closest_indirect(stream, rb, hb, cb, 1)
        accel.trace_any_indirect(stream, rb, ab, cb, 1)
        got = _hits(stream, hb, n)
        occ = np.zeros(n, np.uint8)
        stream << ab.copy_to(occ) << lc.synchronize()
        m = min(count, n)
        np.testing.assert_array_equal(got[:m].view(np.uint32).reshape(-1, 6)[:, :5],
                                      want[:m].view(np.uint32).reshape(-1, 6)[:, :5])
        assert (got[m:].view(np.uint32) == 0xA5A5A5A5).all()
        np.testing.assert_array_equal(occ[:m], want_any[:m])
        assert (occ[m:] == 7).all()
    with pytest.raises(lc.LCError, match="count outside"):
        accel.trace_closest_indirect(stream, rb, hb, cb, 4)


def test_captured_traces_replay_with_new_counts(device):
    """A graph of an indirect trace replays several times, with eager traces and host waits
    between the replays; each replay follows the count present when it runs."""
    s = device.create_stream()
    accel, keep = _soup_accel(device, s)
    n = 40000
    rays = lc.rays_as_floats(lc.rays_box_sphere(n, initseq=35, lo=-5.0, hi=5.0))
    rb = host.upload(device, s, host.as_rays(rays))
    want = host.trace_closest(device, s, accel, rays)
    cb, hb = device.create_buffer(4, 1), device.create_buffer(24, n)
    s << cb.copy_from(np.array([n], np.uint32))
    accel.trace_closest_indirect(s, rb, hb, cb)  # eager first: the stream's working memory exists
    s.synchronize()
    with s.capture() as cap:
        accel.trace_closest_indirect(s, rb, hb, cb)
    g = cap.graph
    for k, count in enumerate((n, 123, 29999, n)):
        print(f"[graph replay] {k} count={count}", flush=True)
        s << cb.copy_from(np.array([count], np.uint32))
        s << hb.copy_from(np.full(n * 6, 0xFFFFFFFF, np.uint32))
        g.launch(s)
        got = _hits(s, hb, n)
        np.testing.assert_array_equal(got[:count].view(np.uint32).reshape(-1, 6)[:, :5],
                                      want[:count].view(np.uint32).reshape(-1, 6)[:, :5])
        assert (got[count:].view(np.uint32) == 0xFFFFFFFF).all()
        # eager work on the same stream between replays
        np.testing.assert_array_equal(host.trace_closest(device, s, accel, rays[:1000])["prim"], want["prim"][:1000])
    del g, cap


def test_capture_refuses_to_grow_working_memory(device):
    """A trace captured on a stream that never traced needs working memory the stream does not
    hold: the call fails with a clear error and the capture is dropped."""
    s = device.create_stream()
    accel, keep = _soup_accel(device, s)
    rays = lc.rays_as_floats(lc.rays_box_sphere(1000, initseq=37, lo=-5.0, hi=5.0))
    rb = host.upload(device, s, host.as_rays(rays))
    hb, cb = device.create_buffer(24, 1000), device.create_buffer(4, 1)
    s << cb.copy_from(np.array([1000], np.uint32))
    s.synchronize()
    with pytest.raises(lc.LCError, match="working memory"):
        with s.capture():
            accel.trace_closest_indirect(s, rb, hb, cb)
    # the stream works normally afterwards
    accel.trace_closest_indirect(s, rb, hb, cb)
    got = _hits(s, hb, 1000)
    np.testing.assert_array_equal(got["prim"], host.trace_closest(device, s, accel, rays)["prim"])


def _room(device, stream):
    from luisacompute_amd import workloads
    room = workloads.BoxRoom(wall_quads=16, block_quads=4)
    accel, meshes, inst = room.build(device, stream)
    return room, accel, meshes, inst


@pytest.mark.parametrize("w,h,spp,max_depth,per_wave,sort", [(64, 48, 4, 8, None, False), (64, 48, 4, 8, 1, True),
                                                              (200, 120, 5, 5, 2, True), (200, 120, 5, 5, 2, False)])
def test_dsl_wavefront_equals_dsl_megakernel(device, w, h, spp, max_depth, per_wave, sort):
    """The wavefront form renders the megakernel's image bit for bit: eagerly, and with a wave's
    launches captured once and replayed (renders of one wave, of one sample per wave, and of
    two-sample waves with a one-sample remainder; eager work between the renders), with the
    traces' queues sorted on the device or traced in queue order."""
    s = device.create_stream()
    room, accel, meshes, inst = _room(device, s)
    bind = programs.PathTracerBindings(device, s, accel, inst, room.scene())
    mega = programs.path_tracer.compile(device)
    img = device.create_buffer(16, w * h)
    seq = 3 << 32
    s << mega(*bind.args(img, spp, max_depth=max_depth, sequence=seq)).dispatch(w, h)
    want = np.zeros((h, w, 4), f32)
    s << img.copy_to(want) << lc.synchronize()
    assert want[..., :3].mean() > 0.01
    wf = programs.WavefrontPathTracer(device, s, bind, w, h, samples_per_wave=per_wave, sort_queues=sort)
    before = (device.ray_sort("closest"), device.ray_sort("any"))
    for label, graph in (("eager", False), ("graph", True), ("graph, replay", True)):
        print(f"[dsl wavefront] {w}x{h} spp={spp} depth={max_depth} per wave {wf.samples_per_wave} sort={sort} "
              f"{label}", flush=True)
        out = device.create_buffer(16, w * h)
        wf.render(s, out, spp, max_depth=max_depth, sequence=seq, graph=graph)
        got = np.zeros((h, w, 4), f32)
        s << out.copy_to(got) << lc.synchronize()
        np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
        assert (device.ray_sort("closest"), device.ray_sort("any")) == before  # the device setting is kept
        # eager work between renders on the same stream
        s << mega(*bind.args(img, 1, max_depth=max_depth, sequence=seq)).dispatch(w, h)
        s.synchronize()


def test_destroys_during_a_capture_wait_for_its_end(device):
    """Objects of the device released while a capture is open (by the program or by Python's
    collector) are destroyed when the capture ends; the capture and its replays are unaffected."""
    import gc
    s = device.create_stream()
    accel, keep = _soup_accel(device, s)
    n = 20000
    rays = lc.rays_as_floats(lc.rays_box_sphere(n, initseq=39, lo=-5.0, hi=5.0))
    rb = host.upload(device, s, host.as_rays(rays))
    want = host.trace_closest(device, s, accel, rays)
    cb, hb = device.create_buffer(4, 1), device.create_buffer(24, n)
    s << cb.copy_from(np.array([n], np.uint32))
    accel.trace_closest_indirect(s, rb, hb, cb)
    s.synchronize()
    junk = [device.create_buffer(4, 4096) for _ in range(3)]
    other = device.create_stream()
    v, t = lc.soup(500, initseq=41)
    mesh = host.mesh_from_arrays(device, other, v, t)
    other.synchronize()
    with s.capture() as cap:
        accel.trace_closest_indirect(s, rb, hb, cb)
        del junk, mesh, other
        gc.collect()
    for _ in range(2):
        s << hb.copy_from(np.full(n * 6, 0xFFFFFFFF, np.uint32))
        cap.graph.launch(s)
        got = _hits(s, hb, n)
        np.testing.assert_array_equal(got.view(np.uint32).reshape(-1, 6)[:, :5],
                                      want.view(np.uint32).reshape(-1, 6)[:, :5])
