"""GPU tests of streams, timeline events and callbacks (DeviceInterface create_event / signal_event /
wait_event / is_event_completed / synchronize_event and the stream callback,
include/luisa/runtime/rhi/device_interface.h:136-168; semantics of the CUDA backend's timeline
semaphore, src/backends/cuda/cuda_event.cpp:21-63, and CUDAStream::callback,
src/backends/cuda/cuda_stream.cpp:142-170), and of traces running on two streams at once."""
import threading
import time

import numpy as np
import pytest

import luisacompute_amd as lc
from luisacompute_amd import host
from oracle import binding as ob
from parity import compare_any, compare_closest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def soup_scene(device, stream):
    verts, tris = lc.soup(1 << 17, initseq=51)
    accel = host.single_mesh_accel(device, stream, verts, tris)
    rays = lc.rays_box_sphere(1 << 19, initseq=52)
    rb = host.upload(device, stream, rays)
    stream.synchronize()
    return accel, rays, rb, ob.single_mesh_scene(verts, tris)


def test_timeline_values(device):
    s = device.create_stream()
    ev = device.create_event()
    assert ev.is_completed(0)          # the initial value
    assert not ev.is_completed(1)      # nothing signalled yet
    s << ev.signaled(1)
    ev.synchronize(1)
    assert ev.is_completed(1)
    assert not ev.is_completed(2)      # a later value is not reached by an earlier signal
    s << ev.signaled(5)
    ev.synchronize(3)                  # reached through the signal of 5
    assert ev.is_completed(5) and ev.is_completed(4) and not ev.is_completed(6)


def test_cross_stream_wait_orders_work(device, soup_scene):
    """Stream b copies stream a's hits after waiting for a's signal: the copy sees the finished
    trace. Without the wait the copy would race the trace (the buffer is cleared before)."""
    accel, rays, rb, sc = soup_scene
    n = len(rays)
    a, b = device.create_stream(), device.create_stream()
    ev = device.create_event()
    hits = device.create_buffer(24, n)
    copy = device.create_buffer(24, n)
    zeros = np.zeros(n, lc.HIT_DTYPE)
    for rep in range(3):
        a << hits.copy_from(zeros)
        b << copy.copy_from(zeros)
        b.synchronize()
        accel.trace_closest(a, rb, hits, n)
        a << ev.signaled(rep + 1)
        b << ev.waited(rep + 1) << copy.copy_from_buffer(hits)
        got = np.zeros(n, lc.HIT_DTYPE)
        b << copy.copy_to(got) << lc.synchronize()
        assert ev.is_completed(rep + 1)
        ref = np.zeros(n, lc.HIT_DTYPE)
        a << hits.copy_to(ref) << lc.synchronize()
        assert np.array_equal(got.view(np.uint8), ref.view(np.uint8))
    compare_closest(ref, sc.trace_closest2(lc.rays_as_floats(rays)), rays)


def test_wait_issued_before_the_signal(device, soup_scene):
    """A wait for a value nobody signalled yet holds its caller until another thread signals it;
    the waiting stream's later work then runs after the signalled point."""
    accel, rays, rb, _ = soup_scene
    n = len(rays)
    a, b = device.create_stream(), device.create_stream()
    ev = device.create_event()
    hits, copy = device.create_buffer(24, n), device.create_buffer(24, n)
    b << copy.copy_from(np.zeros(n, lc.HIT_DTYPE))
    b.synchronize()
    returned = []

    def waiter():
        ev.wait(b, 7)
        returned.append(time.perf_counter())
        b << copy.copy_from_buffer(hits)

    th = threading.Thread(target=waiter)
    th.start()
    time.sleep(0.3)
    assert not returned               # still waiting: 7 was not signalled
    t_signal = time.perf_counter()
    accel.trace_closest(a, rb, hits, n)
    ev.signal(a, 7)
    th.join(timeout=60)
    assert returned and returned[0] >= t_signal
    got, ref = np.zeros(n, lc.HIT_DTYPE), np.zeros(n, lc.HIT_DTYPE)
    b << copy.copy_to(got) << lc.synchronize()
    a << hits.copy_to(ref) << lc.synchronize()
    assert np.array_equal(got.view(np.uint8), ref.view(np.uint8))


def test_callbacks_run_in_order_after_completion(device, soup_scene):
    accel, rays, rb, _ = soup_scene
    n = len(rays)
    s = device.create_stream()
    ev = device.create_event()
    hits = device.create_buffer(24, n)
    log = []
    done = threading.Event()
    for k in range(4):
        accel.trace_closest(s, rb, hits, n)
        s << ev.signaled(k + 1)
        # the callback runs after the work before it: its own signal has completed
        s << (lambda k=k: log.append((k, ev.is_completed(k + 1))))
    s << (lambda: done.set())
    assert done.wait(60)
    assert log == [(0, True), (1, True), (2, True), (3, True)]
    s.synchronize()


def test_two_streams_trace_concurrently():
    """Two accels traced on two streams with no host synchronisation between the launches (each
    stream has its own slot counters, stack spill and sort scratch): both match the oracle."""
    from luisacompute_amd import autotune
    device = lc.Device(0)  # a device of its own: the tuned table (sorted queues) stays local
    stream = device.create_stream()
    autotune.load_default(device)  # sorted queues: the sort scratch is in use too
    scenes = []
    for k, (ntris, nrays) in enumerate(((1 << 16, 300000), (20000, 200000))):
        verts, tris = lc.soup(ntris, initseq=60 + k)
        accel = host.single_mesh_accel(device, stream, verts, tris)
        rays = lc.rays_box_sphere(nrays, initseq=62 + k)
        scenes.append((accel, rays, host.upload(device, stream, rays), ob.single_mesh_scene(verts, tris)))
    stream.synchronize()
    s = [device.create_stream(), device.create_stream()]
    outs = [(device.create_buffer(24, len(r)), device.create_buffer(1, len(r))) for _, r, _, _ in scenes]
    for rep in range(3):
        for (accel, rays, rb, _), (hb, ab), st in zip(scenes, outs, s):
            accel.trace_closest(st, rb, hb, len(rays))
            accel.trace_any(st, rb, ab, len(rays))
        for (accel, rays, rb, sc), (hb, ab), st in zip(scenes, outs, s):
            g, ga = np.zeros(len(rays), lc.HIT_DTYPE), np.zeros(len(rays), np.uint8)
            st << hb.copy_to(g) << ab.copy_to(ga) << lc.synchronize()
            rf = lc.rays_as_floats(rays)
            compare_closest(g, sc.trace_closest2(rf), rays)
            compare_any(ga, sc.trace_any(rf))


def test_instance_count_change_on_update_only_rebuilds(device, stream):
    """update_instance_buffer_only with a different instance count than the last build: the accel
    is built in full (the old TLAS does not describe the new instances)."""
    v, t = lc.icosphere(2)
    m = host.mesh_from_arrays(device, stream, v, t)
    om = ob.Mesh(v, t)
    accel = device.create_accel()
    accel.emplace_back(m, lc.IDENTITY)
    stream << accel.build()
    a1 = np.array([1, 0, 0, 4, 0, 1, 0, 0, 0, 0, 1, 0], np.float32)
    accel.emplace_back(m, a1)
    stream << accel.update_instance_buffer()
    rays = lc.rays_box_sphere(20000, initseq=70, lo=-3.0, hi=7.0)
    g = host.trace_closest(device, stream, accel, rays)
    sc = ob.Scene().add(om).add(om, a1).commit()
    compare_closest(g, sc.trace_closest2(lc.rays_as_floats(rays)), rays)
    assert (g["inst"] == 1).sum() > 100
