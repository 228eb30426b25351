"""GPU: the no-allocation closest-hit entry point (lc_hip_accel_trace_closest_noalloc, called
through the C ABI) gives the same hits, bit for bit, as lc_hip_accel_trace_closest on the same
accel and rays; two calls in a row, the second with fewer rays, on a single-mesh accel and on a
two-level accel."""
import numpy as np
import pytest

import luisacompute_amd as lc
from luisacompute_amd import host

pytestmark = pytest.mark.gpu


def _noalloc(device, stream, accel, rays):
    n = len(rays)
    rb = host.upload(device, stream, rays)
    hb = device.create_buffer(24, n)
    stream.synchronize()  # buffers exist before the call: the entry point itself allocates nothing
    rc = lc.lib().lc_hip_accel_trace_closest_noalloc(device.handle, stream.handle, accel.handle, rb.handle, 0,
                                                     hb.handle, 0, n, 0xFF)
    assert rc == 0, lc.lib().lc_hip_last_error()
    out = np.zeros(n, lc.HIT_DTYPE)
    stream << hb.copy_to(out) << lc.synchronize()
    return out


def _same(device, stream, accel, rays, label):
    print(f"[noalloc {label}] rays={len(rays)}", flush=True)
    want = host.trace_closest(device, stream, accel, rays)
    got = _noalloc(device, stream, accel, rays)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32)), label
    return int((want["prim"] != 0xFFFFFFFF).sum())


def test_noalloc_single_mesh_two_calls(device, stream):
    n = 1 << 18
    verts, tris = lc.soup(n)
    accel = host.single_mesh_accel(device, stream, verts, tris)
    big = lc.rays_box_sphere(1 << 17, skip=13 * n)
    small = lc.rays_box_sphere(5000, initseq=77)
    assert _same(device, stream, accel, big, "soup-262k big") > 0
    assert _same(device, stream, accel, small, "soup-262k small") > 0


def test_noalloc_two_level_two_calls(device, stream):
    verts, tris = lc.icosphere(4)
    mesh = host.mesh_from_arrays(device, stream, verts, tris)
    accel = device.create_accel()
    rng = np.random.default_rng(5)
    for k in range(6):
        aff = np.zeros((3, 4), np.float32)
        aff[:, :3] = np.eye(3, dtype=np.float32) * (0.5 + k * 0.2)
        aff[:, 3] = rng.uniform(-2.0, 2.0, 3)
        accel.emplace_back(mesh, aff)
    stream << accel.build()
    stream.synchronize()
    big = lc.rays_orbit_to_ball(1 << 16)
    small = lc.rays_orbit_to_ball(3000, initseq=9)
    assert _same(device, stream, accel, big, "instances big") > 0
    assert _same(device, stream, accel, small, "instances small") > 0


def test_noalloc_rejects_short_hit_buffer(device, stream):
    verts, tris = lc.icosphere(2)
    accel = host.single_mesh_accel(device, stream, verts, tris)
    rays = lc.rays_orbit_to_ball(100)
    rb = host.upload(device, stream, rays)
    hb = device.create_buffer(24, 50)
    stream.synchronize()
    rc = lc.lib().lc_hip_accel_trace_closest_noalloc(device.handle, stream.handle, accel.handle, rb.handle, 0,
                                                     hb.handle, 0, 100, 0xFF)
    assert rc != 0
    assert b"hit buffer" in lc.lib().lc_hip_last_error()
