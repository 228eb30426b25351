"""DeviceInterface::set_name: named meshes, accels and shaders keep working (their builds and
dispatches are bracketed by roctx marker ranges, visible under rocprofv3 --marker-trace); unknown
resource kinds accept a name and ignore it; a bad handle is an error."""
import numpy as np
import pytest

import luisacompute_amd as lc
from luisacompute_amd import host

pytestmark = pytest.mark.gpu


def test_named_build_update_and_dispatch(device, stream):
    verts, tris = lc.icosphere(3)
    vb = host.upload(device, stream, verts)
    tb = host.upload(device, stream, tris)
    mesh = device.create_mesh(vb, tb, option=lc.AccelOption(lc.AccelOption.FAST_TRACE, 1, 1)).set_name("sphere")
    accel = device.create_accel().set_name("scene")
    accel.emplace_back(mesh, lc.IDENTITY)
    stream << mesh.build() << accel.build()
    stream << vb.copy_from(verts * np.float32(1.5)) << mesh.build() << accel.build()  # refit ("update")
    rays = lc.rays_orbit_to_ball(4096)
    hits = host.trace_closest(device, stream, accel, rays)
    assert (hits["prim"] != 0xFFFFFFFF).sum() > 3000
    sh = device.create_shader('''#include "lc_hip_device.h"
struct Params { alignas(16) lchip::dev::Buffer<uint32_t> out; alignas(16) uint4 ls_kid; };
extern "C" __global__ void __launch_bounds__(64) fill(const Params p) {
    if (!lchip::dev::in_dispatch(p.ls_kid)) return;
    p.out[lchip::dev::dispatch_id().x] = 7u;
}''', "fill", (64, 1, 1)).set_name("fill")
    ob = device.create_buffer(4, 256)
    stream << sh(ob).dispatch(256)
    got = np.zeros(256, np.uint32)
    stream << ob.copy_to(got) << lc.synchronize()
    assert (got == 7).all()
    # a buffer (tag 0) takes a name and ignores it; a bad handle fails
    assert lc.lib().lc_hip_set_name(device.handle, 0, ob.handle, b"out") == 0
    assert lc.lib().lc_hip_set_name(device.handle, 7, 0, b"x") != 0
