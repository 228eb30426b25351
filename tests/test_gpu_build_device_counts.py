"""GPU: the BVH build reads nothing back on the host; the bottom-up passes over device-counted
entry lists and the collapse's one-workgroup walk below the levels launched over the GPU give the
same tree whatever the split between them. With the collapse head forced to 1 or 2 levels, almost
every wide level goes through the one-workgroup walk: the trees pass the device-side invariant
check, hold the same node and primitive counts and depth, and trace exactly like the oracle, for
LBVH and SAH-refined builds, triangles and procedural boxes, and builds around the one-workgroup
and leaf-run size limits."""
import numpy as np
import pytest

import luisacompute_amd as lc
from luisacompute_amd import host
from oracle import binding as ob
from parity import compare_any, compare_closest

pytestmark = pytest.mark.gpu

HINTS = {"fast_build": lc.AccelOption(lc.AccelOption.FAST_BUILD, 1, 0),
         "fast_trace": lc.AccelOption(lc.AccelOption.FAST_TRACE, 1, 0)}


def _build(device, stream, verts, tris, option, head):
    device.set_build_collapse_head(head)
    try:
        mesh = host.mesh_from_arrays(device, stream, verts, tris, option=option)
        accel = device.create_accel(option)
        accel.emplace_back(mesh, lc.IDENTITY)
        stream << accel.build()
        stream.synchronize()
    finally:
        device.set_build_collapse_head(0)
    return mesh, accel


@pytest.mark.parametrize("ntris", [700, 1024, 1025, 5000, 300000])
@pytest.mark.parametrize("hint", ["fast_build", "fast_trace"])
def test_collapse_head_and_device_counted_phases(device, stream, ntris, hint):
    extent = 30.0 if ntris >= 5000 else 8.0
    verts, tris = lc.soup(ntris, initseq=500 + ntris, extent=extent, sigma=0.4)
    rays = lc.rays_box_sphere(16384, initseq=600 + ntris, lo=-1.0, hi=extent + 1.0)
    sc = ob.single_mesh_scene(verts, tris)
    want_closest = sc.trace_closest2(lc.rays_as_floats(rays))
    want_any = sc.trace_any(lc.rays_as_floats(rays))
    ref = None
    for head in (0, 1, 2):
        print(f"[device-counted build] tris={ntris} hint={hint} collapse head={head}", flush=True)
        mesh, accel = _build(device, stream, verts, tris, HINTS[hint], head)
        errs = mesh.validate(stream)
        assert not any(errs.values()), errs
        st = mesh.stats()
        got = (st.node_count, st.leaf_primitive_count, st.depth)
        assert st.leaf_primitive_count == ntris
        if ref is None:
            ref = got
        assert got == ref
        compare_closest(host.trace_closest(device, stream, accel, rays), want_closest, rays)
        compare_any(host.trace_any(device, stream, accel, rays), want_any)


def test_procedural_build_with_one_level_head(device, stream):
    """Box builds (procedural primitives) through the same passes, head forced to one level."""
    rng = np.random.default_rng(17)
    n = 20000
    lo = rng.uniform(0.0, 50.0, (n, 3)).astype(np.float32)
    aabbs = np.concatenate([lo, lo + rng.uniform(0.05, 0.5, (n, 3)).astype(np.float32)], axis=1)
    stats = []
    for head in (0, 1):
        device.set_build_collapse_head(head)
        try:
            prim = host.procedural_from_aabbs(device, stream, aabbs)
            stream.synchronize()
        finally:
            device.set_build_collapse_head(0)
        st = prim.stats()
        stats.append((st.node_count, st.leaf_primitive_count, st.depth))
    assert stats[0] == stats[1] and stats[0][1] == n


def test_collapse_head_range(device):
    with pytest.raises(lc.LCError):
        device.set_build_collapse_head(-1)
    with pytest.raises(lc.LCError):
        device.set_build_collapse_head(99)
    device.set_build_collapse_head(98)
    device.set_build_collapse_head(0)
