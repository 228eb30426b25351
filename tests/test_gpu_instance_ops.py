"""GPU tests of the in-kernel instance ops (RAY_TRACING_[SET_]INSTANCE_TRANSFORM / _USER_ID /
_VISIBILITY_MASK, include/luisa/ast/op.h:290-297) run as device batches: reads see the latest
writes, writes reach traversal at the accel's next build (as the CPU backend's dirty instances
reach Embree at the next update, cpu/accel.rs:335-343,560-579)."""
import numpy as np
import pytest

import luisacompute_amd as lc
from luisacompute_amd import host
from oracle import binding as ob
from parity import compare_any, compare_closest, random_rigid

pytestmark = pytest.mark.gpu


def _scene(device, stream, ninst, seed):
    rng = np.random.default_rng(seed)
    meshes = [lc.icosphere(2), lc.icosphere(3), lc.soup(2000, initseq=31, extent=2.0, sigma=0.2)]
    gm = [host.mesh_from_arrays(device, stream, v, t) for v, t in meshes]
    om = [ob.Mesh(v, t) for v, t in meshes]
    accel = device.create_accel()
    kinds, affines = [], []
    for i in range(ninst):
        k = int(rng.integers(len(meshes)))
        a = random_rigid(rng, -8, 8)
        accel.emplace_back(gm[k], a, visibility_mask=0xFF, user_id=100 + i)
        kinds.append(k)
        affines.append(a)
    stream << accel.build()
    return accel, kinds, np.array(affines, np.float32), (gm, om)


def _oracle(om, kinds, affines, masks):
    sc = ob.Scene()
    for k, a, m in zip(kinds, affines, masks):
        sc.add(om[k], a, int(m))
    return sc.commit()


def _read(device, stream, accel, field, idx):
    ib = host.upload(device, stream, np.asarray(idx, np.uint32))
    width = 12 if field == lc.INSTANCE_TRANSFORM else 1
    ob_ = device.create_buffer(4 * width, len(idx))
    accel.instance_read(stream, field, ib, ob_, len(idx))
    out = np.zeros((len(idx), width), np.float32 if width == 12 else np.uint32)
    stream << ob_.copy_to(out) << lc.synchronize()
    return out if width == 12 else out.ravel()


def _write(device, stream, accel, field, idx, values):
    ib = host.upload(device, stream, np.asarray(idx, np.uint32))
    vb = host.upload(device, stream, np.ascontiguousarray(values))
    accel.instance_write(stream, field, ib, vb, len(idx))
    stream.synchronize()


def test_instance_reads_match_the_instance_table(device, stream):
    accel, kinds, affines, keep = _scene(device, stream, 9, seed=5)
    idx = np.arange(9)
    np.testing.assert_array_equal(_read(device, stream, accel, lc.INSTANCE_TRANSFORM, idx), affines)
    np.testing.assert_array_equal(_read(device, stream, accel, lc.INSTANCE_USER_ID, idx), 100 + idx)
    np.testing.assert_array_equal(_read(device, stream, accel, lc.INSTANCE_VISIBILITY, idx), 0xFF)
    # repeated and permuted indices
    perm = np.array([3, 3, 0, 8, 1], np.uint32)
    np.testing.assert_array_equal(_read(device, stream, accel, lc.INSTANCE_TRANSFORM, perm), affines[perm])


def test_instance_writes_reach_traversal_at_the_next_build(device, stream):
    ninst = 17
    accel, kinds, affines, (gm, om) = _scene(device, stream, ninst, seed=11)
    rays = lc.rays_box_sphere(30000, initseq=77, lo=-10.0, hi=10.0)
    rf = lc.rays_as_floats(rays)
    masks = np.full(ninst, 0xFF, np.uint32)
    rng = np.random.default_rng(12)
    moved = np.array([1, 5, 9, 16], np.uint32)
    new_aff = np.array([random_rigid(rng, -8, 8) for _ in moved], np.float32)
    hidden = np.array([2, 7], np.uint32)
    _write(device, stream, accel, lc.INSTANCE_TRANSFORM, moved, new_aff)
    _write(device, stream, accel, lc.INSTANCE_VISIBILITY, hidden, np.zeros(len(hidden), np.uint32))
    _write(device, stream, accel, lc.INSTANCE_USER_ID, [4], np.array([7777], np.uint32))
    # reads see the writes at once
    np.testing.assert_array_equal(_read(device, stream, accel, lc.INSTANCE_TRANSFORM, moved), new_aff)
    np.testing.assert_array_equal(_read(device, stream, accel, lc.INSTANCE_VISIBILITY, hidden), 0)
    assert _read(device, stream, accel, lc.INSTANCE_USER_ID, [4])[0] == 7777
    # traversal still sees the built instances
    before = _oracle(om, kinds, affines, masks)
    g = host.trace_closest(device, stream, accel, rays)
    compare_closest(g, before.trace_closest2(rf), rays)
    # after the next build it sees the written ones
    stream << accel.build()
    aff2 = affines.copy()
    aff2[moved] = new_aff
    masks2 = masks.copy()
    masks2[hidden] = 0
    after = _oracle(om, kinds, aff2, masks2)
    g = host.trace_closest(device, stream, accel, rays)
    rep = compare_closest(g, after.trace_closest2(rf), rays)
    ga = host.trace_any(device, stream, accel, rays)
    compare_any(ga, after.trace_any(rf))
    assert not np.isin(g["inst"], hidden).any()
    print(f"[instance writes] {rep}", flush=True)
    a, vis, uid = accel.instance_info(4)
    assert uid == 7777 and vis == 0xFF
    a, vis, uid = accel.instance_info(int(moved[0]))
    np.testing.assert_array_equal(a, new_aff[0])
    # a modification of the same build wins over an earlier device write (applied after it)
    _write(device, stream, accel, lc.INSTANCE_USER_ID, [3], np.array([1], np.uint32))
    accel.set_visibility_on_update(3, 0x0F)
    stream << accel.build()
    a, vis, uid = accel.instance_info(3)
    assert uid == 1 and vis == 0x0F


def test_out_of_range_instance_write_fails_the_next_build(device, stream):
    accel, kinds, affines, keep = _scene(device, stream, 3, seed=2)
    _write(device, stream, accel, lc.INSTANCE_USER_ID, [1, 3], np.array([5, 6], np.uint32))
    with pytest.raises(lc.LCError):
        stream << accel.build()
    # the error is reported once; the valid write went through
    stream << accel.build()
    assert accel.instance_info(1)[2] == 5
