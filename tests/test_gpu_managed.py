"""GPU parity of the managed accel (the Python frontend's Accel over ManagedAccel,
src/py/luisa/accel.py:58-160, src/py/managed_accel.cpp): shared BLASes trace like separate ones."""
import numpy as np
import pytest

import luisacompute_amd as lc
from luisacompute_amd import host
from luisacompute_amd.managed import ManagedAccel
from oracle import binding as ob
from parity import compare_closest, compare_query, random_boxes, random_rigid

pytestmark = pytest.mark.gpu


def test_managed_accel_shares_blases_and_traces_like_the_oracle(device, stream):
    rng = np.random.default_rng(50)
    (v1, t1), (v2, t2) = lc.icosphere(3), lc.soup(3000, initseq=17, extent=2.0, sigma=0.2)
    bufs = [(host.upload(device, stream, v1), host.upload(device, stream, t1)),
            (host.upload(device, stream, v2), host.upload(device, stream, t2))]
    om = [ob.Mesh(v1, t1), ob.Mesh(v2, t2)]
    boxes = random_boxes(rng, 150, -2.0, 2.0, 0.1, 0.5)
    ab = host.upload(device, stream, boxes)
    m = ManagedAccel(device)
    sc = ob.Scene()
    affs = []
    for i in range(8):
        a = random_rigid(rng, -4, 4, 0.5, 1.5)
        affs.append(a)
        k = i % 2
        m.add(bufs[k][0], bufs[k][1], a)
        sc.add(om[k], a)
    a = random_rigid(rng, -4, 4, 0.5, 1.5)
    m.add_procedural(ab, transform=a)
    sc.add(ob.Procedural(boxes), a)
    m.update(stream, sync=True)
    sc.commit()
    assert m.blas_count == 3 and len(m) == 9
    rays = lc.rays_as_floats(lc.rays_box_sphere(30000, initseq=51, lo=-5.0, hi=5.0))
    rep = compare_closest(m.trace_closest(stream, rays), sc.trace_closest2(rays), rays)
    pol = dict(surface=lc.QUERY_SURFACE_COMMIT, procedural=lc.QUERY_PROCEDURAL_SPHERE)
    q = compare_query(m.query_all(stream, rays, **pol), sc.query(rays, **pol), sc, rays, **pol)
    print(f"[managed] {rep} {q}", flush=True)
    # re-describe instance 3 as the icosphere and drop the procedural instance
    m.set(3, bufs[0][0], bufs[0][1], affs[3])
    m.pop()
    m.update(stream, sync=True)
    assert m.blas_count == 2 and len(m) == 8
    sc2 = ob.Scene()
    for i in range(8):
        sc2.add(om[0 if i == 3 else i % 2], affs[i])
    sc2.commit()
    compare_closest(m.trace_closest(stream, rays), sc2.trace_closest2(rays), rays)
    np.testing.assert_array_equal(m.instance_transform(5), affs[5])
