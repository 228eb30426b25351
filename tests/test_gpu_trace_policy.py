"""GPU tests of the traversal step policy (triangle-step threshold, early refill, refill interval)
and of the counting run's lane accounting: every policy gives the same hits bit for bit, and the
accounting is consistent with itself and with the plain counting run."""
import numpy as np
import pytest

import luisacompute_amd as lc
from luisacompute_amd import host

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def soup(device, stream):
    verts, tris = lc.soup(200000, initseq=61, extent=20.0, sigma=0.3)
    rays = lc.rays_box_sphere(1 << 17, initseq=62, lo=0.0, hi=20.0)
    accel = host.single_mesh_accel(device, stream, verts, tris)
    rb = host.upload(device, stream, rays)
    return accel, rays, rb


def test_step_policies_give_the_same_hits(device, stream, soup):
    accel, rays, rb = soup
    saved_c, saved_a = device.trace_config("closest"), device.trace_config("any")
    saved_sc, saved_sa = device.trace_steps("closest"), device.trace_steps("any")
    try:
        want = host.trace_closest(device, stream, accel, rays)
        want_any = host.trace_any(device, stream, accel, rays)
        assert (want["prim"] != 0xFFFFFFFF).sum() > 1000
        for tq, refill, steps in ((1, 0, 16), (1, 32, 32), (2, 48, 8), (3, 0, 4), (4, 63, 1)):
            print(f"[policy] tri_quarters={tq} refill_below={refill} steps={steps}", flush=True)
            for kind in ("closest", "any"):
                c = device.trace_config(kind)
                device.set_trace_config(kind, c["block"], c["lds_stack"], steps, c["blocks_per_cu"])
                device.set_trace_steps(kind, tq, refill)
            got = host.trace_closest(device, stream, accel, rays)
            np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
            np.testing.assert_array_equal(host.trace_any(device, stream, accel, rays), want_any)
        with pytest.raises(lc.LCError):
            device.set_trace_steps("closest", 0, 0)
        with pytest.raises(lc.LCError):
            device.set_trace_steps("closest", 2, 65)
    finally:
        for kind, c, st in (("closest", saved_c, saved_sc), ("any", saved_a, saved_sa)):
            device.set_trace_config(kind, c["block"], c["lds_stack"], c["steps"], c["blocks_per_cu"])
            device.set_trace_steps(kind, *st)


def test_lane_accounting(device, stream, soup):
    """Node visits and triangle tests equal the plain counting run's; each is at most one per live
    lane-iteration; live lane-iterations fit in the waves' lane slots; refills happen."""
    accel, rays, rb = soup
    n = len(rays)
    nodes, tris = accel.trace_count(stream, rb, n)
    st = accel.trace_steps(stream, rb, n)
    print(f"[lanes] nodes {st.nodes_visited} tris {st.triangles_tested} live {st.live_lane_steps} "
          f"waves {st.wave_steps} refills {st.refill_rounds}", flush=True)
    assert (st.nodes_visited, st.triangles_tested) == (nodes, tris)
    assert nodes > n and tris > 0
    assert st.nodes_visited + st.triangles_tested <= st.live_lane_steps <= 64 * st.wave_steps
    assert st.refill_rounds > 0 and st.wave_steps >= st.refill_rounds
    sa = accel.trace_steps(stream, rb, n, any_hit=True)
    assert sa.nodes_visited + sa.triangles_tested <= sa.live_lane_steps <= 64 * sa.wave_steps
