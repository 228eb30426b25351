"""Pins the oracle's ray queries (cpu/accel.rs:582-800 with the candidate callbacks of
src/tests/test_procedural.cpp:86-120): against the oracle's pinned BVH trace where the callbacks
reduce to it, against a float64 closed form of ray/sphere intersection for the procedural
callback, and against the cutout condition evaluated in float64 for the surface callback."""
import numpy as np
import pytest

from oracle import binding as ob
from parity import random_boxes, random_rigid

MISS = 0xFFFFFFFF
COMMIT, IGNORE, CUTOUT = 0, 1, 2
P_IGNORE, P_SPHERE = 0, 1


def _mesh_scene(rng, opaque, n_inst=3):
    meshes = [ob.icosphere(2), ob.soup(400, initseq=5, extent=2.0, sigma=0.3)]
    om = [ob.Mesh(v, t) for v, t in meshes]
    sc = ob.Scene()
    for i in range(n_inst):
        sc.add(om[i % 2], random_rigid(rng, -3, 3), 0xFF, opaque=opaque)
    return sc.commit(), om


def _rays(n, seed, lo=-3.0, hi=3.0):
    return ob.rays_box_sphere(n, initstate=seed, initseq=3, lo=lo, hi=hi)


@pytest.mark.parametrize("opaque", [True, False])
def test_commit_all_equals_trace_closest(opaque):
    """Every surface candidate committed, procedural ignored: the nearest commit is trace_closest."""
    rng = np.random.default_rng(1)
    sc, keep = _mesh_scene(rng, opaque)
    rays = _rays(3000, 7)
    q = sc.query(rays, surface=COMMIT, procedural=P_IGNORE)
    h = sc.trace_closest(rays)
    for f in ("inst", "prim", "u", "v", "t"):
        np.testing.assert_array_equal(q[f], h[f])
    hit = h["prim"] != MISS
    assert hit.mean() > 0.2
    np.testing.assert_array_equal(q["hit_type"], np.where(hit, 1, 0))


def test_opaque_instances_skip_the_surface_callback():
    rng = np.random.default_rng(2)
    sc, keep = _mesh_scene(rng, opaque=True)
    rays = _rays(2000, 8)
    h = sc.trace_closest(rays)
    for surface in (IGNORE, CUTOUT):
        q = sc.query(rays, surface=surface, procedural=P_IGNORE)
        np.testing.assert_array_equal(q["prim"], h["prim"])
        np.testing.assert_array_equal(q["t"], h["t"])


def test_ignore_all_misses_with_tmax():
    rng = np.random.default_rng(3)
    sc, keep = _mesh_scene(rng, opaque=False)
    rays = _rays(1000, 9)
    q = sc.query(rays, surface=IGNORE, procedural=P_IGNORE)
    assert (q["hit_type"] == 0).all() and (q["inst"] == MISS).all() and (q["prim"] == MISS).all()
    np.testing.assert_array_equal(q["t"], rays[:, 7])
    assert (q["u"] == 0).all() and (q["v"] == 0).all()


def _cutout64(u, v, r):
    u, v = u.astype(np.float64), v.astype(np.float64)
    w = 1.0 - u - v
    m = np.minimum(np.minimum(r - np.hypot(w, u), r - np.hypot(u, v)), r - np.hypot(v, w))
    return m  # > 0: inside the cutout (committed), < 0: outside


def test_cutout_callback_against_float64():
    """Non-opaque triangles commit iff the three pairwise lengths of (1-u-v, u, v) are below the
    cutout: every committed hit satisfies it; the nearest hit is kept iff it satisfies it, and
    otherwise the query answer lies beyond it."""
    rng = np.random.default_rng(4)
    sc, keep = _mesh_scene(rng, opaque=False)
    rays = _rays(6000, 10)
    r = 0.8
    q = sc.query(rays, surface=CUTOUT, cutout=r, procedural=P_IGNORE)
    h = sc.trace_closest(rays)
    qs = q["hit_type"] == 1
    assert (_cutout64(q["u"][qs], q["v"][qs], r) > -1e-6).all()
    hit = h["prim"] != MISS
    m = _cutout64(h["u"], h["v"], r)
    keep_ = hit & (m > 1e-5)
    drop = hit & (m < -1e-5)
    assert keep_.sum() > 100 and drop.sum() > 100
    np.testing.assert_array_equal(q["prim"][keep_], h["prim"][keep_])
    np.testing.assert_array_equal(q["t"][keep_], h["t"][keep_])
    later = (q["hit_type"][drop] == 0) | (q["t"][drop] > h["t"][drop])
    assert later.all()


def _sphere_closed_form(rays, boxes, affines):
    """Nearest entry distance of the spheres inscribed in the boxes (object space), mapped to world
    space by rigid + uniform-scale affines, in float64; with the callback's rules (centre ahead,
    entry distance in [tmin, tmax)). Returns (inst, prim, t, margin) per ray."""
    n = rays.shape[0]
    best_t = np.full(n, np.inf)
    best_i = np.full(n, MISS, np.int64)
    best_p = np.full(n, MISS, np.int64)
    margin = np.full(n, np.inf)  # distance of the decision from a boundary case (grazing)
    o = rays[:, :3].astype(np.float64)
    d = rays[:, 4:7].astype(np.float64)
    tmin, tmax = rays[:, 3].astype(np.float64), rays[:, 7].astype(np.float64)
    for i, (bx, a) in enumerate(zip(boxes, affines)):
        A = a.reshape(3, 4).astype(np.float64)
        s = np.cbrt(np.linalg.det(A[:, :3]))
        c_obj = 0.5 * (bx[:, :3] + bx[:, 3:]).astype(np.float64)
        r_obj = 0.5 * np.min((bx[:, 3:] - bx[:, :3]).astype(np.float64), axis=1)
        C = c_obj @ A[:, :3].T + A[:, 3]
        R = s * r_obj
        for p in range(len(bx)):
            L = C[p] - o
            dd = np.einsum("ij,ij->i", d, d)
            tc = np.einsum("ij,ij->i", L, d) / dd
            q = np.einsum("ij,ij->i", L, L) - tc * tc * dd
            disc = R[p] ** 2 - q
            ok = (tc > 0) & (disc >= 0)
            t = tc - np.sqrt(np.maximum(disc, 0) / dd)
            ok &= (t >= tmin) & (t < tmax)
            margin = np.minimum(margin, np.abs(disc) / R[p] ** 2)
            better = ok & (t < best_t)
            best_t = np.where(better, t, best_t)
            best_i = np.where(better, i, best_i)
            best_p = np.where(better, p, best_p)
    return best_i, best_p, best_t, margin


def test_sphere_callback_against_closed_form():
    """Procedural instances under rigid + uniform-scale transforms: the nearest committed sphere
    entry equals the float64 closed form (object-space spheres map to world-space spheres)."""
    rng = np.random.default_rng(5)
    boxes, affines, prims = [], [], []
    sc = ob.Scene()
    for i in range(3):
        bx = random_boxes(rng, 40, -2.0, 2.0, 0.2, 0.8)
        a = random_rigid(rng, -2, 2, 0.5, 1.5)
        pr = ob.Procedural(bx)
        sc.add(pr, a)
        boxes.append(bx)
        affines.append(a)
        prims.append(pr)
    sc.commit()
    rays = _rays(4000, 11, -5.0, 5.0)
    rays[:, 3] = 0.0
    q = sc.query(rays, surface=COMMIT, procedural=P_SPHERE)
    gi, gp, gt, margin = _sphere_closed_form(rays, boxes, affines)
    clear = margin > 1e-3
    hit = gp != MISS
    assert (hit & clear).sum() > 300
    sel = clear
    np.testing.assert_array_equal(q["prim"][sel & hit], gp[sel & hit])
    np.testing.assert_array_equal(q["inst"][sel & hit], gi[sel & hit])
    assert (q["hit_type"][sel & hit] == 2).all() and (q["hit_type"][sel & ~hit] == 0).all()
    np.testing.assert_allclose(q["t"][sel & hit], gt[sel & hit], rtol=1e-4, atol=1e-5)
    # procedural hits report bary (0, 0)
    assert (q["u"][q["hit_type"] == 2] == 0).all() and (q["v"][q["hit_type"] == 2] == 0).all()
    # plain traces do not see procedural instances
    assert (sc.trace_closest(rays)["prim"] == MISS).all()
    assert not sc.trace_any(rays).any()


def test_query_candidate_agrees_and_masks_apply():
    rng = np.random.default_rng(6)
    v, t = ob.icosphere(2)
    m = ob.Mesh(v, t)
    pr = ob.Procedural(random_boxes(rng, 30, -2.0, 2.0, 0.3, 0.9))
    sc = ob.Scene()
    sc.add(m, random_rigid(rng, -2, 2), 0x01, opaque=False)
    sc.add(pr, random_rigid(rng, -2, 2), 0x02)
    sc.commit()
    rays = _rays(2000, 12, -5.0, 5.0)
    pol = dict(surface=CUTOUT, cutout=0.8, procedural=P_SPHERE)
    q = sc.query(rays, **pol)
    assert set(np.unique(q["hit_type"])) == {0, 1, 2}
    for i in np.flatnonzero(q["hit_type"] != 0)[:200]:
        c = sc.query_candidate(rays[i], q["inst"][i], q["prim"][i], q["hit_type"][i], **pol)
        assert c is not None and np.float32(c[0]) == q["t"][i]
    # ray masks select instances
    q1 = sc.query(rays, mask=0x01, **pol)
    q2 = sc.query(rays, mask=0x02, **pol)
    assert set(np.unique(q1["inst"])) <= {0, MISS} and set(np.unique(q2["inst"])) <= {1, MISS}
    # the full query's answer is the nearer of the two single-mask answers (a miss reports tmax)
    hit = q["hit_type"] != 0
    nearest = np.where(q1["t"] <= q2["t"], q1["inst"], q2["inst"])
    np.testing.assert_array_equal(q["inst"][hit], nearest[hit])
    # a candidate of a masked-out instance is never committed
    i = int(np.flatnonzero(q2["hit_type"] == 2)[0])
    assert sc.query_candidate(rays[i], q2["inst"][i], q2["prim"][i], 2, mask=0x01, **pol) is None
