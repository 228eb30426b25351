"""Pins the oracle's path tracer (the checker of the HIP path tracer) with what the mathematics
fixes: the sampled sine/cosine against libm, radiance seen directly on the light, the direct
illumination of a floor under a quad light against the form factor (quadrature and closed form),
Russian roulette without bias, and the per-sample RNG stream layout."""
import math

import numpy as np
import pytest

from luisacompute_amd import workloads
from oracle import binding as ob


def room_scene(room):
    meshes = [ob.Mesh(v, t) for v, t in room.meshes]
    sc = ob.Scene()
    for m in meshes:
        sc.add(m)
    sc.commit()
    sc._meshes = meshes
    return sc


def light_of(room):
    return (room.LIGHT_POSITION, room.LIGHT_U, room.LIGHT_V, room.emission)


def test_sincos_accuracy():
    u = np.concatenate([np.arange(0, 1 << 24, 97, dtype=np.int64), [0, 1 << 22, 1 << 23, 3 << 22, (1 << 24) - 1]])
    u = (u.astype(np.float64) * 2.0 ** -24).astype(np.float32)
    worst = 0.0
    for x in u[:: max(1, len(u) // 20000)]:
        s, c = ob.sincos_2pi(x)
        a = 2.0 * math.pi * float(x)
        worst = max(worst, abs(s - math.sin(a)), abs(c - math.cos(a)))
    assert worst < 3e-7, worst
    # quadrant boundaries are exact
    assert ob.sincos_2pi(0.0) == (0.0, 1.0)
    assert ob.sincos_2pi(0.25) == (1.0, -0.0)
    assert ob.sincos_2pi(0.5) == (-0.0, -1.0)
    assert ob.sincos_2pi(0.75) == (-1.0, 0.0)


def test_light_seen_directly_is_its_radiance():
    """Camera rays that land on the light return its radiance exactly (depth-0 emission)."""
    room = workloads.BoxRoom(wall_quads=8, block_quads=2)
    sc = room_scene(room)
    w, h, spp = 96, 64, 4
    img, _, _ = ob.pathtrace(sc, room.albedo, room.emissive, light_of(room), room.camera, w, h, spp)
    # where does the film corner (x + jx, y + jy) land on the light plane y = 1.98?
    frame = min(w, h)
    inside = np.ones((h, w), bool)
    p, u, v = (np.array(a) for a in (room.LIGHT_POSITION, room.LIGHT_U, room.LIGHT_V))
    for jx in (0.0, 1.0):
        for jy in (0.0, 1.0):
            X, Y = np.meshgrid(np.arange(w) + jx, np.arange(h) + jy)
            fx, fy = (2 * X - w) / frame, (2 * Y - h) / frame
            with np.errstate(divide="ignore", invalid="ignore"):
                lam = np.where(fy < 0, -(1.98 - 1.0) / fy, np.inf)
                x, z = lam * fx, 0.98 - lam
            m = 1e-3
            inside &= (x > p[0] + m) & (x < p[0] + u[0] - m) & (z > p[2] + m) & (z < p[2] + v[2] - m)
    assert inside.sum() >= 20
    np.testing.assert_array_equal(img[inside][:, :3], np.float32(15.0))


def form_factor_quadrature(px, pz, height, x0, x1, z0, z1, n=400):
    """Point (px, 0, pz) facing +y, rectangle [x0,x1] x [z0,z1] at y = height facing -y:
    F = 1/pi * integral of cos cos' / r^2 dA = 1/pi * integral of h^2 / r^4 dA (midpoint rule)."""
    xs = x0 + (np.arange(n) + 0.5) * (x1 - x0) / n
    zs = z0 + (np.arange(n) + 0.5) * (z1 - z0) / n
    X, Z = np.meshgrid(xs, zs)
    r2 = (X - px) ** 2 + height ** 2 + (Z - pz) ** 2
    return float(np.sum(height ** 2 / r2 ** 2) * (x1 - x0) * (z1 - z0) / (n * n) / math.pi)


def form_factor_closed(px, pz, height, x0, x1, z0, z1):
    """Differential area to a parallel rectangle with one corner above it (Howell's catalogue,
    B-3), extended to any position by signed superposition of four corner rectangles."""
    def f(a, b):
        A, B = a / height, b / height
        return (A / math.sqrt(1 + A * A) * math.atan(B / math.sqrt(1 + A * A)) +
                B / math.sqrt(1 + B * B) * math.atan(A / math.sqrt(1 + B * B))) / (2 * math.pi)
    return f(x1 - px, z1 - pz) - f(x0 - px, z1 - pz) - f(x1 - px, z0 - pz) + f(x0 - px, z0 - pz)


def test_form_factor_formulas_agree():
    for px, pz in ((0.0, 0.0), (0.3, -0.7), (-1.5, 2.0)):
        q = form_factor_quadrature(px, pz, 1.98, -0.24, 0.23, -0.22, 0.16)
        c = form_factor_closed(px, pz, 1.98, -0.24, 0.23, -0.22, 0.16)
        assert abs(q - c) < 1e-5 * c, (q, c)


def test_direct_lighting_matches_form_factor():
    """A floor under the quad light, camera between them looking down, two bounces: the
    estimate (light sample + BSDF sample, balance heuristic) is the direct illumination
    albedo * Le * F(x) without bias."""
    floor_v = np.array([[-20, 0, -20, 0], [-20, 0, 20, 0], [20, 0, -20, 0], [20, 0, 20, 0]], np.float32)
    floor_t = np.array([[0, 1, 2], [2, 1, 3]], np.uint32)   # faces +y
    p, u, v = (np.array(a, np.float64) for a in (workloads.BoxRoom.LIGHT_POSITION, workloads.BoxRoom.LIGHT_U,
                                                  workloads.BoxRoom.LIGHT_V))
    lv = np.zeros((4, 4), np.float32)
    lv[:, :3] = [p, p + u, p + v, p + u + v]
    light_t = np.array([[0, 1, 2], [1, 3, 2]], np.uint32)
    m0, m1 = ob.Mesh(floor_v, floor_t), ob.Mesh(lv, light_t)
    sc = ob.Scene().add(m0).add(m1).commit()
    t = 0.8
    cam = dict(origin=(0.0, 1.0, 0.0), right=(t, 0.0, 0.0), up=(0.0, 0.0, -t), forward=(0.0, -1.0, 0.0))
    w = h = 16
    img, _, _ = ob.pathtrace(sc, [[0.5, 0.5, 0.5], [0, 0, 0]], [0, 1],
                             (workloads.BoxRoom.LIGHT_POSITION, workloads.BoxRoom.LIGHT_U,
                              workloads.BoxRoom.LIGHT_V, (1.0, 1.0, 1.0)),
                             cam, w, h, 1024, max_depth=2, rr_depth=8)
    # pixel (x, y) sees the floor point (fx t, 0, fy t); average F over the pixel footprint
    sub = (np.arange(4) + 0.5) / 4
    expect = np.zeros((h, w))
    x0, x1, z0, z1 = p[0], p[0] + u[0], p[2], p[2] + v[2]
    for y in range(h):
        for x in range(w):
            acc = 0.0
            for sy in sub:
                for sx in sub:
                    fx, fy = (2 * (x + sx) - w) / w, (2 * (y + sy) - h) / h
                    acc += form_factor_closed(fx * t, fy * t, p[1], x0, x1, z0, z1)
            expect[y, x] = 0.5 * acc / 16
    got = img[..., 0].astype(np.float64)
    assert np.all(img[..., 0] == img[..., 1]) and np.all(img[..., 1] == img[..., 2])
    rel_mean = abs(got.mean() - expect.mean()) / expect.mean()
    rel_pix = np.abs(got - expect) / expect
    assert rel_mean < 0.01, rel_mean
    assert np.percentile(rel_pix, 95) < 0.08, np.percentile(rel_pix, 95)


def test_russian_roulette_is_unbiased():
    room = workloads.BoxRoom(wall_quads=8, block_quads=2)
    sc = room_scene(room)
    args = (sc, room.albedo, room.emissive, light_of(room), room.camera, 32, 18, 256)
    a, _, sa = ob.pathtrace(*args, max_depth=8, rr_depth=3)
    b, _, sb = ob.pathtrace(*args, max_depth=8, rr_depth=8, sequence=7 << 32)
    ma, mb = a[..., :3].mean(), b[..., :3].mean()
    assert abs(ma - mb) / mb < 0.02, (ma, mb)
    assert sa < sb  # roulette ends paths early


def test_sample_offset_selects_the_stream_window():
    """Sample s of a pixel starts 64 s draws into the pixel's stream: rendering samples 0 and 1
    together equals the mean of rendering each alone."""
    room = workloads.BoxRoom(wall_quads=8, block_quads=2)
    sc = room_scene(room)
    args = (sc, room.albedo, room.emissive, light_of(room), room.camera, 24, 16)
    both, _, _ = ob.pathtrace(*args, 2)
    s0, _, _ = ob.pathtrace(*args, 1, sample_offset=0)
    s1, _, _ = ob.pathtrace(*args, 1, sample_offset=1)
    np.testing.assert_array_equal(both[..., :3], ((s0 + s1) / np.float32(2))[..., :3])
    assert not np.array_equal(s0, s1)


def test_box_room_geometry():
    room = workloads.BoxRoom(wall_quads=4, block_quads=2)
    assert room.triangle_count == 6 * 4 * 4 * 2 + 2 * 6 * 2 * 2 * 2 + 2
    # the full room: 6 walls x 512^2 quads x 2 = 3,145,728 triangles, blocks 2 x 6 x 128^2 x 2
    assert workloads.BoxRoom().triangle_count == 3145728 + 393216 + 2
    # walls face inward, blocks outward, the light down
    centre = np.array([0.0, 1.0, 0.0])
    for k, (v, t) in enumerate(room.meshes):
        a, b, c = (v[t[:, i], :3].astype(np.float64) for i in range(3))
        n = np.cross(b - a, c - a)
        mid = (a + b + c) / 3
        if k < 6:
            assert np.all(np.einsum("ij,ij->i", n, centre - mid) > 0)
        elif k < 8:
            box_c = (v[:, :3].min(0) + v[:, :3].max(0)) / 2
            assert np.all(np.einsum("ij,ij->i", n, mid - box_c) > 0)
        else:
            assert np.all(n[:, 1] < 0)
    assert np.all((room.albedo[:8] >= 0.1) & (room.albedo[:8] < 0.9))


def test_offset_ray_origin_closed_forms():
    """offset_ray_origin (src/dsl/rtx/ray.cpp:16-24; Waechter and Binder, "A Fast and Robust Method
    for Avoiding Self-Intersection", Ray Tracing Gems 2019) pinned on values fixed by its three
    constants: components with |p| < 1/32 move by n / 65536 in floating point; the others move by
    trunc(256 n) units in the last place, away from zero where n points that way."""
    f = np.float32

    def off(p, n):
        return ob.offset_ray_origin(np.array([p], np.float32), np.array([n], np.float32))[0]
    # near the origin: p + n * 2^-16
    got = off((0.01, 0.0, -0.02), (1.0, 0.0, 0.0))
    assert got[0] == f(0.01) + f(2.0 ** -16) and got[1] == 0.0 and got[2] == f(-0.02)
    # 0.5 has an ulp of 2^-24: 256 ulps = 2^-16; -3.0 (ulp 2^-22) moved further negative by 2^-14
    got = off((-3.0, 1.0, 0.5), (-1.0, 0.0, 1.0))
    assert got[0] == f(-3.0 - 2.0 ** -14) and got[1] == 1.0 and got[2] == f(0.5 + 2.0 ** -16)
    # |p| = 1/32 exactly is outside the float branch: 256 ulps of 2^-5 (ulp 2^-28) = 2^-20;
    # just below it the float branch adds 2^-16
    got = off((1.0 / 32.0, 0.0, 0.0), (1.0, 0.0, 0.0))
    assert got[0] == f(1.0 / 32.0 + 2.0 ** -20)
    below = np.nextafter(f(1.0 / 32.0), f(0.0))
    got = off((below, 0.0, 0.0), (1.0, 0.0, 0.0))
    assert got[0] == below + f(2.0 ** -16)
    # the integer offset truncates 256 n: n = 0.5 -> 128 ulps, n = 0.999 -> 255 ulps, n = -0.999
    # on a positive coordinate -> 255 steps of the bit pattern toward zero (below 4.0 the ulp is
    # 2^-22)
    got = off((1.0, 2.0, 4.0), (0.5, 0.999, -0.999))
    assert got[0] == f(1.0 + 128 * 2.0 ** -23)
    assert got[1] == f(2.0 + 255 * 2.0 ** -22)
    assert got[2] == f(4.0 - 255 * 2.0 ** -22)
