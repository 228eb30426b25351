"""Kernels written in the Python DSL (luisacompute_amd.dsl) for the DSL tests."""
from luisacompute_amd import dsl
from luisacompute_amd.dsl import Accel, Buffer, float_, uint
from luisacompute_amd.programs import INV_PI, path_tracer  # noqa: F401  (the benchmark's program)


@dsl.kernel
def trace_kernel(rays: Buffer("Ray"), hits: Buffer("SurfaceHit"), occluded: Buffer("uint"), accel: Accel,
                 mask: uint):
    i = dispatch_id().x
    r = rays.read(i)
    hits.write(i, accel.trace_closest(r, mask))
    occluded.write(i, 1 if accel.trace_any(r, mask) else 0)


@dsl.kernel(block=(64, 1, 1))
def control_kernel(xs: Buffer("float"), out: Buffer("float"), n_terms: uint, scale: float_):
    """Loops, branches and math: out[i] = scale * sum_{k < n_terms, k != 3} f(x, k)."""
    i = dispatch_id().x
    x = xs.read(i)
    acc = 0.0
    for k in range(n_terms):
        if k == 3:
            continue
        term = x * k if x > 0.5 else x - k
        if k > 6 and term < -5.0:
            break
        acc += term
    j = 0
    while j < 2:
        acc = acc * 0.5 + abs(x)
        j += 1
    out.write(i, scale * acc)


@dsl.kernel
def shade_kernel(rays: Buffer("Ray"), points: Buffer("float4"), accel: Accel):
    """Hit point from the ray and t, an offset origin and a shadow ray toward +z."""
    i = dispatch_id().x
    r = rays.read(i)
    h = accel.trace_closest(r, 255)
    if hit_miss(h):
        points.write(i, float4(0.0, 0.0, 0.0, -1.0))
        return
    p = ray_origin(r) + ray_direction(r) * h.committed_ray_t
    q = offset_ray_origin(p, normalize(ray_direction(r)) * -1.0)
    blocked = accel.trace_any(make_ray(q, float3(0.0, 0.0, 1.0), 0.0, 100.0), 255)
    points.write(i, float4(q.x, q.y, q.z, 1.0 if blocked else 0.0))


@dsl.kernel
def query_kernel(rays: Buffer("Ray"), out: Buffer("CommittedHit"), accel: Accel, cutout: float_, any_hit: uint):
    """The callbacks of src/tests/test_procedural.cpp:86-120 written in the DSL."""
    i = dispatch_id().x

    def on_surface(c):
        w = (1.0 - c.bary.x) - c.bary.y
        u = c.bary.x
        v = c.bary.y
        return sqrt(w * w + u * u) < cutout and sqrt(u * u + v * v) < cutout and sqrt(v * v + w * w) < cutout

    def on_procedural(c):
        lo = c.aabb_min
        hi = c.aabb_max
        r = 0.5 * min(min(hi.x - lo.x, hi.y - lo.y), hi.z - lo.z)
        center = float3((lo.x + hi.x) * 0.5, (lo.y + hi.y) * 0.5, (lo.z + hi.z) * 0.5)
        L = center - c.origin
        d = c.direction
        dd = dot(d, d)
        tc = dot(L, d) / dd
        if not (tc > 0.0):
            return -1.0
        q = dot(L, L) - (tc * tc) * dd
        rr = r * r
        if not (q <= rr):
            return -1.0
        return tc - sqrt((rr - q) / dd)

    if any_hit != 0:
        out.write(i, accel.query_any(rays.read(i), 255, on_surface, on_procedural))
    else:
        out.write(i, accel.query_all(rays.read(i), 255, on_surface, on_procedural))


@dsl.kernel(block=(64, 1, 1))
def sample_kernel(uv: Buffer("float2"), out: Buffer("float4"), heap: dsl.BindlessArray, n_slots: uint):
    """Samples texture slots 0..n_slots-1 of the heap at uv[i] (out[i * n_slots + k])."""
    i = dispatch_id().x
    for k in range(n_slots):
        out.write(i * n_slots + k, heap.tex2d_sample(k, uv.read(i)))


@dsl.kernel(block=(8, 4, 2))
def volume_kernel(src: dsl.Volume, dst: dsl.Volume, scale: float_):
    """dst(x, y, z) = (scale * src.xy, z, 1) over a 3-D dispatch."""
    p = dispatch_id()
    v = src.read(p.x, p.y, p.z)
    dst.write(p.x, p.y, p.z, float4(v.x * scale, v.y * scale, float(p.z), 1.0))


@dsl.kernel(block=(64, 1, 1))
def division_kernel(out_true: Buffer("float"), out_floor: Buffer("uint"), k: uint):
    """Python division: '/' of two integers is float division (src/py/luisa/builtin.py:191-194),
    '//' the integer division of the operand type."""
    i = dispatch_id().x
    out_true.write(i, i / k)
    out_floor.write(i, i // k)
