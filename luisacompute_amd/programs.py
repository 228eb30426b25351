"""Programs the benchmark runs through the run-time shader path, shared with the tests.

  path_tracer : the path tracer of src/tests/test_path_tracing.cpp:165-258 written in the Python DSL
                (the AST -> HIP codegen path; one thread per pixel, the megakernel form)
  QUERY       : a ray-query shader in HIP source with the candidate callbacks of
                src/tests/test_procedural.cpp:86-120, as a codegen would inline them
"""
from . import dsl
from .dsl import Accel, Buffer, uint

INV_PI = 0.318309886183790672


@dsl.kernel(block=(8, 8, 1))
def path_tracer(image: Buffer("float4"), accel: Accel, heap: dsl.BindlessArray, xforms: Buffer("float"),
                materials: Buffer("float4"), light_position: dsl.float3_, light_u: dsl.float3_, light_v: dsl.float3_,
                light_emission: dsl.float3_, camera_origin: dsl.float3_, camera_right: dsl.float3_,
                camera_up: dsl.float3_, camera_forward: dsl.float3_, spp: uint, sample_offset: uint, max_depth: uint,
                rr_depth: uint, seed: dsl.ulong, sequence: dsl.ulong):
    """The path tracer of src/tests/test_path_tracing.cpp:165-258 in the DSL, with the fixed
    choices of the library's path tracer (PCG32 streams per pixel, 64 draws per sample, polynomial
    sine / cosine, luminance Russian roulette): heap slots 2i / 2i + 1 hold instance i's vertex
    (float4) and triangle buffers, xforms 12 floats (object -> world) and materials (albedo,
    emissive) per instance."""
    size = dispatch_size()
    xy = dispatch_id()
    x = xy.x
    y = xy.y
    pixel = y * size.x + x
    state = ulong(0)
    inc = ulong(0)

    def pcg_next():
        old = state
        state = old * 6364136223846793005 + inc
        xs = uint(((old >> 18) ^ old) >> 27)
        rot = uint(old >> 59)
        return (xs >> rot) | (xs << ((0 - rot) & 31))

    def pcg_float():
        return float(pcg_next() >> 8) * 5.9604644775390625e-08

    def to_world(i, v):
        m = 12 * i
        return float3(xforms.read(m) * v.x + xforms.read(m + 1) * v.y + xforms.read(m + 2) * v.z + xforms.read(m + 3),
                      xforms.read(m + 4) * v.x + xforms.read(m + 5) * v.y + xforms.read(m + 6) * v.z + xforms.read(m + 7),
                      xforms.read(m + 8) * v.x + xforms.read(m + 9) * v.y + xforms.read(m + 10) * v.z + xforms.read(m + 11))

    lc_ = cross(light_u, light_v)
    light_area = length(lc_)
    light_n = normalize(lc_)
    frame = float(min(size.x, size.y))
    total = float3(0.0, 0.0, 0.0)
    for s in range(spp):
        state = ulong(0)
        inc = ((sequence + pixel) << 1) | 1
        pcg_next()
        state += seed
        pcg_next()
        delta = ulong(64) * (sample_offset + s)
        cur_mult = ulong(6364136223846793005)
        cur_plus = inc
        acc_mult = ulong(1)
        acc_plus = ulong(0)
        while delta > 0:
            if (delta & 1) != 0:
                acc_mult *= cur_mult
                acc_plus = acc_plus * cur_mult + cur_plus
            cur_plus = (cur_mult + 1) * cur_plus
            cur_mult *= cur_mult
            delta >>= 1
        state = acc_mult * state + acc_plus
        rx = pcg_float()
        ry = pcg_float()
        fx = ((float(x) + rx) * 2.0 - float(size.x)) / frame
        fy = ((float(y) + ry) * 2.0 - float(size.y)) / frame
        d = normalize(float3(fx * camera_right.x - fy * camera_up.x + camera_forward.x,
                             fx * camera_right.y - fy * camera_up.y + camera_forward.y,
                             fx * camera_right.z - fy * camera_up.z + camera_forward.z))
        o = camera_origin
        beta = float3(1.0, 1.0, 1.0)
        L = float3(0.0, 0.0, 0.0)
        pdf = 0.0
        for depth in range(max_depth):
            h = accel.trace_closest(make_ray(o, d, 0.0, FLT_MAX), 255)
            if hit_miss(h):
                break
            inst = h.inst
            tri = heap.buffer_read("uint3", 2 * inst + 1, h.prim)
            p0 = to_world(inst, heap.buffer_read("float4", 2 * inst, tri.x))
            p1 = to_world(inst, heap.buffer_read("float4", 2 * inst, tri.y))
            p2 = to_world(inst, heap.buffer_read("float4", 2 * inst, tri.z))
            p = triangle_interpolate(hit_bary(h), p0, p1, p2)
            n = normalize(cross(p1 - p0, p2 - p0))
            cos_wo = -d.x * n.x + -d.y * n.y + -d.z * n.z
            if cos_wo < 1e-4:
                break
            mat = materials.read(inst)
            if mat.w != 0.0:
                if depth == 0:
                    L = float3(L.x + light_emission.x, L.y + light_emission.y, L.z + light_emission.z)
                else:
                    dp = p - o
                    pdf_light = dot(dp, dp) / (light_area * cos_wo)
                    mis = pdf / max(pdf + pdf_light, 1e-4)
                    L = float3(L.x + mis * beta.x * light_emission.x, L.y + mis * beta.y * light_emission.y,
                               L.z + mis * beta.z * light_emission.z)
                break
            # light sample
            ulx = pcg_float()
            uly = pcg_float()
            pl = float3(light_position.x + ulx * light_u.x + uly * light_v.x,
                        light_position.y + ulx * light_u.y + uly * light_v.y,
                        light_position.z + ulx * light_u.z + uly * light_v.z)
            pp = offset_ray_origin(p, n)
            ppl = offset_ray_origin(pl, light_n)
            dl = ppl - pp
            d_light = sqrt(dot(dl, dl))
            wi = normalize(dl)
            cos_wi_light = dot(wi, n)
            cos_light = -dot(light_n, wi)
            shadow = cos_wi_light > 1e-4 and cos_light > 1e-4
            contrib = float3(0.0, 0.0, 0.0)
            if shadow:
                pdf_l = (d_light * d_light) / (light_area * cos_light)
                pdf_b = cos_wi_light * INV_PI
                mis_l = pdf_l / max(pdf_l + pdf_b, 1e-4)
                den = max(pdf_l, 1e-4)
                contrib = float3(beta.x * (mat.x * INV_PI * cos_wi_light) * mis_l * light_emission.x / den,
                                 beta.y * (mat.y * INV_PI * cos_wi_light) * mis_l * light_emission.y / den,
                                 beta.z * (mat.z * INV_PI * cos_wi_light) * mis_l * light_emission.z / den)
            # cosine-weighted direction in the make_onb frame
            ux = pcg_float()
            uy = pcg_float()
            r = sqrt(ux)
            x4 = uy * 4.0
            k = int(x4)
            f = x4 - float(k)
            t = f * 1.57079632679489662
            t2 = t * t
            ps = t * (1.0 + t2 * (-1.0 / 6.0 + t2 * (1.0 / 120.0 + t2 * (-1.0 / 5040.0 + t2 * (1.0 / 362880.0 +
                                                                                              t2 * (-1.0 / 39916800.0))))))
            pc = 1.0 + t2 * (-1.0 / 2.0 + t2 * (1.0 / 24.0 + t2 * (-1.0 / 720.0 + t2 * (1.0 / 40320.0 + t2 * (
                -1.0 / 3628800.0 + t2 * (1.0 / 479001600.0))))))
            q4 = k & 3
            sp = ps
            cp = pc
            if q4 == 1:
                sp = pc
                cp = -ps
            elif q4 == 2:
                sp = -ps
                cp = -pc
            elif q4 == 3:
                sp = -pc
                cp = ps
            wl = float3(r * cp, r * sp, sqrt(1.0 - ux))
            bn = normalize(float3(-n.y, n.x, 0.0) if abs(n.x) > abs(n.z) else float3(0.0, -n.z, n.y))
            tn = normalize(cross(bn, n))
            nd = float3(wl.x * tn.x + wl.y * bn.x + wl.z * n.x, wl.x * tn.y + wl.y * bn.y + wl.z * n.y,
                        wl.x * tn.z + wl.y * bn.z + wl.z * n.z)
            pdf = abs(wl.z) * INV_PI
            beta = float3(beta.x * mat.x, beta.y * mat.y, beta.z * mat.z)
            cont = depth + 1 < max_depth
            if depth >= rr_depth:
                lum = 0.212671 * beta.x + 0.715160 * beta.y + 0.072169 * beta.z
                if lum == 0.0:
                    cont = False
                else:
                    qq = max(lum, 0.05)
                    rr = pcg_float()
                    if rr >= qq:
                        cont = False
                    else:
                        iq = 1.0 / qq
                        beta = float3(beta.x * iq, beta.y * iq, beta.z * iq)
            if shadow:
                if not accel.trace_any(make_ray(offset_ray_origin(pp, n), wi, 0.0, d_light), 255):
                    L = float3(L.x + contrib.x, L.y + contrib.y, L.z + contrib.z)
            if not cont:
                break
            o = pp
            d = nd
        if isnan(L.x) or isnan(L.y) or isnan(L.z):
            L = float3(0.0, 0.0, 0.0)
        total = float3(total.x + L.x, total.y + L.y, total.z + L.z)
    fs = float(spp)
    image.write(pixel, float4(total.x / fs, total.y / fs, total.z / fs, 1.0))


# accel.traverse(ray, {}).on_surface_candidate(...).on_procedural_candidate(...).trace() with the
# callbacks of src/tests/test_procedural.cpp:86-120 written as the codegen would inline them
QUERY = r'''
#include "lc_hip_device.h"
using namespace lchip;
struct Params {
    alignas(16) dev::Buffer<dev::Ray> rays;
    alignas(16) dev::Buffer<rq::CommittedHit> out;
    alignas(16) dev::Accel accel;
    alignas(16) float cutout;
    alignas(16) uint32_t any;
    alignas(16) uint4 ls_kid;
};
extern "C" __global__ void __launch_bounds__(128) query_kernel(const Params p) {
    if (!dev::in_dispatch(p.ls_kid)) return;
    const uint32_t i = dev::dispatch_id().x;
    const float cutout = p.cutout;
    auto on_surface = [cutout](uint32_t, uint32_t, float u, float v, float) {
        const float w = (1.0f - u) - v;
        return sqrtf(w * w + u * u) < cutout && sqrtf(u * u + v * v) < cutout && sqrtf(v * v + w * w) < cutout;
    };
    auto on_procedural = [](uint32_t, uint32_t, const float *o, const float *d, const float *, const float *,
                            float4 lo, float4 hi, float, float &t) {
        const float c[3] = {(lo.x + hi.x) * 0.5f, (lo.y + hi.y) * 0.5f, (lo.z + hi.z) * 0.5f};
        const float r = 0.5f * fminf(fminf(hi.x - lo.x, hi.y - lo.y), hi.z - lo.z);
        const float L[3] = {c[0] - o[0], c[1] - o[1], c[2] - o[2]};
        const float dd = (d[0] * d[0] + d[1] * d[1]) + d[2] * d[2];
        const float ld = (L[0] * d[0] + L[1] * d[1]) + L[2] * d[2];
        const float ll = (L[0] * L[0] + L[1] * L[1]) + L[2] * L[2];
        const float tc = ld / dd;
        if (!(tc > 0.0f)) return false;
        const float q = ll - (tc * tc) * dd;
        const float rr = r * r;
        if (!(q <= rr)) return false;
        t = tc - sqrtf((rr - q) / dd);
        return true;
    };
    p.out[i] = dev::query(p.accel, p.rays[i], 0xFFu, p.any != 0u, on_surface, on_procedural);
}
'''


class PathTracerBindings:
    """The resources `path_tracer` reads for one scene: heap slots 2i / 2i + 1 = vertex / triangle
    buffers of instance i's mesh, xforms = the instances' object -> world affines, materials =
    (albedo, emissive) per instance (the lc_hip_pt_instance list the library path tracer shades
    the same accel with). `args(image, spp, ...)` binds one request."""

    def __init__(self, device, stream, accel, pt_instances, scene):
        import numpy as np
        from . import host
        n = len(accel)
        if len(pt_instances) != n:
            raise ValueError("one material per accel instance")
        self.accel, self.scene = accel, scene
        self.heap = device.create_bindless_array(2 * n)
        for i, m in enumerate(accel._meshes):
            self.heap.emplace(2 * i, m.vb)
            self.heap.emplace(2 * i + 1, m.tb)
        stream << self.heap.update()
        xf = np.concatenate([accel.instance_info(i)[0] for i in range(n)]).astype(np.float32)
        self.xforms = host.upload(device, stream, xf)
        mats = np.array([list(p.albedo) + [1.0 if p.emissive else 0.0] for p in pt_instances], np.float32)
        self.materials = host.upload(device, stream, mats)

    def vec(self, name):
        import numpy as np
        return np.array(getattr(self.scene, name)[:], np.float32)

    def args(self, image, spp, sample_offset=0, max_depth=8, rr_depth=3, seed=42, sequence=0):
        v = self.vec
        return (image, self.accel, self.heap, self.xforms, self.materials, v("light_position"), v("light_u"),
                v("light_v"), v("light_emission"), v("camera_origin"), v("camera_right"), v("camera_up"),
                v("camera_forward"), spp, sample_offset, max_depth, rr_depth, seed, sequence)
