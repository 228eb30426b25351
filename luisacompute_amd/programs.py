"""Programs the benchmark runs through the run-time shader path, shared with the tests.

  path_tracer : the path tracer of src/tests/test_path_tracing.cpp:165-258 written in the Python DSL
                (the AST -> HIP codegen path; one thread per pixel, the megakernel form)
  wf_*, WavefrontPathTracer : the same path tracer in wavefront form, written in the DSL: camera,
                shade and resolve kernels over ray queues in device memory, traversal by the
                library's persistent trace over the device-counted queues, one sample's launch
                sequence captured once into a HIP graph and replayed per sample
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


# ---- the wavefront form. A wave holds `samples` consecutive samples of every pixel (path
# i = j * n + pixel for sample frame[0] + j). Per wave: wf_camera writes one camera ray per path
# into queue 0; per bounce, the closest hits of the live queue (trace_closest_indirect over the
# count the previous kernel left in counters[q]), wf_shade (the megakernel's loop body between the
# closest hit and the shadow ray, appending continuation rays to the other queue and shadow rays to
# the shadow queue, one atomic per wave and queue), the shadow queue's any-hits and wf_resolve
# (adds the unoccluded shadow contributions); then wf_accumulate adds the wave's samples to each
# pixel's sum in sample order. Path state (ray, throughput and pdf, radiance, PCG state) lives in
# per-path buffers between the kernels; every value is computed by the megakernel's expressions in
# the same order, so the image equals the megakernel's bit for bit.
# counters (one 128-B line each, so that their atomics do not serialise on one line): [0] and [32]
# the continuation queue lengths, [64] the shadow queue length; frame[0]: the wave's first sample.

@dsl.kernel(block=(256, 1, 1))
def wf_clear(total: Buffer("float4"), frame: Buffer("uint"), n: uint):
    """Clears the running sums and restarts the sample counter."""
    i = dispatch_id().x
    if i >= n:
        return
    total.write(i, float4(0.0, 0.0, 0.0, 0.0))
    if i == 0:
        frame.write(0, 0)


@dsl.kernel(block=(256, 1, 1))
def wf_camera(rays: Buffer("Ray"), paths: Buffer("uint"), throughput: Buffer("float4"), radiance: Buffer("float4"),
              rng: Buffer("ulong"), counters: Buffer("uint"), frame: Buffer("uint"), camera_origin: dsl.float3_,
              camera_right: dsl.float3_, camera_up: dsl.float3_, camera_forward: dsl.float3_, width: uint,
              height: uint, samples: uint, sample_offset: uint, seed: dsl.ulong, sequence: dsl.ulong):
    """Path i = j * n + pixel traces sample frame[0] + j of its pixel: the megakernel's PCG stream
    (seeded per pixel, then advanced by 64 draws per earlier sample) and camera ray."""
    i = dispatch_id().x
    n = width * height
    if i >= n * samples:
        return
    pixel = i % n
    s = frame.read(0) + i // n
    x = pixel % width
    y = pixel // width
    state = ulong(0)
    inc = ((sequence + pixel) << 1) | 1

    def pcg_next():
        old = state
        state = old * 6364136223846793005 + inc
        xs = uint(((old >> 18) ^ old) >> 27)
        rot = uint(old >> 59)
        return (xs >> rot) | (xs << ((0 - rot) & 31))

    def pcg_float():
        return float(pcg_next() >> 8) * 5.9604644775390625e-08

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
    fr = float(min(width, height))
    fx = ((float(x) + rx) * 2.0 - float(width)) / fr
    fy = ((float(y) + ry) * 2.0 - float(height)) / fr
    d = normalize(float3(fx * camera_right.x - fy * camera_up.x + camera_forward.x,
                         fx * camera_right.y - fy * camera_up.y + camera_forward.y,
                         fx * camera_right.z - fy * camera_up.z + camera_forward.z))
    rays.write(i, make_ray(camera_origin, d, 0.0, FLT_MAX))
    paths.write(i, i)
    throughput.write(i, float4(1.0, 1.0, 1.0, 0.0))  # beta, pdf of the last bounce
    radiance.write(i, float4(0.0, 0.0, 0.0, 0.0))
    rng.write(2 * i, state)
    rng.write(2 * i + 1, inc)
    if i == 0:
        counters.write(0, n * samples)


@dsl.kernel(block=(64, 1, 1))
def wf_reset(counters: Buffer("uint"), out_queue: uint):
    """Empties the queue the next wf_shade appends continuation rays to, and the shadow queue."""
    if dispatch_id().x == 0:
        counters.write(32 * out_queue, 0)
        counters.write(64, 0)


@dsl.kernel(block=(512, 1, 1))
def wf_shade(rays_in: Buffer("Ray"), paths_in: Buffer("uint"), hits: Buffer("SurfaceHit"), rays_out: Buffer("Ray"),
             paths_out: Buffer("uint"), shadow_rays: Buffer("Ray"), shadow_paths: Buffer("uint"),
             shadow_contrib: Buffer("float4"), throughput: Buffer("float4"), radiance: Buffer("float4"),
             rng: Buffer("ulong"), counters: Buffer("uint"), heap: dsl.BindlessArray, xforms: Buffer("float"),
             materials: Buffer("float4"), light_position: dsl.float3_, light_u: dsl.float3_, light_v: dsl.float3_,
             light_emission: dsl.float3_, depth: uint, max_depth: uint, rr_depth: uint, in_queue: uint):
    """One bounce for the entries of queue `in_queue`: the megakernel's loop body from the closest
    hit to the shadow ray for each entry's path (a one-pass loop whose `break`s are the
    megakernel's early exits, so that every thread reaches the appends). Continuation and shadow
    rays are appended with one global atomic per workgroup and queue: lanes count through a
    shared array, one thread reserves the workgroup's slots (the dispatch covers whole
    workgroups)."""
    j = dispatch_id().x
    shadow = False
    cont = False
    p = uint(0)
    pp = float3(0.0, 0.0, 0.0)
    n = float3(0.0, 0.0, 0.0)
    wi = float3(0.0, 0.0, 0.0)
    nd = float3(0.0, 0.0, 0.0)
    d_light = 0.0
    contrib = float3(0.0, 0.0, 0.0)
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

    while True:
        if j >= counters.read(32 * in_queue):
            break
        p = paths_in.read(j)
        h = hits.read(j)
        if hit_miss(h):
            break
        r = rays_in.read(j)
        o = ray_origin(r)
        d = ray_direction(r)
        state = rng.read(2 * p)
        inc = rng.read(2 * p + 1)
        tp = throughput.read(p)
        beta = float3(tp.x, tp.y, tp.z)
        pdf = tp.w
        lr = radiance.read(p)
        L = float3(lr.x, lr.y, lr.z)
        lc_ = cross(light_u, light_v)
        light_area = length(lc_)
        light_n = normalize(lc_)
        inst = h.inst
        tri = heap.buffer_read("uint3", 2 * inst + 1, h.prim)
        p0 = to_world(inst, heap.buffer_read("float4", 2 * inst, tri.x))
        p1 = to_world(inst, heap.buffer_read("float4", 2 * inst, tri.y))
        p2 = to_world(inst, heap.buffer_read("float4", 2 * inst, tri.z))
        pt = triangle_interpolate(hit_bary(h), p0, p1, p2)
        n = normalize(cross(p1 - p0, p2 - p0))
        cos_wo = -d.x * n.x + -d.y * n.y + -d.z * n.z
        if cos_wo < 1e-4:
            break
        mat = materials.read(inst)
        if mat.w != 0.0:
            if depth == 0:
                L = float3(L.x + light_emission.x, L.y + light_emission.y, L.z + light_emission.z)
            else:
                dp = pt - o
                pdf_light = dot(dp, dp) / (light_area * cos_wo)
                mis = pdf / max(pdf + pdf_light, 1e-4)
                L = float3(L.x + mis * beta.x * light_emission.x, L.y + mis * beta.y * light_emission.y,
                           L.z + mis * beta.z * light_emission.z)
            radiance.write(p, float4(L.x, L.y, L.z, 0.0))
            break
        # light sample
        ulx = pcg_float()
        uly = pcg_float()
        pl = float3(light_position.x + ulx * light_u.x + uly * light_v.x,
                    light_position.y + ulx * light_u.y + uly * light_v.y,
                    light_position.z + ulx * light_u.z + uly * light_v.z)
        pp = offset_ray_origin(pt, n)
        ppl = offset_ray_origin(pl, light_n)
        dl = ppl - pp
        d_light = sqrt(dot(dl, dl))
        wi = normalize(dl)
        cos_wi_light = dot(wi, n)
        cos_light = -dot(light_n, wi)
        shadow = cos_wi_light > 1e-4 and cos_light > 1e-4
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
        rr_ = sqrt(ux)
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
        wl = float3(rr_ * cp, rr_ * sp, sqrt(1.0 - ux))
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
        throughput.write(p, float4(beta.x, beta.y, beta.z, pdf))
        rng.write(2 * p, state)
        break
    # appends: lanes count per wave into the shared counts, one thread reserves the workgroup's
    # slots of each queue, lanes write in wave order within the workgroup's range
    wg = shared_array("uint", 4)
    tid = thread_id().x
    if tid == 0:
        wg[0] = 0
        wg[1] = 0
    sync_block()
    ns = warp_active_count_bits(shadow)
    nc = warp_active_count_bits(cont)
    sw = uint(0)
    cw = uint(0)
    if warp_is_first_active_lane():
        sw = wg.atomic_fetch_add(0, ns)
        cw = wg.atomic_fetch_add(1, nc)
    sw = warp_read_first_active_lane(sw)
    cw = warp_read_first_active_lane(cw)
    sync_block()
    if tid == 0:
        wg[2] = counters.atomic_fetch_add(64, wg[0])
        wg[3] = counters.atomic_fetch_add(32 * (1 - in_queue), wg[1])
    sync_block()
    if shadow:
        si = wg[2] + sw + warp_prefix_count_bits(shadow)
        shadow_rays.write(si, make_ray(offset_ray_origin(pp, n), wi, 0.0, d_light))
        shadow_paths.write(si, p)
        shadow_contrib.write(si, float4(contrib.x, contrib.y, contrib.z, 0.0))
    if cont:
        ci = wg[3] + cw + warp_prefix_count_bits(cont)
        rays_out.write(ci, make_ray(pp, nd, 0.0, FLT_MAX))
        paths_out.write(ci, p)


@dsl.kernel(block=(256, 1, 1))
def wf_resolve(shadow_paths: Buffer("uint"), shadow_contrib: Buffer("float4"), occluded: Buffer("ubyte"),
               radiance: Buffer("float4"), counters: Buffer("uint")):
    """Adds the contribution of every unoccluded shadow ray to its path (at most one shadow ray per
    path and bounce, so the updates do not collide)."""
    j = dispatch_id().x
    if j >= counters.read(64):
        return
    if occluded.read(j) == 0:
        p = shadow_paths.read(j)
        c = shadow_contrib.read(j)
        lr = radiance.read(p)
        radiance.write(p, float4(lr.x + c.x, lr.y + c.y, lr.z + c.z, 0.0))


@dsl.kernel(block=(256, 1, 1))
def wf_accumulate(radiance: Buffer("float4"), total: Buffer("float4"), frame: Buffer("uint"), n: uint, samples: uint):
    """Adds the wave's sample radiances of each pixel (NaN samples dropped) to its running sum in
    sample order and advances the sample counter (read by no other thread of this launch)."""
    i = dispatch_id().x
    if i >= n:
        return
    tt = total.read(i)
    sum_ = float3(tt.x, tt.y, tt.z)
    for j in range(samples):
        lr = radiance.read(j * n + i)
        L = float3(lr.x, lr.y, lr.z)
        if isnan(L.x) or isnan(L.y) or isnan(L.z):
            L = float3(0.0, 0.0, 0.0)
        sum_ = float3(sum_.x + L.x, sum_.y + L.y, sum_.z + L.z)
    total.write(i, float4(sum_.x, sum_.y, sum_.z, 0.0))
    if i == 0:
        frame.write(0, frame.read(0) + samples)


@dsl.kernel(block=(256, 1, 1))
def wf_finish(total: Buffer("float4"), image: Buffer("float4"), n: uint, spp: uint):
    """image = running sum / spp, alpha 1 (the megakernel's last write)."""
    i = dispatch_id().x
    if i >= n:
        return
    t = total.read(i)
    fs = float(spp)
    image.write(i, float4(t.x / fs, t.y / fs, t.z / fs, 1.0))


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


class WavefrontPathTracer:
    """Renders `path_tracer`'s image with the wavefront kernels above for one scene's bindings at
    width x height. A wave carries `samples_per_wave` samples of every pixel (default: about 32M
    paths, at most 64 samples); queue and path-state buffers hold one wave (about 200 B per path).
    Traversal runs through Accel.trace_closest_indirect / trace_any_indirect over the
    device-counted queues. With graph=True a wave's launch sequence (camera, max_depth x {reset,
    closest hits, shade, shadow any-hits, resolve}, accumulate) runs once eagerly, is then captured
    into a graph (Stream.capture) and replayed for the following waves; the first sample of a wave
    advances in device memory, so every replay is the same launch sequence. sort_queues: the
    traces sort their queues (origin cell, direction octant; the library's radix sort, which
    reads the device-side count) before traversal; the image is the same either way."""

    KERNELS = ("wf_clear", "wf_camera", "wf_reset", "wf_shade", "wf_resolve", "wf_accumulate", "wf_finish")

    def __init__(self, device, stream, bindings, width, height, samples_per_wave=None, sort_queues=False):
        self.device, self.b = device, bindings
        self.sort_queues = bool(sort_queues)
        self.width, self.height = width, height
        n = self.n = width * height
        k = samples_per_wave or max(1, min(64, (32 << 20) // n))
        if not 1 <= k or n * k >= 1 << 31:
            raise ValueError("samples_per_wave x pixels must be below 2^31")
        self.samples_per_wave = k
        m = n * k
        B = device.create_buffer
        self.rays = [B(32, m), B(32, m)]
        self.paths = [B(4, m), B(4, m)]
        self.hits = B(24, m)
        self.shadow_rays, self.shadow_paths = B(32, m), B(4, m)
        self.shadow_contrib, self.occluded = B(16, m), B(1, m)
        self.throughput, self.radiance = B(16, m), B(16, m)
        self.rng = B(8, 2 * m)
        self.total = B(16, n)
        self.counters, self.frame = B(4, 96), B(4, 4)
        g = globals()
        self.k = {name: g[name].compile(device) for name in self.KERNELS}
        self._waves = {}

    def _wave_commands(self, samples, max_depth, rr_depth, seed, sequence, sample_offset):
        """Callables issuing one wave's launches on a stream, in order."""
        v, n, k, b = self.b.vec, self.n, self.k, self.b
        m = n * samples
        accel = b.accel
        cam = k["wf_camera"](self.rays[0], self.paths[0], self.throughput, self.radiance, self.rng, self.counters,
                             self.frame, v("camera_origin"), v("camera_right"), v("camera_up"), v("camera_forward"),
                             self.width, self.height, samples, sample_offset, seed, sequence).dispatch(m)
        cmds = [lambda s, c=cam: s << c]
        # the callables hold the buffers themselves, not self (no reference cycle through self)
        hits, ctr, srays, occ = self.hits, self.counters, self.shadow_rays, self.occluded
        dev, sort = self.device, self.sort_queues

        def traced(kind, fn):
            def call(s):
                was = dev.ray_sort(kind)
                dev.set_ray_sort(kind, sort)
                try:
                    fn(s)
                finally:
                    dev.set_ray_sort(kind, was)
            return call

        # wf_shade's workgroups all reach its barriers: whole workgroups, the tail idle
        shade_threads = -(-m // 512) * 512
        for depth in range(max_depth):
            qi = depth & 1
            reset = k["wf_reset"](self.counters, 1 - qi).dispatch(1)
            shade = k["wf_shade"](self.rays[qi], self.paths[qi], self.hits, self.rays[1 - qi], self.paths[1 - qi],
                                  self.shadow_rays, self.shadow_paths, self.shadow_contrib, self.throughput,
                                  self.radiance, self.rng, self.counters, b.heap, b.xforms, b.materials,
                                  v("light_position"), v("light_u"), v("light_v"), v("light_emission"), depth,
                                  max_depth, rr_depth, qi).dispatch(shade_threads)
            resolve = k["wf_resolve"](self.shadow_paths, self.shadow_contrib, self.occluded, self.radiance,
                                      self.counters).dispatch(m)
            cmds += [lambda s, c=reset: s << c,
                     traced("closest", lambda s, r=self.rays[qi], q=qi: accel.trace_closest_indirect(
                         s, r, hits, ctr, 32 * q, capacity=m)),
                     lambda s, c=shade: s << c,
                     traced("any", lambda s: accel.trace_any_indirect(s, srays, occ, ctr, 64, capacity=m)),
                     lambda s, c=resolve: s << c]
        acc = k["wf_accumulate"](self.radiance, self.total, self.frame, n, samples).dispatch(n)
        cmds.append(lambda s, c=acc: s << c)
        return cmds

    def render(self, stream, image, spp, sample_offset=0, max_depth=8, rr_depth=3, seed=42, sequence=0, graph=True):
        """image: Buffer of width x height float4 (as path_tracer writes it)."""
        if not 1 <= max_depth <= 64:
            raise ValueError("max_depth must be in 1..64")
        if spp <= 0:
            raise ValueError("spp must be positive")
        n, s = self.n, stream
        s << self.k["wf_clear"](self.total, self.frame, n).dispatch(n)
        done = 0
        while done < spp:
            samples = min(self.samples_per_wave, spp - done)
            key = (samples, max_depth, rr_depth, seed, sequence, sample_offset)
            entry = self._waves.get(key)
            if entry is None:
                entry = self._waves[key] = {"cmds": self._wave_commands(*key), "graph": None, "runs": 0}
            if graph and entry["graph"] is not None:
                entry["graph"].launch(s)
            elif graph and entry["runs"] > 0:
                # traces have sized their working memory in an eager run: capture, then replay
                with s.capture() as cap:
                    for c in entry["cmds"]:
                        c(s)
                entry["graph"] = cap.graph
                entry["graph"].launch(s)
            else:
                for c in entry["cmds"]:
                    c(s)
            entry["runs"] += 1
            done += samples
        s << self.k["wf_finish"](self.total, image, n, spp).dispatch(n)
