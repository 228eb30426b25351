"""Autotuner: every launch knob of the ray-tracing path is chosen by timing the real kernels with
HIP events on the workloads they serve, and the shipped table is this module's output file.

Stages (each on the best configuration of the stages before it):
  1. traversal, per trace kind (closest / any) on the 16.7M-triangle soup, searched jointly: block
     size, LDS stack depth, steps between ray refills, sorted or unsorted queue, and how the launch
     hands out queue slots (slots per claim, one queue or a segment per XCD);
  1b. the traversal step policy on the best stage-1 launch shape and hand-out: the share of live
     lanes that must hold triangles for a closest-hit wave to take a triangle step (1..3 quarters),
     the early refill of waves whose live lanes fall below a count (off, 32, 48), and the steps
     between refills again;
  2. BVH width (children per wide node, 4 / 6 / 8) on the soup step it serves: BLAS build +
     closest-hit trace;
  3. the wavefront path tracer on the box room: samples per wave (queue chunking), the fetch of its
     traces and the bucket sort of continuation rays;
  4. the serving loop: paths per shared wave over a backlog of requests on the four resident scenes.

The table is JSON ({"closest": {...}, "any": {...}, "bvh": {...}, "pathtracer": {...},
"serving": {...}, "device": ..., "workload": ...}); `apply(device, table)` sets the device knobs
(traversal, BVH width) and `pathtracer_options(table)` / `serving_options(table)` give the
per-call ones. `python -m luisacompute_amd.autotune --out table.json` runs every stage.
"""
import argparse
import itertools
import json
import os
import sys
import time

from . import Device
from . import host

BLOCKS = (64, 128, 256)
STACKS = (8, 16, 24)
STEPS = (8, 16, 32, 128)
BLOCKS_PER_CU = (0,)
SORT = (False, True)
FETCH = ((0, 1), (64, 1), (256, 1), (64, 8), (256, 8))  # (slots per claim, queue segments)
TRI_QUARTERS = (1, 2, 3)
REFILL_BELOW = (0, 32, 48)
WIDTHS = (4, 6, 8)
PT_CHUNKS = (0, 2, 4, 8, 16)        # samples per wave (0: about 32M paths per wave)
PT_FETCH = ((0, 1), (64, 8), (128, 8), (256, 8))
PT_SORT = (0, 1)
SERVE_WAVES = (8 << 20, 16 << 20, 32 << 20, 64 << 20)

DEFAULT_TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "autotune_table.json")


def apply(device, table):
    for kind in ("closest", "any"):
        c = table.get(kind)
        if c:
            device.set_trace_config(kind, c["block"], c["lds_stack"], c["steps"], c.get("blocks_per_cu", 0))
            device.set_ray_sort(kind, c.get("sort_rays", False))
            device.set_fetch(kind, c.get("fetch_chunk", 0), c.get("fetch_segments", 1))
            device.set_trace_steps(kind, c.get("tri_quarters", 2), c.get("refill_below", 0))
    b = table.get("bvh")
    if b and "width" in b:
        device.set_build_width(b["width"])


def load_default(device):
    """Applies the shipped table if present; returns it (or None)."""
    t = load_table()
    if t is not None:
        apply(device, t)
    return t


def load_table(path=DEFAULT_TABLE):
    if os.path.exists(path):
        with open(path) as f:
            return json.load(f)
    return None


def pathtracer_options(table):
    """(chunk_spp, sort, (fetch_chunk, fetch_segments)) of the wavefront path tracer."""
    p = (table or {}).get("pathtracer") or {}
    return p.get("chunk_spp", 0), p.get("sort", 0), (p.get("fetch_chunk", 128), p.get("fetch_segments", 8))


def serving_options(table):
    """wave_paths of the serving loop."""
    p = (table or {}).get("serving") or {}
    return p.get("wave_paths", 32 << 20)


def _time(device, stream, fn, name, repeats):
    fn()  # warm
    stream.synchronize()
    device.reset_timing()
    device.enable_timing(True)
    for _ in range(repeats):
        fn()
    stream.synchronize()
    device.enable_timing(False)
    ms, n = device.timing(name)
    return ms / max(n, 1)


def tune(device, stream, accel, ray_buffer, hit_buffer, any_buffer, count, repeats=3, space=None, log=print):
    """Stage 1: every combination of launch shape, queue order and queue hand-out, per kind (the
    best shape depends on the hand-out, so the search is joint, not one knob at a time)."""
    space = space or list(itertools.product(BLOCKS, STACKS, STEPS, BLOCKS_PER_CU))
    result = {}
    for kind, buf in (("closest", hit_buffer), ("any", any_buffer)):
        fn = (lambda: accel.trace_closest(stream, ray_buffer, buf, count)) if kind == "closest" else \
             (lambda: accel.trace_any(stream, ray_buffer, buf, count))
        timed = []
        for srt in SORT:
            device.set_ray_sort(kind, srt)
            for (block, slds, steps, bpc), (chunk, segs) in itertools.product(space, FETCH):
                device.set_trace_config(kind, block, slds, steps, bpc)
                device.set_fetch(kind, chunk, segs)
                ms = _time(device, stream, fn, "trace_" + kind, repeats)
                log(f"  {kind:7s} sort={int(srt)} block={block:3d} stack={slds:2d} steps={steps:3d} bpc={bpc} "
                    f"fetch={chunk:3d}/{segs}: {ms:.3f} ms ({count / ms / 1e3:.1f} Mrays/s)")
                timed.append((ms, dict(block=block, lds_stack=slds, steps=steps, blocks_per_cu=bpc, sort_rays=srt,
                                       fetch_chunk=chunk, fetch_segments=segs)))
        ms, c = min(timed, key=lambda x: x[0])
        device.set_trace_config(kind, c["block"], c["lds_stack"], c["steps"], c["blocks_per_cu"])
        device.set_ray_sort(kind, c["sort_rays"])
        device.set_fetch(kind, c["fetch_chunk"], c["fetch_segments"])
        result[kind] = dict(c, ms=round(ms, 4), mrays_s=round(count / ms / 1e3, 2))
    return result


def tune_steps(device, stream, accel, ray_buffer, hit_buffer, any_buffer, count, result, repeats=3, log=print):
    """Stage 1b: the step policy on the best stage-1 configuration of each kind (the triangle-step
    threshold applies to closest hit only)."""
    for kind, buf in (("closest", hit_buffer), ("any", any_buffer)):
        fn = (lambda: accel.trace_closest(stream, ray_buffer, buf, count)) if kind == "closest" else \
             (lambda: accel.trace_any(stream, ray_buffer, buf, count))
        c = result[kind]
        timed = []
        for tq, rb, steps in itertools.product(TRI_QUARTERS if kind == "closest" else (2,), REFILL_BELOW, STEPS):
            device.set_trace_config(kind, c["block"], c["lds_stack"], steps, c["blocks_per_cu"])
            device.set_trace_steps(kind, tq, rb)
            ms = _time(device, stream, fn, "trace_" + kind, repeats)
            log(f"  {kind:7s} tri_quarters={tq} refill_below={rb:2d} steps={steps:3d}: {ms:.3f} ms "
                f"({count / ms / 1e3:.1f} Mrays/s)")
            timed.append((ms, tq, rb, steps))
        ms, tq, rb, steps = min(timed)
        device.set_trace_config(kind, c["block"], c["lds_stack"], steps, c["blocks_per_cu"])
        device.set_trace_steps(kind, tq, rb)
        c.update(steps=steps, tri_quarters=tq, refill_below=rb, ms=round(ms, 4), mrays_s=round(count / ms / 1e3, 2))
    return result


def tune_width(device, stream, mesh, accel, ray_buffer, hit_buffer, count, repeats=3, log=print):
    """Stage 2: the soup step (BLAS build + TLAS + closest hits) at each BVH width."""
    timed = []
    for w in WIDTHS:
        device.set_build_width(w)

        def step():
            stream << mesh.build() << accel.build()
            accel.trace_closest(stream, ray_buffer, hit_buffer, count)
        step()
        stream.synchronize()
        t0 = time.perf_counter()
        for _ in range(repeats):
            step()
        stream.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / repeats
        st = mesh.stats()
        log(f"  bvh width={w}: step {ms:.3f} ms, {st.node_count} wide nodes, depth {st.depth}")
        timed.append((ms, w, st.node_count, st.depth))
    ms, w, nodes, depth = min(timed)
    device.set_build_width(w)
    stream << mesh.build() << accel.build()
    stream.synchronize()
    return {"width": w, "step_ms": round(ms, 4), "wide_nodes": nodes, "depth": depth,
            "tried": {str(x[1]): round(x[0], 4) for x in timed}}


def tune_pathtracer(device, stream, width=1920, height=1080, spp=16, repeats=2, log=print):
    """Stage 3: box room, wavefront mode, hipGraph replay."""
    import luisacompute_amd as lc
    from . import workloads
    room = workloads.BoxRoom()
    accel, _, inst = room.build(device, stream)
    scene = room.scene()
    img = device.create_buffer(16, width * height)
    stream.synchronize()
    timed = []
    for chunk, (fc, fs), srt in itertools.product(PT_CHUNKS, PT_FETCH, PT_SORT):
        o = lc.PtOptions()
        o.width, o.height, o.spp, o.sample_offset, o.max_depth, o.rr_depth = width, height, spp, 0, 8, 3
        o.seed, o.sequence, o.mode, o.chunk_spp, o.sort, o.graph = 42, 3 << 32, lc.PT_WAVEFRONT, chunk, srt, 1
        o.fetch_chunk, o.fetch_segments = fc, fs
        accel.pathtrace(stream, inst, scene, o, img)
        stream.synchronize()
        t0 = time.perf_counter()
        for _ in range(repeats):
            accel.pathtrace(stream, inst, scene, o, img)
        stream.synchronize()
        ms = (time.perf_counter() - t0) * 1e3 / repeats
        log(f"  pathtracer chunk_spp={chunk:2d} fetch=({fc},{fs}) sort={srt}: {ms:.3f} ms")
        timed.append((ms, chunk, fc, fs, srt))
    ms, chunk, fc, fs, srt = min(timed)
    return {"chunk_spp": chunk, "fetch_chunk": fc, "fetch_segments": fs, "sort": srt, "ms": round(ms, 4),
            "workload": f"box room {width}x{height} {spp} spp depth 8"}


def tune_serving(device, stream, requests=24, budget=1 << 30, log=print):
    """Stage 4: makespan of a backlog of requests (all queued at t = 0) per wave size."""
    import luisacompute_amd as lc
    from . import workloads
    scenes = workloads.serving_scenes(device, stream)
    base = workloads.serving_requests(requests, 0.0)
    bufs = [device.create_buffer(16, r["res"] * r["res"]) for r in base]
    stream.synchronize()
    timed = []
    for wave in SERVE_WAVES:
        srv = lc.Server(device, hbm_budget=budget, wave_paths=wave, max_active=64, sort=0, graph=1, fetch=(128, 8))
        ids = [srv.add_scene(acc, inst, sc) for _, acc, inst, sc in scenes]
        for sid in ids:
            srv.submit(sid, 16, 16, 1, bufs[0], seed=1, sequence=1)
        srv.drain()
        t0 = time.perf_counter()
        for r in base:
            srv.submit(ids[r["scene"]], r["res"], r["res"], r["spp"], bufs[r["index"]], seed=42, sequence=r["sequence"])
        srv.drain(timeout=600)
        mk = time.perf_counter() - t0
        srv.close()
        log(f"  serving wave_paths={wave >> 20}M: {requests} requests in {mk:.3f} s")
        timed.append((mk, wave))
    mk, wave = min(timed)
    return {"wave_paths": wave, "makespan_s": round(mk, 4), "requests": requests}


def main():
    import luisacompute_amd as lc
    ap = argparse.ArgumentParser()
    ap.add_argument("--tris", type=int, default=1 << 24)
    ap.add_argument("--rays", type=int, default=1 << 23)
    ap.add_argument("--out", default=DEFAULT_TABLE)
    ap.add_argument("--skip-serving", action="store_true")
    a = ap.parse_args()

    def log(msg):
        print(msg, file=sys.stderr, flush=True)
    dev = Device(0)
    s = dev.create_stream()
    v, t = lc.soup(a.tris)
    rays = lc.rays_box_sphere(a.rays, skip=13 * a.tris)
    # the soup step's BLAS (FAST_BUILD, as the timed step builds it)
    opt = lc.AccelOption(lc.AccelOption.FAST_BUILD, 1, 0)
    mesh = host.mesh_from_arrays(dev, s, v, t, option=opt)
    accel = dev.create_accel()
    accel.emplace_back(mesh, lc.IDENTITY)
    s << accel.build()
    rb = host.upload(dev, s, rays)
    hb = dev.create_buffer(24, a.rays)
    ab = dev.create_buffer(1, a.rays)
    log("[autotune] stage 1: traversal")
    res = tune(dev, s, accel, rb, hb, ab, a.rays, log=log)
    log("[autotune] stage 1b: step policy")
    res = tune_steps(dev, s, accel, rb, hb, ab, a.rays, res, log=log)
    log("[autotune] stage 2: BVH width")
    res["bvh"] = tune_width(dev, s, mesh, accel, rb, hb, a.rays, log=log)
    del accel, mesh, rb, hb, ab
    dev.release_build_arena()
    log("[autotune] stage 3: path tracer")
    res["pathtracer"] = tune_pathtracer(dev, s, log=log)
    if not a.skip_serving:
        log("[autotune] stage 4: serving")
        res["serving"] = tune_serving(dev, s, log=log)
    res["device"] = dev.query("device_name")
    res["workload"] = f"soup {a.tris} tris, {a.rays} rays"
    with open(a.out, "w") as f:
        json.dump(res, f, indent=1)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
