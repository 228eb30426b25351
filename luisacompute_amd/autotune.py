"""Traversal autotuner: block size, LDS stack depth, steps between ray refills, blocks per CU and
whether to traverse a sorted ray queue, chosen per trace kind (closest / any) by timing the real
kernels (sort included) on a workload with HIP events.

The table is JSON ({"closest": {...}, "any": {...}, "device": ..., "workload": ...}) and is applied
with `apply(device, table)`. `python -m luisacompute_amd.autotune --out table.json` tunes on the
soup workload (scaled by --tris/--rays).
"""
import argparse
import itertools
import json
import os

from . import Device
from . import host

BLOCKS = (64, 128, 256)
STACKS = (8, 16, 24)
STEPS = (8, 32, 128)
BLOCKS_PER_CU = (0,)
SORT = (False, True)
FETCH = ((0, 1), (64, 1), (256, 1), (64, 8), (256, 8))  # (slots per claim, queue segments)

DEFAULT_TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "autotune_table.json")


def apply(device, table):
    for kind in ("closest", "any"):
        c = table.get(kind)
        if c:
            device.set_trace_config(kind, c["block"], c["lds_stack"], c["steps"], c.get("blocks_per_cu", 0))
            device.set_ray_sort(kind, c.get("sort_rays", False))
            device.set_fetch(kind, c.get("fetch_chunk", 0), c.get("fetch_segments", 1))


def load_default(device):
    """Applies the shipped table if present; returns it (or None)."""
    if os.path.exists(DEFAULT_TABLE):
        with open(DEFAULT_TABLE) as f:
            t = json.load(f)
        apply(device, t)
        return t
    return None


def _time(device, stream, fn, kind, repeats):
    fn()  # warm
    stream.synchronize()
    device.reset_timing()
    device.enable_timing(True)
    for _ in range(repeats):
        fn()
    stream.synchronize()
    device.enable_timing(False)
    ms, n = device.timing("trace_" + kind)
    return ms / max(n, 1)


def tune(device, stream, accel, ray_buffer, hit_buffer, any_buffer, count, repeats=3, space=None, log=print):
    """Two stages per kind: launch shape without sorting, then the sort flag on the best shapes."""
    space = space or list(itertools.product(BLOCKS, STACKS, STEPS, BLOCKS_PER_CU))
    result = {}
    for kind, buf in (("closest", hit_buffer), ("any", any_buffer)):
        fn = (lambda: accel.trace_closest(stream, ray_buffer, buf, count)) if kind == "closest" else \
             (lambda: accel.trace_any(stream, ray_buffer, buf, count))
        timed = []
        for srt in SORT:
            device.set_ray_sort(kind, srt)
            for block, slds, steps, bpc in space:
                device.set_trace_config(kind, block, slds, steps, bpc)
                ms = _time(device, stream, fn, kind, repeats)
                log(f"  {kind:7s} sort={int(srt)} block={block:3d} stack={slds:2d} steps={steps:3d} bpc={bpc}: "
                    f"{ms:.3f} ms ({count / ms / 1e3:.1f} Mrays/s)")
                timed.append((ms, dict(block=block, lds_stack=slds, steps=steps, blocks_per_cu=bpc, sort_rays=srt)))
        best = min(timed, key=lambda x: x[0])
        c = dict(best[1])
        device.set_trace_config(kind, c["block"], c["lds_stack"], c["steps"], c["blocks_per_cu"])
        device.set_ray_sort(kind, c["sort_rays"])
        # third stage: how the launch hands out queue slots, on the best shape
        fetched = []
        for chunk, segs in FETCH:
            device.set_fetch(kind, chunk, segs)
            ms = _time(device, stream, fn, kind, repeats)
            log(f"  {kind:7s} fetch chunk={chunk:3d} segments={segs}: {ms:.3f} ms")
            fetched.append((ms, chunk, segs))
        ms, chunk, segs = min(fetched)
        device.set_fetch(kind, chunk, segs)
        c.update(fetch_chunk=chunk, fetch_segments=segs)
        result[kind] = dict(c, ms=round(ms, 4), mrays_s=round(count / ms / 1e3, 2))
    return result


def main():
    import luisacompute_amd as lc
    ap = argparse.ArgumentParser()
    ap.add_argument("--tris", type=int, default=1 << 24)
    ap.add_argument("--rays", type=int, default=1 << 23)
    ap.add_argument("--out", default=DEFAULT_TABLE)
    a = ap.parse_args()
    dev = Device(0)
    s = dev.create_stream()
    v, t = lc.soup(a.tris)
    rays = lc.rays_box_sphere(a.rays, skip=13 * a.tris)
    accel = host.single_mesh_accel(dev, s, v, t)
    rb = host.upload(dev, s, rays)
    hb = dev.create_buffer(24, a.rays)
    ab = dev.create_buffer(1, a.rays)
    res = tune(dev, s, accel, rb, hb, ab, a.rays)
    res["device"] = dev.query("device_name")
    res["workload"] = f"soup {a.tris} tris, {a.rays} rays"
    with open(a.out, "w") as f:
        json.dump(res, f, indent=1)
    print(json.dumps(res))


if __name__ == "__main__":
    main()
