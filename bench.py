"""Benchmark: incoherent closest-hit on a 16,777,216-triangle soup with the BVH build in the step.

Workload (the configuration the headline metric is quoted on): 16,777,216 fp32 triangles with
centroids uniform in [0,100]^3 and vertices offset by N(0, 0.5^2) per axis; 8,388,608 rays with
origins uniform in the cube and directions uniform on the sphere; PCG32 stream (42, 2).
One step = mesh (BLAS) build + accel build + trace_closest of all rays. Inputs are resident in
HBM before timing starts. Prints one JSON line.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
NODE_BYTES = 80       # compressed 8-wide node
TRI_BYTES = 48        # leaf triangle record
RAY_BYTES = 32
HIT_BYTES = 24


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--tris", type=int, default=1 << 24)
    ap.add_argument("--rays", type=int, default=1 << 23)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-sample-rays", type=int, default=1 << 18)
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run 3 device-to-device copies of this many MiB (counter calibration) and exit")
    ap.add_argument("--workload", default="soup", choices=["soup", "instanced", "pathtracer", "serving", "query", "build"],
                    help="soup: the headline line; instanced: 64 ray batches over a 1,024-instance TLAS; "
                         "pathtracer: box room 1920x1080, 64 spp, depth 8 (wavefront vs megakernel); "
                         "serving: render requests over 4 resident scenes (served vs plain)")
    ap.add_argument("--serve-requests", type=int, default=200)
    ap.add_argument("--no-serving", action="store_true", help="soup line: skip the serving leg")
    ap.add_argument("--serve-rate", type=float, default=0.0, help="arrivals/s (0: 0.8 x plain saturated throughput)")
    ap.add_argument("--serve-budget-mib", type=int, default=1024)
    ap.add_argument("--serve-wave-paths", type=int, default=None, help="paths per shared wave (default: tuned)")
    ap.add_argument("--pt-width", type=int, default=1920)
    ap.add_argument("--pt-height", type=int, default=1080)
    ap.add_argument("--pt-spp", type=int, default=64)
    ap.add_argument("--pt-chunk", type=int, default=None, help="wavefront samples per wave (0: about 32M paths; default: tuned)")
    ap.add_argument("--pt-sort", type=int, default=None, help="wavefront: bucket-sort continuation rays (default: tuned)")
    ap.add_argument("--pt-fetch", default=None, help="wavefront traces: slots per claim (0: exact), queue segments (default: tuned)")
    ap.add_argument("--pt-no-mega", action="store_true", help="skip the megakernel comparison")
    ap.add_argument("--pt-no-dsl-wavefront", action="store_true", help="skip the DSL wavefront form")
    a = ap.parse_args()
    # per-call knobs the autotuner chose (luisacompute_amd/autotune_table.json) unless given
    from luisacompute_amd import autotune
    table = autotune.load_table()
    chunk, sort, fetch = autotune.pathtracer_options(table)
    if a.pt_chunk is None:
        a.pt_chunk = chunk
    if a.pt_sort is None:
        a.pt_sort = sort
    if a.pt_fetch is None:
        a.pt_fetch = f"{fetch[0]},{fetch[1]}"
    if a.serve_wave_paths is None:
        a.serve_wave_paths = autotune.serving_options(table)
    return a


def ncores_here():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def bench_pathtracer(a):
    """Box room (6 walls x 512^2 quads + 2 blocks x 6 x 128^2 quads + light: 3,538,946 triangles,
    9 instances), 1920x1080 at 64 spp, max depth 8, Russian roulette from bounce 3. One step = one
    full render. Mrays/s counts every ray traced (camera, bounce and shadow rays)."""
    import luisacompute_amd as lc
    from luisacompute_amd import autotune, host, workloads
    room = workloads.BoxRoom()
    dev = lc.Device(0)
    autotune.load_default(dev)
    s = dev.create_stream()
    accel, meshes, inst = room.build(dev, s)
    scene = room.scene()
    w, h, spp = a.pt_width, a.pt_height, a.pt_spp
    img = dev.create_buffer(16, w * h)
    counters = dev.create_buffer(8, 2)
    s.synchronize()

    def opts(mode, graph=1):
        o = lc.PtOptions()
        o.width, o.height, o.spp, o.sample_offset, o.max_depth, o.rr_depth = w, h, spp, 0, 8, 3
        o.seed, o.sequence, o.mode, o.chunk_spp, o.sort, o.graph = 42, 3 << 32, mode, a.pt_chunk, a.pt_sort, graph
        o.fetch_chunk, o.fetch_segments = (int(x) for x in a.pt_fetch.split(","))
        return o

    def run(mode, reps, warm):
        o = opts(mode)
        for _ in range(warm):
            accel.pathtrace(s, inst, scene, o, img, counters=counters)
        s.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            accel.pathtrace(s, inst, scene, o, img, counters=counters)
        s.synchronize()
        ms = (time.perf_counter() - t0) / reps * 1e3
        cnt = np.zeros(2, np.uint64)
        s << counters.copy_to(cnt) << lc.synchronize()
        return ms, int(cnt[0]), int(cnt[1])

    wf_ms, nc, ns = run(lc.PT_WAVEFRONT, a.steps, a.warmup)
    image = np.zeros((h, w, 4), np.float32)
    s << img.copy_to(image) << lc.synchronize()
    mk_ms, mc, ms_ = (None, nc, ns) if a.pt_no_mega else run(lc.PT_MEGAKERNEL, max(1, a.steps // 2), 1)
    rays = nc + ns
    paths = w * h * spp
    # the same program written in the Python DSL (AST -> HIP source -> hiprtc; one thread per
    # pixel): the AST-codegen megakernel, at the same image size and spp, same image
    dsl_ms = dsl_same = None
    if not a.pt_no_mega:
        from luisacompute_amd import programs
        heap = dev.create_bindless_array(2 * len(meshes))
        for i, m in enumerate(meshes):
            heap.emplace(2 * i, m.vb)
            heap.emplace(2 * i + 1, m.tb)
        s << heap.update()
        xf = host.upload(dev, s, np.concatenate([accel.instance_info(i)[0] for i in range(len(meshes))]).astype(np.float32))
        mats = np.array([list(room.albedo[i]) + [1.0 if room.emissive[i] else 0.0] for i in range(len(meshes))], np.float32)
        mb = host.upload(dev, s, mats)
        vec = lambda name: np.array(getattr(scene, name)[:], np.float32)
        kdsl = programs.path_tracer.compile(dev)
        dimg = dev.create_buffer(16, w * h)
        dsl_cmd = kdsl(dimg, accel, heap, xf, mb, vec("light_position"), vec("light_u"), vec("light_v"),
                       vec("light_emission"), vec("camera_origin"), vec("camera_right"), vec("camera_up"),
                       vec("camera_forward"), spp, 0, 8, 3, 42, 3 << 32)
        s << dsl_cmd.dispatch(w, h)
        s.synchronize()
        t0 = time.perf_counter()
        s << dsl_cmd.dispatch(w, h)
        s.synchronize()
        dsl_ms = (time.perf_counter() - t0) * 1e3
        dimage = np.zeros((h, w, 4), np.float32)
        s << dimg.copy_to(dimage) << lc.synchronize()
        dsl_same = bool((dimage.view(np.uint32) == image.view(np.uint32)).all())
    # the wavefront form written in the DSL (camera / shade / resolve kernels over queues, the
    # library's persistent trace on the device-counted queues; one sample's launches captured
    # into a graph and replayed per sample): same image as the DSL megakernel
    # (queues traced in the order the shade kernel appended them, and sorted on the device by
    # the library's radix sort before each trace)
    dwf_ms = dwf_same = dwfs_ms = dwfs_same = None
    if not a.pt_no_dsl_wavefront:
        from luisacompute_amd import programs
        bind = programs.PathTracerBindings(dev, s, accel, inst, scene)
        for sort_queues in (False, True):
            wfd = programs.WavefrontPathTracer(dev, s, bind, w, h, sort_queues=sort_queues)
            wimg = dev.create_buffer(16, w * h)
            wfd.render(s, wimg, spp, sequence=3 << 32)  # first render: eager sample + capture
            s.synchronize()
            reps = max(1, a.steps // 2)
            t0 = time.perf_counter()
            for _ in range(reps):
                wfd.render(s, wimg, spp, sequence=3 << 32)
            s.synchronize()
            ms = (time.perf_counter() - t0) / reps * 1e3
            wimage = np.zeros((h, w, 4), np.float32)
            s << wimg.copy_to(wimage) << lc.synchronize()
            same = bool((wimage.view(np.uint32) == image.view(np.uint32)).all())
            if sort_queues:
                dwfs_ms, dwfs_same = ms, same
            else:
                dwf_ms, dwf_same = ms, same
            del wfd
    out = {
        "metric": "Mrays/s path tracing, box room 1920x1080 64 spp depth 8",
        "value": round(rays / (wf_ms * 1e-3) / 1e6, 2), "unit": "Mrays/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(wf_ms, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "samples_per_s": round(paths / (wf_ms * 1e-3), 1),
        "rays_closest": nc, "rays_shadow": ns,
        "megakernel_ms": mk_ms and round(mk_ms, 3),
        "megakernel_mrays_s": mk_ms and round((mc + ms_) / (mk_ms * 1e-3) / 1e6, 2),
        "wavefront_over_megakernel": mk_ms and round(mk_ms / wf_ms, 3),
        "same_ray_counts": (mc, ms_) == (nc, ns),
        "dsl_megakernel_ms": dsl_ms and round(dsl_ms, 3),
        "dsl_megakernel_mrays_s": dsl_ms and round(rays / (dsl_ms * 1e-3) / 1e6, 2),
        "dsl_megakernel_same_image": dsl_same,
        "dsl_wavefront_ms": dwf_ms and round(dwf_ms, 3),
        "dsl_wavefront_mrays_s": dwf_ms and round(rays / (dwf_ms * 1e-3) / 1e6, 2),
        "dsl_wavefront_over_dsl_megakernel": (dwf_ms and dsl_ms) and round(dsl_ms / dwf_ms, 3),
        "dsl_wavefront_same_image": dwf_same,
        "dsl_wavefront_sorted_queues_ms": dwfs_ms and round(dwfs_ms, 3),
        "dsl_wavefront_sorted_queues_same_image": dwfs_same,
        "image_mean": float(image[..., :3].mean()),
        "config": {"workload": f"box room {room.triangle_count} tris, {w}x{h}, {spp} spp, depth 8, RR from 3, "
                               f"chunk_spp {a.pt_chunk or 'auto'}, sort {a.pt_sort}, fetch {a.pt_fetch}, hipGraph replay"},
    }
    if not a.no_cpu:
        from oracle import binding as ob
        nco = ncores_here()
        t0 = time.perf_counter()
        om = [ob.Mesh(v, t, nthreads=nco) for v, t in room.meshes]
        sc = ob.Scene()
        for m in om:
            sc.add(m)
        sc.commit()
        t_build = time.perf_counter() - t0
        rows = 8
        t0 = time.perf_counter()
        _, c_c, c_s = ob.pathtrace(sc, room.albedo, room.emissive,
                                   (room.LIGHT_POSITION, room.LIGHT_U, room.LIGHT_V, room.emission), room.camera,
                                   w, h, spp, pixel_range=((h // 2) * w, (h // 2 + rows) * w), nthreads=nco)
        t_pt = time.perf_counter() - t0
        out["cpu_baseline"] = {
            "value": round((c_c + c_s) / t_pt / 1e6, 4), "unit": "Mrays/s", "cores": nco, "kind": "port",
            "sample": f"oracle path tracer, {rows} image rows x {w} px x {spp} spp ({c_c + c_s} rays, {t_pt:.2f} s) "
                      f"on a BVH built in {t_build:.2f} s (build excluded, as for the GPU)"}
    print(json.dumps(out), flush=True)


def serving_measure(a, n):
    """Render serving over four resident scenes (icosphere L6, 16.7M-triangle soup, box room,
    1,024-instance TLAS). Requests: scene U{4}, square resolution U{256,512,1024,2048}, spp
    U{1,4,16,64}, depth 8, PCG32 stream (42, 5). Two disciplines on the same requests:
      plain  : the path tracer written in the DSL (AST -> HIP -> hiprtc megakernel), one dispatch
               per request in arrival order on one stream;
      served : the serving loop (HBM-budget admission, continuous batching into shared wavefront
               waves, scenes on their own streams).
    Throughput is measured with every request queued at t = 0 (requests/s = N / makespan);
    latency (p50/p99 of completion - arrival) with Poisson arrivals at a fixed rate, 0.8 x the
    plain discipline's saturated throughput. Returns the result dict."""
    import threading
    import luisacompute_amd as lc
    from luisacompute_amd import autotune, programs, workloads
    dev = lc.Device(0)
    autotune.load_default(dev)
    s = dev.create_stream()
    t0 = time.perf_counter()
    scenes = workloads.serving_scenes(dev, s)
    binds = [programs.PathTracerBindings(dev, s, acc, inst, sc) for _, acc, inst, sc in scenes]
    kdsl = programs.path_tracer.compile(dev)
    s.synchronize()
    t_setup = time.perf_counter() - t0
    base = workloads.serving_requests(n, 0.0)
    bufs = [dev.create_buffer(16, r["res"] * r["res"]) for r in base]

    def dispatch_plain(r):
        cmd = kdsl(*binds[r["scene"]].args(bufs[r["index"]], r["spp"], 0, 8, 3, 42, r["sequence"]))
        s << cmd.dispatch(r["res"], r["res"])

    def plain(reqs, rate_mode):
        lat = []
        t_start = time.perf_counter()
        for r in reqs:
            if rate_mode:
                wait = t_start + r["arrival"] - time.perf_counter()
                if wait > 0:
                    time.sleep(wait)
            dispatch_plain(r)
            s.synchronize()
            lat.append(time.perf_counter() - t_start - r["arrival"])
        return time.perf_counter() - t_start, lat

    fetch = tuple(int(x) for x in a.pt_fetch.split(","))

    def served(reqs, rate_mode):
        srv = lc.Server(dev, hbm_budget=a.serve_budget_mib << 20, wave_paths=a.serve_wave_paths, max_active=64,
                        sort=a.pt_sort, graph=1, fetch=fetch)
        ids = [srv.add_scene(acc, inst, sc) for _, acc, inst, sc in scenes]
        # warm: capture each scene's wave graph
        for sid in ids:
            srv.submit(sid, 16, 16, 1, bufs[0], seed=1, sequence=1)
        srv.drain()
        st0 = srv.stats()
        rid = {}
        t_start = time.perf_counter()

        def submitter():
            for r in reqs:
                if rate_mode:
                    wait = t_start + r["arrival"] - time.perf_counter()
                    if wait > 0:
                        time.sleep(wait)
                rid[r["index"]] = srv.submit(ids[r["scene"]], r["res"], r["res"], r["spp"], bufs[r["index"]], seed=42,
                                             sequence=r["sequence"])
        th = threading.Thread(target=submitter)
        th.start()
        th.join()
        srv.drain(timeout=1200)
        makespan = time.perf_counter() - t_start
        lat = []
        for r in reqs:
            q = srv.query(rid[r["index"]])
            lat.append(q.t_done - q.t_submit)
        st = srv.stats()
        rays = (st.rays_closest + st.rays_shadow) - (st0.rays_closest + st0.rays_shadow)
        out = dict(waves=st.waves - st0.waves, peak_active=st.peak_active, peak_accumulator_mib=st.peak_bytes / 2 ** 20)
        srv.close()
        return makespan, lat, rays, out

    # warm the plain path (module load) on a tiny request per scene
    for k in range(len(scenes)):
        dispatch_plain({"scene": k, "index": 0, "spp": 1, "sequence": 1, "res": 16})
    s.synchronize()
    print(f"[serving] setup {t_setup:.1f} s; {n} requests; plain backlog", file=sys.stderr, flush=True)
    mk_plain, _ = plain(base, False)
    print(f"[serving] plain backlog {mk_plain:.2f} s; served backlog", file=sys.stderr, flush=True)
    mk_served, _, rays_served, info_b = served(base, False)
    rate = a.serve_rate or 0.8 * n / mk_plain
    timed = workloads.serving_requests(n, rate)
    print(f"[serving] served backlog {mk_served:.2f} s; rate {rate:.2f}/s plain", file=sys.stderr, flush=True)
    _, lat_plain = plain(timed, True)
    print("[serving] served at rate", file=sys.stderr, flush=True)
    _, lat_served, _, info_r = served(timed, True)

    def pct(x, q):
        return round(float(np.percentile(np.asarray(x), q)) * 1e3, 2)

    return {
        "requests": n,
        "served_requests_s": round(n / mk_served, 3), "plain_requests_s": round(n / mk_plain, 3),
        "served_over_plain": round(mk_plain / mk_served, 3),
        # both disciplines trace the same paths (each served image equals its alone render)
        "served_mrays_s": round(rays_served / mk_served / 1e6, 1), "plain_mrays_s": round(rays_served / mk_plain / 1e6, 1),
        "rate_per_s": round(rate, 3),
        "latency_ms": {"served_p50": pct(lat_served, 50), "served_p99": pct(lat_served, 99),
                       "plain_p50": pct(lat_plain, 50), "plain_p99": pct(lat_plain, 99)},
        "plain": "DSL path tracer (AST -> HIP -> hiprtc megakernel), one dispatch per request",
        "served_backlog": info_b, "served_at_rate": info_r,
        "workload": f"{n} requests over 4 resident scenes (icosphere L6, soup 16.7M, box room 3.5M, "
                    f"1,024-instance TLAS); res U{{256..2048}}^2, spp U{{1,4,16,64}}, depth 8; "
                    f"budget {a.serve_budget_mib} MiB, wave {a.serve_wave_paths} paths",
    }


def bench_serving(a):
    """The serving measurement alone as its own line (requests/s is the value)."""
    r = serving_measure(a, a.serve_requests)
    out = {
        "metric": "served renders: requests/s (saturated) and p50/p99 latency at a fixed Poisson rate",
        "value": r["served_requests_s"], "unit": "requests/s", "n_gpus": 1, "steps": r["requests"], "warmup": 1,
        "ms_per_step": round(1e3 / r["served_requests_s"], 2), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": r.pop("workload")},
    }
    out.update(r)
    print(json.dumps(out), flush=True)


def bench_instanced(a):
    """1,024 instances of 8 meshes, 64 closest-hit batches (log-uniform 4,096..4,194,304 rays):
    one persistent launch for all batches vs one launch per batch."""
    import luisacompute_amd as lc
    from luisacompute_amd import autotune, host, workloads
    w = workloads.Instanced()
    dev = lc.Device(0)
    autotune.load_default(dev)
    s = dev.create_stream()
    meshes = [host.mesh_from_arrays(dev, s, v, t) for v, t in w.meshes]
    accel = dev.create_accel()
    for aff, k in zip(w.affines, w.mesh_ids):
        accel.emplace_back(meshes[int(k)], aff)
    s << accel.build()
    bufs = []
    for j in range(len(w.batch_sizes)):
        r = w.batch_rays(j)
        bufs.append((host.upload(dev, s, r), dev.create_buffer(24, len(r)), len(r)))
    s.synchronize()
    total = w.total_rays

    def timed(fn, reps):
        fn()
        s.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        s.synchronize()
        return (time.perf_counter() - t0) / reps * 1e3

    one = timed(lambda: accel.trace_batched(s, bufs), a.steps)
    per = timed(lambda: [accel.trace_closest(s, rb, hb, n) for rb, hb, n in bufs], a.steps)
    st = accel.stats()
    print(json.dumps({
        "metric": "Mrays/s instanced closest-hit, 64 batches",
        "value": round(total / (one * 1e-3) / 1e6, 2), "unit": "Mrays/s", "n_gpus": 1, "steps": a.steps,
        "ms_per_step": round(one, 3), "per_batch_launches_ms": round(per, 3),
        "per_batch_launches_mrays_s": round(total / (per * 1e-3) / 1e6, 2),
        "rays": total, "batches": len(w.batch_sizes), "instances": len(w.mesh_ids),
        "unique_triangles": int(sum(len(t) for _, t in w.meshes)), "tlas_depth": st.depth,
        "config": {"workload": "1,024 instances of icospheres L3-L8 + soups 200k/500k; 64 batches log-uniform "
                               "4,096..4,194,304 rays, origin U[-60,60]^3, dir U(S^2)"},
        "dtype": "f32", "data": "synthetic"}), flush=True)


def bench_query(a):
    """Ray queries (accel.traverse / traverse_any) on the reference's procedural test scene at scale:
    a non-opaque triangle soup with the rounded-cutout surface callback plus procedural spheres
    inscribed in AABBs, src/tests/test_procedural.cpp:29-120."""
    import luisacompute_amd as lc
    from luisacompute_amd import autotune, host
    n_tris, n_boxes, n_rays = 1 << 20, 1 << 20, a.rays
    verts, tris = lc.soup(n_tris, initseq=5)
    rng = np.random.default_rng(7)
    c = rng.uniform(0.0, 100.0, (n_boxes, 3))
    boxes = np.concatenate([c - 0.2, c + 0.2], axis=1).astype(np.float32)
    rays = lc.rays_box_sphere(n_rays, initseq=6)
    dev = lc.Device(0)
    autotune.load_default(dev)
    s = dev.create_stream()
    mesh = host.mesh_from_arrays(dev, s, verts, tris)
    prim = host.procedural_from_aabbs(dev, s, boxes)
    accel = dev.create_accel()
    accel.emplace_back(mesh, lc.IDENTITY, opaque=False)
    accel.emplace_back(prim, lc.IDENTITY)
    s << accel.build()
    rb = host.upload(dev, s, rays)
    hb = dev.create_buffer(24, n_rays)
    s.synchronize()
    res = {}
    for name, first in (("query", False), ("query_any", True)):
        def run():
            accel.query(s, rb, hb, n_rays, surface=lc.QUERY_SURFACE_CUTOUT, cutout=0.8,
                        procedural=lc.QUERY_PROCEDURAL_SPHERE, terminate_on_first=first)
        for _ in range(a.warmup):
            run()
        s.synchronize()
        dev.reset_timing()
        dev.enable_timing(True)
        for _ in range(a.steps):
            run()
        s.synchronize()
        dev.enable_timing(False)
        ms, n = dev.timing(name)
        res[name] = ms / max(n, 1)
    accel.query(s, rb, hb, n_rays, surface=lc.QUERY_SURFACE_CUTOUT, cutout=0.8,
                procedural=lc.QUERY_PROCEDURAL_SPHERE, terminate_on_first=False)
    out = np.zeros(n_rays, lc.COMMITTED_HIT_DTYPE)
    s << hb.copy_to(out) << lc.synchronize()
    kinds = np.bincount(out["hit_type"].astype(np.int64), minlength=3)
    # the same query as a run-time shader (hiprtc, one query per thread, callbacks in the shader)
    from luisacompute_amd import programs
    sh = dev.create_shader(programs.QUERY, "query_kernel")
    hs = dev.create_buffer(24, n_rays)

    def run_shader():
        s << sh(rb, hs, accel, np.float32(0.8), np.uint32(0)).dispatch(n_rays)
    run_shader()
    s.synchronize()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        run_shader()
    s.synchronize()
    shader_ms = (time.perf_counter() - t0) * 1e3 / a.steps
    out_sh = np.zeros(n_rays, lc.COMMITTED_HIT_DTYPE)
    s << hs.copy_to(out_sh) << lc.synchronize()
    shader_same = bool((out_sh.view(np.uint32) == out.view(np.uint32)).all())
    # the same rays as plain closest hits (procedural instance unseen, every triangle committed)
    hc = dev.create_buffer(24, n_rays)
    dev.reset_timing()
    dev.enable_timing(True)
    for _ in range(a.steps):
        accel.trace_closest(s, rb, hc, n_rays)
    s.synchronize()
    dev.enable_timing(False)
    tms, nt = dev.timing("trace_closest")
    tms /= max(nt, 1)
    print(json.dumps({
        "metric": "Mrays/s ray queries (traverse, surface + procedural callbacks)",
        "value": round(n_rays / (res["query"] * 1e-3) / 1e6, 2), "unit": "Mrays/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": round(res["query"], 3),
        "query_any_ms": round(res["query_any"], 3),
        "query_any_mrays_s": round(n_rays / (res["query_any"] * 1e-3) / 1e6, 2),
        "trace_closest_same_rays_ms": round(tms, 3),
        "trace_closest_same_rays_mrays_s": round(n_rays / (tms * 1e-3) / 1e6, 2),
        "committed": {"miss": int(kinds[0]), "surface": int(kinds[1]), "procedural": int(kinds[2])},
        "shader_query_ms": round(shader_ms, 3),
        "shader_query_mrays_s": round(n_rays / (shader_ms * 1e-3) / 1e6, 2),
        "shader_query_same_answers": shader_same,
        "config": {"workload": f"non-opaque soup {n_tris:,} tris (U[0,100]^3, sigma 0.5) with the rounded cutout "
                               f"(0.8) + {n_boxes:,} AABBs (half extent 0.2) with inscribed spheres; "
                               f"{n_rays:,} rays origin U[0,100]^3 dir U(S^2)"},
        "dtype": "f32", "data": "synthetic"}), flush=True)


def bench_build(a):
    """BLAS builds of the soup (16,777,216 triangles) under each hint: FAST_BUILD (LBVH, 3-triangle
    leaves), FAST_TRACE without and with the binned-SAH refinement; `steps` timed builds each."""
    import luisacompute_amd as lc
    from luisacompute_amd import host
    verts, tris = lc.soup(a.tris)
    dev = lc.Device(0)
    s = dev.create_stream()
    res = {}
    for name, hint, sah in (("fast_build", lc.AccelOption.FAST_BUILD, True),
                            ("fast_trace_lbvh", lc.AccelOption.FAST_TRACE, False),
                            ("fast_trace_sah", lc.AccelOption.FAST_TRACE, True)):
        dev.set_build_sah(sah)
        m = host.mesh_from_arrays(dev, s, verts, tris, build=False, option=lc.AccelOption(hint, 1, 0))
        for _ in range(max(1, a.warmup)):
            s << m.build()
        s.synchronize()
        dev.reset_timing()
        dev.enable_timing(True)
        for _ in range(a.steps):
            s << m.build()
        s.synchronize()
        dev.enable_timing(False)
        ms, n = dev.timing("mesh_build")
        st = m.stats()
        res[name] = {"build_ms": round(ms / max(n, 1), 3), "wide_nodes": st.node_count, "depth": st.depth}
        del m
    dev.set_build_sah(True)
    print(json.dumps({"metric": "BLAS build ms, soup 16.7M triangles, per hint", "value": res["fast_build"]["build_ms"],
                      "unit": "ms", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup, "higher_is_better": False,
                      "hints": res, "config": {"workload": f"soup {a.tris:,} tris (U[0,100]^3, sigma 0.5)"},
                      "dtype": "f32", "data": "synthetic"}), flush=True)


def copy_probe(mib):
    """Plain copy of known size: reading FETCH_SIZE / WRITE_SIZE for it calibrates the counters."""
    import luisacompute_amd as lc
    dev = lc.Device(0)
    s = dev.create_stream()
    n = mib << 20
    a, b = dev.create_buffer(1, n), dev.create_buffer(1, n)
    for _ in range(3):
        s << b.copy_from_buffer(a)
    s.synchronize()
    print(json.dumps({"copy_probe_bytes": n, "copies": 3}), flush=True)


def cpu_baseline(verts, tris, rays_f, sample_rays, total_rays):
    """The CPU oracle (C++ restatement of the reference CPU backend's semantics, std::thread
    pool) on a bounded sample: full BVH build over all triangles + closest-hit of a ray sample;
    the job time is estimated as build + sample trace time scaled to all rays."""
    from oracle import binding as ob
    ncores = os.cpu_count() or 1
    try:
        ncores = len(os.sched_getaffinity(0))
    except AttributeError:
        pass
    t0 = time.perf_counter()
    mesh = ob.Mesh(verts, tris, nthreads=ncores)
    scene = ob.Scene().add(mesh).commit()
    t_build = time.perf_counter() - t0
    sample = np.ascontiguousarray(rays_f[:sample_rays])
    t0 = time.perf_counter()
    scene.trace_closest(sample, nthreads=ncores)
    t_trace = time.perf_counter() - t0
    est = t_build + t_trace * (total_rays / sample_rays)
    return {
        "value": round(total_rays / est / 1e6, 4),
        "unit": "Mrays/s",
        "cores": ncores,
        "cpu_model": cpu_model(),
        "kind": "port",
        "sample": (f"full {len(tris)}-triangle BVH build ({t_build:.2f} s) + closest-hit of the first "
                   f"{sample_rays} rays ({t_trace:.2f} s, {sample_rays / t_trace / 1e6:.3f} Mrays/s); "
                   f"job time estimated as build + trace scaled to {total_rays} rays"),
    }


def build_hints(dev, s, mesh, rb, n_rays):
    """The soup's BLAS under each build hint: build ms, wide nodes, node visits and triangle
    tests per ray and closest-hit ms. FAST_BUILD (LBVH, 3-triangle leaves) is the timed step's;
    FAST_TRACE (single-triangle leaves) with and without the SAH refinement of its clusters."""
    import luisacompute_amd as lc
    out = {}
    for name, hint, sah in (("fast_build", lc.AccelOption.FAST_BUILD, True),
                            ("fast_trace_lbvh", lc.AccelOption.FAST_TRACE, False),
                            ("fast_trace_sah", lc.AccelOption.FAST_TRACE, True)):
        dev.set_build_sah(sah)
        m = dev.create_mesh(mesh.vb, mesh.tb, option=lc.AccelOption(hint, 1, 0))
        acc = dev.create_accel()
        acc.emplace_back(m, lc.IDENTITY)
        s << m.build() << acc.build()
        s.synchronize()
        dev.reset_timing()
        dev.enable_timing(True)
        for _ in range(3):
            s << m.build()
        s << acc.build()
        for _ in range(3):
            acc.trace_closest(s, rb, _hits(dev, n_rays), n_rays)
        s.synchronize()
        dev.enable_timing(False)
        bms, nb = dev.timing("mesh_build")
        tms, nt = dev.timing("trace_closest")
        nn, ntr = acc.trace_count(s, rb, n_rays)
        st = m.stats()
        out[name] = {"build_ms": round(bms / max(nb, 1), 3), "trace_closest_ms": round(tms / max(nt, 1), 3),
                     "wide_nodes": st.node_count, "depth": st.depth,
                     "nodes_per_ray": round(nn / n_rays, 3), "tris_per_ray": round(ntr / n_rays, 3)}
        del acc, m
    dev.set_build_sah(True)
    return out


_HITS = {}


def _hits(dev, n):
    if n not in _HITS:
        _HITS[n] = dev.create_buffer(24, n)
    return _HITS[n]


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def main():
    a = parse()
    if a.copy_probe_mib:
        copy_probe(a.copy_probe_mib)
        return
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported (rays and scenes would shard "
              f"by replicating the BVH per GPU)", file=sys.stderr)
        sys.exit(2)
    if a.workload == "instanced":
        bench_instanced(a)
        return
    if a.workload == "pathtracer":
        bench_pathtracer(a)
        return
    if a.workload == "serving":
        bench_serving(a)
        return
    if a.workload == "query":
        bench_query(a)
        return
    if a.workload == "build":
        bench_build(a)
        return
    import luisacompute_amd as lc
    from luisacompute_amd import host

    t0 = time.perf_counter()
    verts, tris = lc.soup(a.tris)
    rays = lc.rays_box_sphere(a.rays, skip=13 * a.tris)
    gen_s = time.perf_counter() - t0

    dev = lc.Device(0)
    from luisacompute_amd import autotune
    tuned = autotune.load_default(dev)  # block size / LDS stack / refill steps from the autotuner
    s = dev.create_stream()
    # the build is inside every step: the BLAS is built with the FAST_BUILD hint (3 triangles per leaf)
    build_opt = lc.AccelOption(lc.AccelOption.FAST_BUILD, 1, 0)
    mesh = host.mesh_from_arrays(dev, s, verts, tris, build=False, option=build_opt)
    accel = dev.create_accel()
    accel.emplace_back(mesh, lc.IDENTITY)
    rb = host.upload(dev, s, rays)
    hb = dev.create_buffer(24, a.rays)
    hab = dev.create_buffer(1, a.rays)
    s.synchronize()

    def step():
        s << mesh.build() << accel.build()
        accel.trace_closest(s, rb, hb, a.rays)

    for _ in range(a.warmup):
        step()
    s.synchronize()
    dev.reset_timing()
    dev.enable_timing(True)
    t0 = time.perf_counter()
    for _ in range(a.steps):
        step()
    s.synchronize()
    wall = time.perf_counter() - t0
    dev.enable_timing(False)
    build_ms, nb = dev.timing("mesh_build")
    trace_ms, nt = dev.timing("trace_closest")          # entry point: ray sort + traversal
    kern_ms, nk = dev.timing("trace_closest_kernel")    # traversal kernel alone (roofline)
    build_ms /= max(nb, 1)
    trace_ms /= max(nt, 1)
    kern_ms /= max(nk, 1)
    ms_per_step = wall * 1e3 / a.steps

    # any-hit on the same BVH (reported beside the headline)
    dev.reset_timing()
    dev.enable_timing(True)
    for _ in range(3):
        accel.trace_any(s, rb, hab, a.rays)
    s.synchronize()
    dev.enable_timing(False)
    any_ms, na = dev.timing("trace_any")
    any_ms /= max(na, 1)

    # BLAS update (refit) of the same triangles on a mesh that allows updates, reported beside
    # the full build
    umesh = dev.create_mesh(mesh.vb, mesh.tb, option=lc.AccelOption(lc.AccelOption.FAST_BUILD, 1, 1))
    s << umesh.build()
    s.synchronize()
    dev.reset_timing()
    dev.enable_timing(True)
    for _ in range(3):
        s << umesh.build()
    s.synchronize()
    dev.enable_timing(False)
    refit_ms, nr = dev.timing("mesh_refit")
    refit_ms /= max(nr, 1)
    umem = umesh.memory()
    del umesh

    # algorithmic bytes: counter-mode run of the same traversal (same sorted queue, grid and
    # wave-uniform steps) on the same BVH and rays
    steps = accel.trace_steps(s, rb, a.rays)
    n_nodes, n_tris = steps.nodes_visited, steps.triangles_tested
    lane_slots = 64 * max(steps.wave_steps, 1)
    st = mesh.stats()
    mem = mesh.memory()
    nodes_per_ray = n_nodes / a.rays
    tris_per_ray = n_tris / a.rays
    bytes_per_ray = RAY_BYTES + HIT_BYTES + NODE_BYTES * nodes_per_ray + TRI_BYTES * tris_per_ray
    achieved = a.rays * bytes_per_ray / (kern_ms * 1e-3) / 1e9
    floor_gbs = (a.rays * (RAY_BYTES + HIT_BYTES) + st.node_bytes + st.primitive_bytes) / (kern_ms * 1e-3) / 1e9

    # HBM bytes per launch of the traversal kernel: FETCH_SIZE + WRITE_SIZE from two rocprofv3
    # --pmc runs of this command, each scaled by what the counter reads for a 1 GiB device copy
    # (profiles/trace_closest_soup16m_traffic.md)
    traffic = l2_hit = None
    tpath = os.path.join(ROOT, "profiles", "trace_closest_soup16m_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as f:
            tj = json.load(f)
        traffic, l2_hit = tj.get("hbm_bytes_per_launch"), tj.get("l2_hit_rate")

    hints = build_hints(dev, s, mesh, rb, a.rays)

    result = {
        "metric": "Mrays/s incoherent closest-hit @16M tris (BVH build included)",
        "value": round(a.rays / (ms_per_step * 1e-3) / 1e6, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "triangle soup 16,777,216 tris (centroid U[0,100]^3, sigma 0.5) + 8,388,608 rays "
                        "(origin U[0,100]^3, dir U(S^2)); step = BLAS build (FAST_BUILD hint) + TLAS + trace_closest",
            "triangles": a.tris,
            "rays": a.rays,
            "seed": "pcg32(42, 2)",
        },
        "build_ms": round(build_ms, 3),
        "build_mtris_s": round(a.tris / (build_ms * 1e-3) / 1e6, 2),
        # per triangle: 60 B indices + vertices read, 8 B key/id written, 4 radix passes of 16 B,
        # 48 B record written; plus 80 B per wide node
        "build_algorithmic_gbs": round((a.tris * (60 + 8 + 4 * 16 + 48) + 80 * st.node_count) / (build_ms * 1e-3) / 1e9, 1),
        "refit_ms": round(refit_ms, 3),
        "trace_closest_ms": round(trace_ms, 3),
        "trace_closest_kernel_ms": round(kern_ms, 3),
        "trace_closest_mrays_s": round(a.rays / (trace_ms * 1e-3) / 1e6, 2),
        "trace_any_ms": round(any_ms, 3),
        "trace_any_mrays_s": round(a.rays / (any_ms * 1e-3) / 1e6, 2),
        "bvh": {"wide_nodes": st.node_count, "depth": st.depth, "node_bytes": st.node_bytes,
                "tri_bytes": st.primitive_bytes, "held_bytes": mem["total_bytes"],
                "held_bytes_allow_update": umem["total_bytes"], "build_arena_bytes": dev.build_arena_bytes()},
        "nodes_per_ray": round(nodes_per_ray, 3),
        "tris_per_ray": round(tris_per_ray, 3),
        "roofline": {
            "bound": "hbm",
            "kernel": "k_trace<closest> (persistent wave64 traversal, sorted ray queue)",
            "kernel_ms": round(kern_ms, 3),
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "l2_hit_rate": l2_hit,
            "traffic_source": "profiles/trace_closest_soup16m_traffic.json: FETCH_SIZE + WRITE_SIZE of k_trace<closest> "
                              "from rocprofv3 --pmc runs of this command, scaled by a 1 GiB copy's counts"
                              if traffic is not None else None,
            # the HBM roofline is the one priced here; the counter profile shows the kernel is
            # limited by VALU issue (divergent lanes), not by bandwidth
            "limiter": "valu_issue",
            # lane accounting of the counting run: of the wave iterations' lane slots, the share
            # with a ray, and the shares taking a node visit / a triangle test
            "lane_steps": {"live": round(steps.live_lane_steps / lane_slots, 4),
                           "node_visits": round(n_nodes / lane_slots, 4),
                           "triangle_tests": round(n_tris / lane_slots, 4),
                           "wave_iterations_per_ray": round(steps.wave_steps * 64 / a.rays, 2),
                           "refill_rounds": steps.refill_rounds},
            "bytes_per_ray": round(bytes_per_ray, 2),
            "compulsory_floor_gbs": round(floor_gbs, 2),
        },
        "input_gen_s": round(gen_s, 2),
        "trace_config": {"closest": dev.trace_config("closest"), "any": dev.trace_config("any"),
                         "source": "luisacompute_amd/autotune_table.json" if tuned else "defaults"},
        "build_hints": hints,
    }
    if not a.no_cpu:
        result["cpu_baseline"] = cpu_baseline(verts, tris, lc.rays_as_floats(rays), a.cpu_sample_rays, a.rays)
    if not a.no_serving:
        # second half of the metric: served renders (requests/s, p50/p99) beside the plain
        # per-request DSL megakernel, on the four resident scenes
        del mesh, accel, rb, hb, hab
        dev.release_build_arena()
        result["serving"] = serving_measure(a, a.serve_requests)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
