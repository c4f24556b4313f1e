#!/usr/bin/env python3
"""Path-tracer benchmark: Mpaths/s at 1920x1080, 64 spp, on one MI355X.

Workload (the single-large-scene configuration): a seeded fBm heightfield
of 1024x512 quads (1,048,576 triangles) plus 200 subdivision-3 icospheres
(256,000 triangles) and 64 emissive triangles, glTF metallic-roughness
materials (metallic ~ Bernoulli(0.3), roughness ~ U(0.05, 1)), a
procedural 1024x512 sky with its sampling CDFs, 1920x1080, 64 samples per
pixel, 12 bounces, roulette from bounce 3.

A step is one full frame: reset + refine(64) + finish, the scene resident
in HBM. `value` = frame paths x steps / wall time. Reported beside it:
per-frame latency p50/p99, the LDR PSNR against the CPU oracle on 16 seeded
64x64 tiles at the same 64 spp, the trace kernel's roofline (algorithmic
bytes per path counted by the instrumented oracle, divided by the kernel's
HIP-event time), and the oracle's own Mpaths/s on those tiles.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "oxideav-render-vulkan_amd"))

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def psnr(a, b):
    d = a[..., :3].astype(np.float64) - b[..., :3].astype(np.float64)
    mse = float((d * d).mean())
    return float("inf") if mse == 0 else 10.0 * np.log10(255.0 ** 2 / mse)


def workload(name):
    import oxpt

    if name == "c1":
        scene = oxpt.Scene.terrain(seed=1, gx=1024, gz=512, n_spheres=200, subdiv=3, n_emissive=64)
        env = oxpt.EnvironmentMap.sky(1024, 512, 0)
        opts = oxpt.RenderOptions(width=1920, height=1080, samples_per_pixel=64, max_bounces=12, rr_start=3)
        desc = ("single large scene: fBm heightfield 1024x512 quads + 200 icospheres (subdiv 3) + 64 emissive "
                "tris, MR materials, 1024x512 sky CDF; 1920x1080, 64 spp, depth 12")
    elif name == "c2":
        scene = oxpt.Scene.layered(seed=2, n_tris=4_000_000, n_textures=16, tex_size=2048, n_lights=16)
        env = None
        opts = oxpt.RenderOptions(width=1920, height=1080, samples_per_pixel=256, max_bounces=16, rr_start=3)
        desc = ("layered/texture scene: 4 M tris baked from instanced meshes, materials 30% transmission+volume, "
                "20% clearcoat, 20% sheen, 10% alpha BLEND, 20% MR; 16 value-noise RGBA 2048^2 textures with "
                "mips in the f16 atlas; 16 punctual lights; 1920x1080, 256 spp, depth 16")
    elif name == "c0":
        scene = oxpt.Scene.room(0)
        env = None
        opts = oxpt.RenderOptions(width=256, height=256, samples_per_pixel=16, max_bounces=8, rr_start=3)
        desc = "closed room: 5 quads + 2 rotated boxes + emissive quad (36 tris); 256x256, 16 spp, depth 8"
    else:
        raise SystemExit(f"unknown workload {name}")
    return scene, env, opts, desc


# Committed PMC readings per workload: file and the default form they were taken on.
PMC_FILES = {
    "c1": ("trace_persistent_c1_pmc.json", "trace_persistent<6,early,opaque>"),  # no punctual lights
    "c2": ("trace_persistent_c2_pmc.json", "trace_persistent<4,queue>"),  # 16 punctual lights, alpha blending
    "c0": ("trace_batch_c0_pmc.json", "trace_batch<3,early,opaque>"),  # the one-job batch of a small frame
}


def pmc_traffic(args, paths_per_launch):
    """HBM bytes per launch of the default kernel on the single-scene
    workloads, from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE readings
    of that launch, each scaled by its reading on a plain copy of known size
    (profiles/counter_calibration_copy_1GiB.md). None for other configs."""
    if (args.workload not in PMC_FILES or args.kernel != 0 or args.spp or args.block or args.waves_per_cu
            or args.tail or args.early_off):
        return None
    name, kernel = PMC_FILES[args.workload]
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    if d.get("kernel") != kernel:
        return None
    per_path = (d["fetch_kib"] * d["fetch_scale"] + d["write_kib"] * d["write_scale"]) * 1024.0 / d["paths_per_dispatch"]
    return int(round(per_path * paths_per_launch))


def cpu_threads():
    """Host threads for the CPU baseline: every CPU this process may run on,
    at most 16 (the GPU box's CPU share per GPU)."""
    try:
        n = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def pmc_measured(args):
    """Counter readings of the default kernel on this workload, from the
    committed rocprofv3 PMC summary (None when absent or for other forms):
    HBM bytes per path, VALU busy and lane utilisation, L2 hit rate, vector
    store bytes per path, and the limiter they show."""
    if (args.workload not in PMC_FILES or args.kernel != 0 or args.spp or args.block or args.waves_per_cu
            or args.tail or args.early_off):
        return None
    name, kernel = PMC_FILES[args.workload]
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        d = json.load(f)
    if d.get("kernel") != kernel:
        return None
    out = dict(d.get("derived", {}))
    out["source"] = "profiles/" + name
    return out


def tiles(w, h, n, size=64, seed=1234):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        x0 = int(rng.integers(0, max(w - size, 0) + 1))
        y0 = int(rng.integers(0, max(h - size, 0) + 1))
        out.append((x0, y0, min(x0 + size, w), min(y0 + size, h)))
    return out


def bench_uneven_batch(args):
    """64 scenes of uneven size: per-request launches (the plain per-pixel
    kernel; the regenerating persistent megakernel, forced for every frame)
    vs one batched refine under the batch autotuner's choice, eager and
    graph-replayed; per-job latency, PSNR against the oracle on tiles of the
    smallest / median / largest jobs, and the oracle's own rate."""
    import oxpt
    from oxpt import workloads

    t = time.perf_counter()
    jobs = workloads.uneven_batch(n=64, seed=3, threads=min(16, os.cpu_count() or 1))
    log(f"64 scenes generated + packed in {time.perf_counter() - t:.1f} s "
        f"({sum(j.n_tris for j in jobs)} tris, {sum(j.paths for j in jobs) / 1e6:.1f} M paths)")
    dev = oxpt.Device(0)
    log("device:", dev.summary())
    ds = [dev.upload(j.packed) for j in jobs]
    acc = [dev.accum(j.opts.width, j.opts.height) for j in jobs]
    total_paths = sum(j.paths for j in jobs)
    items = [(s, a, s.params, 0, j.opts.samples_per_pixel) for j, s, a in zip(jobs, ds, acc)]

    def per_request(kernel):
        def run():
            dev.set_tuning(kernel=kernel, max_paths=0xFFFFFFFF)
            done = []
            t0 = time.perf_counter()
            for j, s, a in zip(jobs, ds, acc):
                dev.refine(s, a, s.params, 0, j.opts.samples_per_pixel)
                dev.finish()
                done.append(time.perf_counter() - t0)
            return done
        return run

    tuned = None
    if args.kernel == 0:
        dev.set_tuning(kernel=0, max_paths=0xFFFFFFFF)
        # 64-spp trials: at 16 the jobs' work mix (most paths in the 256-spp
        # jobs) is not the full run's, and the trials picked the 5-wave form
        # where the full run is faster at 6.
        tuned, tune_ms = dev.autotune_batch(items, spp_per_trial=64)
        log(f"autotune_batch: kernel {tuned.kernel}, max_paths {tuned.max_paths}, stack_lds {tuned.stack_lds} "
            f"({tune_ms:.1f} ms per 64-spp trial)")

    def batched(graph):
        def run():
            if tuned is not None:
                dev.set_tuning(kernel=tuned.kernel, max_paths=tuned.max_paths, stack_lds=tuned.stack_lds)
            else:
                dev.set_tuning(kernel=args.kernel, max_paths=0xFFFFFFFF)
            dev.set_graph(graph)
            t0 = time.perf_counter()
            dev.refine_batch(items)
            dev.finish()
            dev.set_graph(False)
            return [time.perf_counter() - t0] * len(jobs)
        return run

    res = {}
    for name, fn in (("plain", per_request(1)), ("per_request_persistent", per_request(4)),
                     ("batched", batched(False)), ("batched_graph", batched(True))):
        for _ in range(args.warmup):
            for a in acc:
                a.clear()
            fn()
        times, lat = [], []
        for i in range(args.steps):
            for a in acc:
                a.clear()
            dev.finish()
            t0 = time.perf_counter()
            done = fn()
            times.append(time.perf_counter() - t0)
            lat.extend(done)
            log(f"{name} {i + 1}/{args.steps}: {times[-1] * 1e3:.1f} ms")
        res[name] = {"s": sorted(times)[len(times) // 2], "lat": lat, "kernel": dev.last_kernel(),
                     "images": [a.read() for a in acc] if name in ("plain", "batched") else None}
    same = all(np_equal(a, b) for a, b in zip(res["batched"]["images"], res["plain"]["images"]))
    same_graph = all(np_equal(a.read(), b) for a, b in zip(acc, res["plain"]["images"]))

    def lat_ms(name):
        xs = sorted(x * 1e3 for x in res[name]["lat"])
        return {"p50": round(xs[len(xs) // 2], 3), "p99": round(xs[min(len(xs) - 1, int(0.99 * (len(xs) - 1) + 0.5))], 3)}

    # Oracle: PSNR on two 64x64 tiles of the smallest, median and largest
    # jobs at full spp, and its rate on those tiles.
    psnr_by_job, cpu_paths, cpu_s = {}, 0, 0.0
    threads = cpu_threads()
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle

        order = sorted(range(len(jobs)), key=lambda i: jobs[i].paths)
        for name, i in (("smallest", order[0]), ("median", order[len(order) // 2]), ("largest", order[-1])):
            j = jobs[i]
            w, h, spp = j.opts.width, j.opts.height, j.opts.samples_per_pixel
            osc = oracle.Scene.from_packed(j.packed)
            tl = tiles(w, h, 2, seed=100 + i)
            cacc = np.zeros((h, w, 4), np.float32)
            tc = time.perf_counter()
            for r in tl:
                osc.render(0, spp, region=r, accum=cacc, threads=threads)
            cpu_s += time.perf_counter() - tc
            cpu_paths += sum((r[2] - r[0]) * (r[3] - r[1]) for r in tl) * spp
            rp = oxpt.ResolveParams((1, 3, 0, 0), (1.0, float(spp), float(w), 0.0),
                                    tuple(c / 255.0 for c in j.opts.background))
            gi, ci = oracle.resolve(res["batched"]["images"][i], rp), oracle.resolve(cacc, rp)
            gt = np.concatenate([gi[r[1]:r[3], r[0]:r[2]].reshape(-1, 4) for r in tl])
            ct = np.concatenate([ci[r[1]:r[3], r[0]:r[2]].reshape(-1, 4) for r in tl])
            psnr_by_job[f"{name} (job {i}: {j.n_tris} tris, {w}x{h}, {spp} spp)"] = round(psnr(gt, ct), 2)
            log(f"oracle {name} job {i}: PSNR {psnr_by_job[list(psnr_by_job)[-1]]} dB")
    tp, tr = res["plain"]["s"], res["per_request_persistent"]["s"]
    tb, tg = res["batched"]["s"], res["batched_graph"]["s"]
    best = min(tb, tg)
    out = {
        "metric": "Mpaths/s on the uneven batch (64 scenes): one batched refine vs per-request launches",
        "value": round(total_paths / best / 1e6, 3), "unit": "Mpaths/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(best * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded procedural scene generators)",
        "config": {"workload": "uneven batch: 64 terrain scenes, tris log-uniform 1k..2M, side U(128..1024), "
                               "spp in {4,16,64,256}, depth 12, per-scene seeds", "paths": total_paths,
                   "batch_tuning": None if tuned is None else {"kernel": tuned.kernel, "max_paths": tuned.max_paths,
                                                               "stack_lds": tuned.stack_lds},
                   "batch_kernel": res["batched"]["kernel"]},
        "batched_mpaths_s": {"eager": round(total_paths / tb / 1e6, 3), "graph": round(total_paths / tg / 1e6, 3)},
        "per_request_persistent_mpaths_s": round(total_paths / tr / 1e6, 3),
        "per_request_persistent_kernel": res["per_request_persistent"]["kernel"],
        "speedup_vs_per_request_persistent": round(tr / best, 3),
        "plain_mpaths_s": round(total_paths / tp / 1e6, 3), "speedup_vs_plain": round(tp / best, 3),
        "latency_ms": {"batched": lat_ms("batched"), "per_request_persistent": lat_ms("per_request_persistent")},
        "bitwise_equal_to_plain": bool(same and same_graph),
        "psnr_db_vs_cpu": psnr_by_job or None,
        "cpu_baseline": None if args.no_cpu else {
            "value": round(cpu_paths / cpu_s / 1e6, 4), "unit": "Mpaths/s", "cores": threads, "kind": "port",
            "sample": f"2 seeded 64x64 tiles of the smallest, median and largest jobs at full spp ({cpu_paths} "
                      "paths), CPU oracle, std::thread row tiles", "cpu_model": cpu_model()},
    }
    print(json.dumps(out), flush=True)


def np_equal(a, b):
    return a.shape == b.shape and bool((a.view(np.uint32) == b.view(np.uint32)).all())


def bench_serving(args):
    """Serving: Poisson render arrivals drawn from the uneven-batch
    distribution plus refine(4) on a quarter of the rendered scenes, replayed
    in real time against the continuous-batching server."""
    import oxpt
    from oxpt import workloads

    t = time.perf_counter()
    pool = workloads.uneven_batch(n=64, seed=3, threads=min(16, os.cpu_count() or 1), pack=False)
    t_gen = time.perf_counter() - t
    srv = oxpt.Server(0, hbm_budget=args.budget_gb << 30, spp_per_tick=args.spp_tick, device_build=args.device_build,
                      tick_ms=args.tick_ms)
    t = time.perf_counter()
    ids = [srv.add_scene(j.scene, j.env) for j in pool]
    t_add = time.perf_counter() - t
    log(f"64 pool scenes generated in {t_gen:.1f} s, registered in {t_add:.1f} s "
        f"({'device' if args.device_build else 'host'} scene preparation)")
    reqs = workloads.serving_stream(duration=args.duration, rate=args.rate, pool=len(pool))
    n_render = sum(r.kind == "render" for r in reqs)
    scene_of = {r.seq: r.job for r in reqs if r.kind == "render"}
    offered = sum(pool[r.job].paths if r.kind == "render" else
                  r.spp * pool[scene_of[r.job]].opts.width * pool[scene_of[r.job]].opts.height for r in reqs)
    log(f"{len(reqs)} requests ({n_render} renders) over {args.duration:.0f} s at {args.rate} req/s, "
        f"{offered / 1e9:.2f} G paths offered")
    render_of = {}
    handles = []
    t0 = time.perf_counter()
    for r in reqs:
        delay = t0 + r.arrival - time.perf_counter()
        if delay > 0:
            time.sleep(delay)
        if r.kind == "render":
            h = srv.submit_render(ids[r.job], pool[r.job].opts)
            render_of[r.seq] = h
        else:
            h = srv.submit_refine(render_of[r.job], r.spp)
        handles.append(h)
    log(f"all {len(handles)} requests submitted after {time.perf_counter() - t0:.1f} s")
    lat, lat_render, lat_refine, paths, failed = [], [], [], 0, 0
    kinds = [r.kind for r in reqs]
    for i, h in enumerate(handles):
        try:
            info = srv.wait(h, timeout=args.drain_timeout)
        except oxpt.Error:
            failed += 1
            continue
        lat.append(info.latency_s)
        (lat_render if kinds[i] == "render" else lat_refine).append(info.latency_s)
        paths += info.paths
        if (i + 1) % 100 == 0:
            log(f"{i + 1}/{len(handles)} requests done")
    t1 = time.perf_counter()
    st = srv.stats()
    _, kern_ms = srv.kernel_time()
    lat_ms = sorted(x * 1e3 for x in lat)

    def pct(q, xs=None):
        xs = lat_ms if xs is None else sorted(x * 1e3 for x in xs)
        return round(xs[min(len(xs) - 1, int(round(q / 100.0 * (len(xs) - 1))))], 3) if xs else None

    busy = st.busy_s
    out = {
        "metric": "served Mpaths/s; p50/p99 request latency (arrival to result ready)",
        "value": round(paths / (t1 - t0) / 1e6, 3), "unit": "Mpaths/s", "n_gpus": 1, "steps": len(handles),
        "warmup": 0, "ms_per_step": round((t1 - t0) * 1e3 / max(len(handles), 1), 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded procedural scene generators)",
        "config": {"workload": f"serving: Poisson {args.rate} req/s for {args.duration:.0f} s from the uneven-batch "
                               f"distribution (64-scene pool) + refine(4) on 25% of renders; HBM budget "
                               f"{args.budget_gb} GB; " + (f"ticks of ~{args.tick_ms} ms, shortest remaining first"
                                                           if args.tick_ms else f"{args.spp_tick} spp per tick"),
                   "scene_preparation": "device, at first use" if args.device_build else "host, at registration",
                   "registration_s": round(t_add, 3),
                   "requests": len(handles), "failed": failed, "offered_paths": int(offered),
                   "offered_mpaths_s": round(offered / args.duration / 1e6, 3)},
        "latency_ms": {"p50": pct(50), "p90": pct(90), "p99": pct(99), "max": pct(100),
                       "render": {"p50": pct(50, lat_render), "p99": pct(99, lat_render)},
                       "refine": {"p50": pct(50, lat_refine), "p99": pct(99, lat_refine)}},
        "busy_mpaths_s": round(st.paths / busy / 1e6, 3) if busy > 0 else None,
        "server": {"ticks": st.ticks, "uploads": st.uploads, "evictions": st.evictions, "busy_s": round(busy, 3),
                   "kernel_s": round(kern_ms / 1e3, 3),
                   "io_s": round(st.io_s, 3), "failed_ticks": st.failed_ticks, "wall_s": round(t1 - t0, 3),
                   "arrivals_s": round(args.duration, 3), "tick_ms": args.tick_ms,
                   "spp_per_tick": args.spp_tick},
    }
    print(json.dumps(out), flush=True)


def bench_build(args):
    """Scene preparation (BVH, leaf-ordered triangle / attribute words, 4-wide
    nodes, f16 atlas, tables) on the device vs on the host, for the
    single-large-scene and layered/texture workloads; the device words are
    checked equal to the host packer's."""
    import oxpt

    res = {}
    for name in args.build_scenes.split(","):
        scene, env, opts, desc = workload(name)

        def prepare(device):
            pt = oxpt.GpuPathTracer(0)
            pt.set_device_build(device)
            if env is not None:
                pt.set_environment(env)
            t0 = time.perf_counter()
            pt.sync(scene, opts)
            pt.params()  # tables bound and uploaded
            return pt, time.perf_counter() - t0

        dev_t, host_t, infos = [], [], []
        for i in range(args.warmup + args.steps):
            pt, t = prepare(True)
            if i >= args.warmup:
                dev_t.append(t)
                infos.append(pt.build_info())
            log(f"{name}: device prepare {i + 1}/{args.warmup + args.steps}: {t * 1e3:.1f} ms")
        for i in range(min(args.steps, 2)):
            hp, t = prepare(False)
            host_t.append(t)
            log(f"{name}: host prepare {i + 1}: {t * 1e3:.1f} ms")
        equal = all(np.array_equal(pt.scene_words(w).view(np.uint16 if w == 3 else np.uint32),
                                   hp.scene_words(w).view(np.uint16 if w == 3 else np.uint32)) for w in range(4))
        med = sorted(dev_t)[len(dev_t) // 2]
        bi = infos[len(infos) // 2]
        res[name] = {"workload": desc, "triangles": scene.triangle_count(), "device_ms": round(med * 1e3, 2),
                     "device_soup_upload_ms": round(bi.upload_ms, 2), "device_bvh_gather_wide_ms": round(bi.build_ms, 2),
                     "device_atlas_ms": round(bi.atlas_ms, 2), "host_ms": round(min(host_t) * 1e3, 2),
                     "speedup": round(min(host_t) / med, 2), "words_equal": bool(equal)}
        log(json.dumps(res[name]))
    first = res[next(iter(res))]
    out = {"metric": "scene preparation ms (device build: BVH, triangle/attribute words, 4-wide nodes, f16 atlas, "
                     "tables; upload included)", "value": first["device_ms"], "unit": "ms", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": first["device_ms"], "higher_is_better": False,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic (seeded procedural scene generators)",
           "config": {"workload": "scene preparation: " + ", ".join(res)}, "scenes": res}
    print(json.dumps(out), flush=True)


def bench_display(args):
    """The progressive viewer's per-frame cost at 1920x1080 on the room
    scene (the reference's Cornell 1080p rows): refine(1) + draw (resolve
    into device memory, draw_texture) vs refine(1) + image() (resolve +
    readback), and each resolve alone. Median of --steps frames."""
    import oxpt

    scene = oxpt.Scene.room(0)
    opts = oxpt.RenderOptions(width=1920, height=1080, samples_per_pixel=1 << 20, max_bounces=8, rr_start=3)
    pt = oxpt.GpuPathTracer(0)
    pt.sync(scene, opts)
    target = oxpt.DeviceBuffer(1920 * 1080 * 4)

    def draw():
        pt.draw(target.ptr, target.nbytes)
        pt.finish()

    def image():
        pt.image()

    def refine_draw():
        pt.refine(1)
        draw()

    def refine_image():
        pt.refine(1)
        image()

    res = {}
    for name, fn in (("refine1_draw", refine_draw), ("refine1_image", refine_image), ("draw", draw), ("image", image)):
        for _ in range(max(args.warmup, 1)):
            fn()
        times = []
        for _ in range(max(args.steps, 5) * 4):
            t0 = time.perf_counter()
            fn()
            times.append(time.perf_counter() - t0)
        res[name] = round(sorted(times)[len(times) // 2] * 1e3, 3)
        log(f"{name}: {res[name]} ms")
    same = bool(np.array_equal(target.read().reshape(1080, 1920, 4), pt.image()))
    out = {"metric": "progressive frame ms at 1920x1080: refine(1) + draw (device resolve)", "value": res["refine1_draw"],
           "unit": "ms", "n_gpus": 1, "steps": max(args.steps, 5) * 4, "warmup": max(args.warmup, 1),
           "ms_per_step": res["refine1_draw"], "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32", "data": "synthetic (seeded procedural scene generators)",
           "config": {"workload": "viewer loop: closed room (36 tris), 1920x1080, 1 spp per frame, depth 8"},
           "frame_ms": res, "draw_equals_image": same}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="c1", choices=["c0", "c1", "c2", "c3", "c4", "build", "display"])
    ap.add_argument("--build-scenes", default="c1,c2", help="build: scenes to prepare (c0, c1, c2)")
    ap.add_argument("--rate", type=float, default=50.0, help="c4: Poisson arrival rate (req/s)")
    ap.add_argument("--duration", type=float, default=60.0, help="c4: arrival window (s)")
    ap.add_argument("--budget-gb", type=int, default=200, help="c4: HBM admission budget")
    ap.add_argument("--spp-tick", type=int, default=0,
                    help="c4: samples per request per tick (with --tick-ms: a cap, 0 = none)")
    ap.add_argument("--tick-ms", type=int, default=10, help="c4: target tick duration (0: fixed --spp-tick)")
    ap.add_argument("--device-build", action="store_true", help="c4: prepare pool scenes on the device at first use")
    ap.add_argument("--drain-timeout", type=float, default=600.0, help="c4: wait per request (s)")
    ap.add_argument("--cpu-tiles", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle leg (PSNR, baseline, bytes/path)")
    ap.add_argument("--kernel", type=int, default=0, help="0 persistent megakernel, 1 plain per-pixel kernel")
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--waves-per-cu", type=int, default=0)
    ap.add_argument("--wf-variant", type=int, default=0, help="wavefront kernel forms (kernel 7)")
    ap.add_argument("--early-off", action="store_true",
                    help="run the general megakernel forms (no early continuation; same images)")
    ap.add_argument("--tail", type=int, default=0, help="persistent kernel tail in 1/16 of the frame "
                    "(0 default, 65535 none)")
    ap.add_argument("--spp", type=int, default=0, help="override samples per pixel (profiling runs)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device with a plain copy kernel (counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch

        n = args.copy_calibration * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copied {args.copy_calibration} MiB ({n * 4} bytes read, {n * 4} bytes written)")
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(the path would shard by sample range or by scene)", file=sys.stderr)
        sys.exit(2)

    import oxpt

    if args.workload == "c3":
        return bench_uneven_batch(args)
    if args.workload == "c4":
        return bench_serving(args)
    if args.workload == "build":
        return bench_build(args)
    if args.workload == "display":
        return bench_display(args)
    scene, env, opts, desc = workload(args.workload)
    pt = oxpt.GpuPathTracer(0)
    log("device:", pt.adapter_summary())
    pt.set_tuning(block=args.block, waves_per_cu=args.waves_per_cu, kernel=args.kernel, wf_variant=args.wf_variant,
                  tail=args.tail, early=oxpt.EARLY_OFF if args.early_off else 0)
    if args.spp:
        opts.samples_per_pixel = args.spp
    t = time.perf_counter()
    if env is not None:
        pt.set_environment(env)
    pt.sync(scene, opts)
    pt.refine(1)
    pt.finish()
    log(f"scene prepared + uploaded in {time.perf_counter() - t:.2f} s ({scene.triangle_count()} tris)")
    spp = opts.samples_per_pixel
    for i in range(args.warmup):
        pt.reset()
        pt.refine(spp)
        pt.finish()
        log(f"warmup {i + 1}/{args.warmup} done")
    pt.timing(reset=True)
    lat = []
    t0 = time.perf_counter()
    for i in range(args.steps):
        ts = time.perf_counter()
        pt.reset()
        pt.refine(spp)
        pt.finish()
        lat.append(time.perf_counter() - ts)
        log(f"step {i + 1}/{args.steps}: {lat[-1] * 1e3:.1f} ms")
    t1 = time.perf_counter()
    launches, kern_ms = pt.timing()
    kernel_name = pt.last_kernel()
    w, h = opts.width, opts.height
    frame_paths = w * h * spp
    total_paths = frame_paths * args.steps
    elapsed = t1 - t0
    value = total_paths / elapsed / 1e6
    img = pt.image()
    per_launch_paths = total_paths / max(launches, 1)
    avg_launch_s = kern_ms / 1e3 / max(launches, 1)

    roofline = None
    cpu_baseline = None
    psnr_db = None
    bytes_per_path = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle

        pk = oxpt.Packed(scene, opts, env)
        osc = oracle.Scene.from_packed(pk)
        threads = cpu_threads()
        tl = tiles(w, h, args.cpu_tiles) if args.workload != "c0" else [(0, 0, w, h)]
        cpu_paths = sum((r[2] - r[0]) * (r[3] - r[1]) for r in tl) * spp
        # Algorithmic bytes per path: an instrumented pass over every tile
        # (it also warms the CPU caches before the timed runs).
        cacc = np.zeros((h, w, 4), dtype=np.float32)
        cnt_paths = cnt_bytes = 0
        for r in tl:
            _, cnt = osc.render(0, spp, region=r, accum=cacc, threads=threads, counts=True)
            cnt_paths += cnt.paths
            cnt_bytes += cnt.bytes
        bytes_per_path = cnt_bytes / max(cnt_paths, 1)
        log(f"oracle: {bytes_per_path:.0f} algorithmic bytes/path over {len(tl)} tiles")
        # Timed CPU baseline: median of 5 runs after the warm-up pass (as the
        # reference's pt_bench times its renders), every tile at full spp.
        runs = []
        for k in range(5):
            acc = np.zeros((h, w, 4), dtype=np.float32)
            tc = time.perf_counter()
            for r in tl:
                osc.render(0, spp, region=r, accum=acc, threads=threads)
            runs.append(time.perf_counter() - tc)
            log(f"oracle run {k + 1}/5: {cpu_paths} paths in {runs[-1]:.2f} s on {threads} threads")
        cpu_s = sorted(runs)[2]
        rp = oxpt.ResolveParams((1, 3, 0, 0), (1.0, float(spp), float(w), 0.0),
                                tuple(c / 255.0 for c in opts.background))
        cimg = oracle.resolve(acc, rp)
        gt = np.concatenate([img[r[1]:r[3], r[0]:r[2]].reshape(-1, 4) for r in tl])
        ct = np.concatenate([cimg[r[1]:r[3], r[0]:r[2]].reshape(-1, 4) for r in tl])
        psnr_db = psnr(gt, ct)
        sample = (f"{len(tl)} seeded 64x64 tiles of the same frame at {spp} spp ({cpu_paths} paths)"
                  if args.workload != "c0" else f"the full frame at {spp} spp ({cpu_paths} paths)")
        cpu_baseline = {
            "value": round(cpu_paths / cpu_s / 1e6, 4), "unit": "Mpaths/s", "cores": threads, "kind": "port",
            "sample": sample + ", CPU oracle (C++ restatement of the reference estimator), std::thread row "
                      "tiles, median of 5 runs after a warm-up",
            "cpu_model": cpu_model(), "host_cpus_visible": os.cpu_count(),
        }
        achieved = bytes_per_path * per_launch_paths / avg_launch_s / 1e9 if avg_launch_s > 0 else 0.0
        measured = pmc_measured(args)
        cache_resident = args.workload == "c0"
        roofline = {
            "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            # A working set of a few KB lives in the caches: an HBM fraction
            # of it would mean nothing.
            "frac": None if cache_resident else round(achieved / HBM_PEAK_GBS, 4),
            "traffic": pmc_traffic(args, per_launch_paths),
            "model": "no-cache algorithmic bytes per path (instrumented oracle, every tile) x paths per launch / "
                     "the launch's HIP-event time",
            "kernel": kernel_name,
            "bytes_per_path": round(bytes_per_path, 1), "paths_per_launch": int(per_launch_paths),
            "avg_launch_ms": round(avg_launch_s * 1e3, 3),
        }
        if measured is not None:
            roofline["measured"] = measured
            roofline["limiter"] = measured.get("limiter")
    lat_ms = sorted(x * 1e3 for x in lat)

    def pct(p):
        if not lat_ms:
            return None
        k = min(len(lat_ms) - 1, max(0, int(round(p / 100.0 * (len(lat_ms) - 1)))))
        return round(lat_ms[k], 3)

    out = {
        "metric": "Mpaths/s at 1080p 64spp; p50/p99 render latency; PSNR>=60dB vs CPU tracer",
        "value": round(value, 3),
        "unit": "Mpaths/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural scene generators)",
        "config": {"workload": desc, "width": w, "height": h, "spp": spp, "max_bounces": opts.max_bounces,
                   "triangles": scene.triangle_count(), "kernel": kernel_name},
        "latency_ms": {"p50": pct(50), "p99": pct(99)},
        "psnr_db_vs_cpu": None if psnr_db is None else round(psnr_db, 2),
        "roofline": roofline,
        "cpu_baseline": cpu_baseline,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
