"""Closest-hit throughput of the MI355X tracer on the large instanced scene.

    python bench.py [--gpus 1] [--steps K] [--warmup W]

Workload: 64 displaced icospheres (30,420 triangles each) placed as 1,000 rigid
instances in a 20^3 volume, realized into one 30.4M-triangle list and BVH as
compile_scene does; a 1920x1080 pinhole frame. One step traces two batches of
closest-hit rays, each 2,073,600 rays:
  primary  the frame's pixel-centre rays (generated in the kernel);
  bounce   cosine-distributed rays leaving each primary hit (rays that missed
           are re-aimed at the volume), the incoherent second bounce.
Both go to the library in one call (lr_trace_frame_and_closest_device): one
persistent launch takes the frame's tiles and then the batch's rays from one
counter, so they share one ramp-up and one drain (each batch in a call of its
own is timed and reported too). Inputs are resident in HBM before timing; HIP
events on the device's stream bracket the timed steps. One JSON line is
printed.

Beside the metric, reported in `config`: the full path tracer on the same
scene (path_tracing) and the serving loop under Poisson arrivals (serving:
p50/p99 time to the target sample count, p99 time to 64 samples, rays/s).
"""
import argparse
import json
import math
import os
import sys
import time
from dataclasses import replace
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path[:0] = [str(ROOT / "luce-render_amd"), str(ROOT)]

WIDTH, HEIGHT = 1920, 1080
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` (name<width,stack,group>) from the
    committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary, or None."""
    path = ROOT / "profiles" / "closest_hit_instanced_30M.md"
    if not path.exists():
        return None
    for line in path.read_text().splitlines():
        parts = line.split()
        if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch" and parts[1] == kernel:
            return int(parts[2])
    return None


def host_cpu():
    """The host's CPU model, its logical CPUs and the ones this process may use."""
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return {"model": model, "nproc": os.cpu_count() or 1, "usable": usable}


def timed_median(fn, runs):
    """One warm-up call, then the median wall time of `runs` calls (steady clock)."""
    fn()
    times = []
    for _ in range(runs):
        t1 = time.perf_counter()
        fn()
        times.append(time.perf_counter() - t1)
    return float(np.median(times)), times


def cpu_protocol(O, render, scenes, threads, full=False, runs=5):
    """The reference's CPU path, timed on this host: the f64 brute-force
    intersector (tests/gpu/main.lucb:82-119) on the 256x256 Cornell probe with
    its 10,000 random triangles (65,536 x 10,034 ray-triangle tests) on all
    usable cores (rows split over threads) and on one core, and the
    single-threaded binned-SAH build (src/render/bvh.lucb:66-217, restated in
    the oracle) on that scene and on the 2M-unique-triangle geometry of the
    large scene (its 64 icospheres once each). One warm-up, then the median of
    `runs`. The one-core brute force runs on the image's first 4,096 rays
    unless `full` (the rate is per ray: every ray tests every triangle); with
    `full` the build also runs once on the realized 30.4M triangles."""
    out = {"cpu": host_cpu(), "threads": threads, "runs": runs}
    c1 = render.compile_scene(scenes.cornell_probe_set())
    f64 = O.ray_frame(c1.cameras[0].matrix, False, 36.0, 30.0, 256, 256)
    tests = 256 * 256 * c1.triangle_count()
    med, _ = timed_median(lambda: O.brute_force(c1.positions, c1.indices, c1.faces, frame=f64, width=256, height=256,
                                                threads=threads), runs)
    out["brute_force_c1_all_cores"] = {"threads": threads, "s": round(med, 4), "mrays_s": round(65536 / med / 1e6, 5),
                                       "ns_per_test": round(med * 1e9 * threads / tests, 3)}
    rows = 256 if full else 16
    d = np.array([O.frame_ray(f64, x + 0.5, y + 0.5)[1] for y in range(rows) for x in range(256)], np.float64)
    o = np.tile(np.asarray(O.frame_ray(f64, 0.5, 0.5)[0], np.float64), (len(d), 1))
    r64 = np.concatenate([o, d], 1)
    med1, _ = timed_median(lambda: O.brute_force(c1.positions, c1.indices, c1.faces, rays=r64, threads=1), runs)
    out["brute_force_c1_one_core"] = {"rays": len(d), "s": round(med1, 4), "mrays_s": round(len(d) / med1 / 1e6, 6),
                                      "ns_per_test": round(med1 * 1e9 / (len(d) * c1.triangle_count()), 3),
                                      "s_for_65536_rays": round(med1 * 65536 / len(d), 2)}
    inv = np.empty_like(c1.order)
    inv[c1.order] = np.arange(len(c1.order), dtype=np.uint32)
    idx1 = np.ascontiguousarray(c1.indices[inv])
    med, _ = timed_median(lambda: O.bvh_build(c1.positions, idx1), runs)
    nodes1 = O.bvh_build(c1.positions, idx1)[0]
    out["bvh_build_c1_one_thread"] = {"triangles": c1.triangle_count(), "s": round(med, 5), "nodes": len(nodes1)}
    c3 = render.compile_scene(scenes.instanced_set(instances=64), geometry=False)
    med, _ = timed_median(lambda: O.bvh_build(c3.positions, c3.indices), runs)
    nodes3 = O.bvh_build(c3.positions, c3.indices)[0]
    out["bvh_build_c3_unique_one_thread"] = {"triangles": c3.triangle_count(), "s": round(med, 3),
                                             "nodes": len(nodes3)}
    if full:
        big = render.compile_scene(scenes.instanced_set(instances=1000), geometry=False)
        t1 = time.perf_counter()
        nodes = O.bvh_build(big.positions, big.indices)[0]
        out["bvh_build_c3_realized_one_thread"] = {"triangles": big.triangle_count(),
                                                   "s": round(time.perf_counter() - t1, 2), "nodes": len(nodes)}
    return out


def bounce_rays(sc, frame, faces_t, rng):
    """Cosine-distributed rays from each primary hit (8 floats a ray)."""
    n = WIDTH * HEIGHT
    hit = faces_t[:, 0] != 0xFFFFFFFF
    t = faces_t[:, 1].view(np.float32).astype(np.float64)
    ys, xs = np.divmod(np.arange(n), WIDTH)
    d = frame.corner[None, :] + (xs + 0.5)[:, None] * frame.du[None, :] + (ys + 0.5)[:, None] * frame.dv[None, :]
    p = frame.position[None, :] + d * np.where(hit, t, 0.0)[:, None]
    # hit points: offset along the unit direction back toward the camera, and a
    # random outward normal proxy (the sphere's radial direction is unknown
    # here, so the cosine lobe is taken about -d)
    dn = d / np.linalg.norm(d, axis=1, keepdims=True)
    normal = -dn
    u1, u2 = rng.random(n), rng.random(n)
    r = np.sqrt(u1)
    phi = 2 * np.pi * u2
    local = np.stack([r * np.cos(phi), r * np.sin(phi), np.sqrt(np.maximum(0.0, 1 - u1))], 1)
    a = np.where(np.abs(normal[:, 0:1]) > 0.9, np.array([[0.0, 1.0, 0.0]]), np.array([[1.0, 0.0, 0.0]]))
    tx = np.cross(normal, a)
    tx /= np.linalg.norm(tx, axis=1, keepdims=True)
    ty = np.cross(normal, tx)
    wd = local[:, 0:1] * tx + local[:, 1:2] * ty + local[:, 2:3] * normal
    origin = p + normal * 1e-3
    # rays that missed start inside the volume instead
    miss = ~hit
    origin[miss] = rng.uniform(-10, 10, (miss.sum(), 3))
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = origin
    rays[:, 3] = 3.0e38
    rays[:, 4:7] = wd
    return rays


def guiding_leg(device, stream, render, gset, width, height, ref_spp, kw, spp=64, methods=("plain", "grid", "learned")):
    """Plain, grid-guided and learned-guided rendering against a long plain
    reference, measured (no extrapolation): the plain method in the
    reference's schedule (one sample a pass) at `spp` samples sets the time
    budget T and the MSE to reach; every method is then rendered batched
    (passes of several samples) for T of wall time (its sample count from its
    measured time a sample; the time each took is reported), and at doubling
    sample counts until its MSE falls below plain's, the time to plain's MSE
    read between the two measured renders around it (log-log)."""
    from luce_render.loading import RenderSettings
    sc = render.compile_scene(gset)
    gpu = render.GpuScene.upload(device, sc)

    def run(samples, seed, guiding, pass_samples=0):
        st = RenderSettings(width, height, samples, spectral=False, guiding=guiding, seed=seed,
                            pass_samples=pass_samples, **kw)
        r = render.Renderer.create(device, sc, st, gpu_scene=gpu)
        t = render.Timer()
        render.synchronize(device)
        t.begin(stream)
        r.finish()
        t.end(stream)
        ms = t.elapsed_ms()
        t.destroy()
        img = r.read()[:, :, :3].astype(np.float64)
        r.destroy()
        return ms, img

    _, ref = run(ref_spp, 1, 0)
    out = {"workload": f"{width}x{height}, {sc.triangle_count()} triangles, reference {ref_spp} spp plain",
           "spp": spp}
    budget_ms, plain_img = run(spp, 2, 0, pass_samples=1)
    plain_mse = float(((plain_img - ref) ** 2).mean())
    out["plain_one_sample_a_pass"] = {"ms": round(budget_ms, 2), "mse": plain_mse}
    for name, mode in (("plain", 0), ("grid", 1), ("learned", 2)):
        if name not in methods:
            continue
        ms_spp, img = run(spp, 2, mode)
        mse_spp = float(((img - ref) ** 2).mean())
        n = max(1, int(spp * budget_ms / ms_spp))
        ms_eq, img = run(n, 2, mode)
        eq_mse = float(((img - ref) ** 2).mean())
        # doubling (or halving) sample counts until two renders lie either
        # side of plain's MSE
        points = [(ms_spp, mse_spp)]
        samples = spp
        if mse_spp > plain_mse:
            while points[-1][1] > plain_mse and samples < 16 * spp:
                samples *= 2
                ms, img = run(samples, 2, mode)
                points.append((ms, float(((img - ref) ** 2).mean())))
        else:
            while points[0][1] <= plain_mse and samples > 1:
                samples //= 2
                ms, img = run(samples, 2, mode)
                points.insert(0, (ms, float(((img - ref) ** 2).mean())))
        reach = None
        for (t0, e0), (t1, e1) in zip(points, points[1:]):
            if e0 > plain_mse >= e1 and e1 > 0 and e0 > e1:
                f = (math.log(e0) - math.log(plain_mse)) / (math.log(e0) - math.log(e1))
                reach = math.exp(math.log(t0) + f * (math.log(t1) - math.log(t0)))
        out[name] = {"ms_at_spp": round(ms_spp, 2), "mse_at_spp": mse_spp,
                     "equal_time": {"samples": n, "ms": round(ms_eq, 2), "mse": eq_mse,
                                    "mse_over_plain": round(eq_mse / plain_mse, 3)},
                     "time_to_plain_mse_ratio": None if reach is None else round(reach / budget_ms, 3),
                     "measured": [[round(t, 2), e] for t, e in points]}
    gpu.destroy()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--instances", type=int, default=1000)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-full", action="store_true",
                    help="CPU protocol in full: one-core brute force over all 65,536 probe rays and a one-thread "
                         "build of the realized 30.4M triangles")
    ap.add_argument("--config", default="", help="width,group,stack[,refill] to skip the sweep")
    ap.add_argument("--render-spp", type=int, default=16,
                    help="samples of the full path-traced render of the same scene (0: skip)")
    ap.add_argument("--render-tune", type=int, default=1,
                    help="tune the render's own traversal variant on a ten-sample pass (0: the batch pick)")
    ap.add_argument("--serving", type=int, default=80,
                    help="requests of the Poisson serving leg (40 requests/s over the four serving scenes; 0: skip)")
    ap.add_argument("--serving-sustained", type=int, default=48,
                    help="requests of a second serving run at 3/4 of the measured capacity (0: skip)")
    ap.add_argument("--guiding", default="1920x1080:4096",
                    help="WxH:reference-spp of the guiding leg on the glossy workload (empty: skip)")
    ap.add_argument("--guide-methods", default="plain,grid,learned",
                    help="methods of the guiding leg (profiling one alone: learned)")
    ap.add_argument("--guide-scenes", default="glossy,lamp_room", help="workloads of the guiding leg")
    ap.add_argument("--build", type=int, default=3,
                    help="device scene builds of the same scene to time against the host pipeline (0: skip)")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="also run the plain copy kernel over this many MiB (a known-traffic dispatch "
                         "for scaling hardware byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; --gpus above 1 is not supported "
                                   "(rays shard by sample or ray range, films reduce once: see the design notes)"}))
        return 2

    from luce_render import render, scenes

    t0 = time.time()
    gset = scenes.instanced_set(instances=args.instances)
    sc = render.compile_scene(gset)
    log(f"scene: {sc.triangle_count()} triangles, {len(sc.nodes)} nodes, compiled in {time.time() - t0:.1f} s")
    device = render.Device.open()
    gpu = render.GpuScene.upload(device, sc)
    info = gpu.info()
    frame = render.ray_frame(sc.cameras[0], WIDTH, HEIGHT)
    stream = render.device_stream(device)
    n = WIDTH * HEIGHT
    hits_p = render.Buffer(device, n * 8)
    hits_b = render.Buffer(device, n * 8)

    render.probe_device(gpu, frame, hits_p, stream)
    render.synchronize(device)
    primary = hits_p.read(np.uint32).reshape(n, 2)
    rng = np.random.default_rng(7)
    rays_b_host = bounce_rays(sc, frame, primary, rng)
    rays_b = render.Buffer.create(device, rays_b_host)

    def step():
        # both batches from one counter in one persistent launch
        render.trace_frame_and_closest_device(gpu, frame, hits_p, rays_b, n, hits_b, stream)

    def time_steps(k):
        timer = render.Timer()
        render.synchronize(device)
        timer.begin(stream)
        for _ in range(k):
            step()
        timer.end(stream)
        ms = timer.elapsed_ms()
        timer.destroy()
        return ms / k

    # Traversal variant: the autotuner's sweep over BVH width x workgroup x
    # LDS stack depth x ray schedule, timed on the step itself (both batches
    # in one call; every variant returns the same hits).
    if args.config:
        best = tuple(int(x) for x in args.config.split(","))
        best = best + (0,) * (4 - len(best))
        sweep = None
    else:
        from luce_render import autotune
        split = {}
        _, _ = autotune.tune(device, gpu, rays_b, n, hits=hits_b, frame=frame, repeats=2, split=split)
        sweep = split["fused"]
        best = min(sweep, key=sweep.get)
        log("sweep (ms, primary and bounce in one call):",
            {f"w{k[0]}g{k[1]}s{k[2]}r{k[3]}": round(v, 3) for k, v in sweep.items()})
    gpu.set_trace_config(*best)

    for _ in range(args.warmup):
        step()
    render.synchronize(device)
    # Per-launch times on the device stream: the step's one launch, and the
    # primary and bounce batches each in a call of its own.
    tp, tb, tf = render.Timer(), render.Timer(), render.Timer()
    kp, kb, kf = [], [], []
    for _ in range(args.steps):
        tf.begin(stream)
        step()
        tf.end(stream)
        tp.begin(stream)
        render.probe_device(gpu, frame, hits_p, stream)
        tp.end(stream)
        tb.begin(stream)
        render.trace_closest_device(gpu, rays_b, n, hits_b, stream)
        tb.end(stream)
        kf.append(tf.elapsed_ms())
        kp.append(tp.elapsed_ms())
        kb.append(tb.elapsed_ms())
    ms_step = time_steps(args.steps)
    render.synchronize(device)
    total_rays = 2 * n
    value = total_rays / (ms_step * 1e-3) / 1e6
    ms_p, ms_b, ms_f = float(np.mean(kp)), float(np.mean(kb)), float(np.mean(kf))

    # Algorithmic bytes per ray, counted by the oracle's f32 mirror of the
    # reference traversal on samples of both batches: 32 B ray + 8 B result,
    # 64 B per interior step (two 32-byte child nodes), 36 B per triangle test.
    from oracle import oracle as O
    sample = rng.choice(n, 16384, replace=False)
    ys, xs = np.divmod(sample, WIDTH)
    f16 = frame.pack().astype(np.float64)
    pd = np.stack([f16[4] + (xs + 0.5) * f16[8] + (ys + 0.5) * f16[12], f16[5] + (xs + 0.5) * f16[9] + (ys + 0.5) * f16[13],
                   f16[6] + (xs + 0.5) * f16[10] + (ys + 0.5) * f16[14]], 1)
    prays = np.concatenate([np.tile(f16[0:3], (len(sample), 1)), pd], 1).astype(np.float32)
    _, _, cp = O.trace_rays(sc.nodes, sc.positions, sc.indices, sc.faces, prays)
    brays = rays_b_host[sample][:, [0, 1, 2, 4, 5, 6]]
    _, _, cb = O.trace_rays(sc.nodes, sc.positions, sc.indices, sc.faces, brays)

    def per_ray(c):
        return 40.0 + (64.0 * c[0] + 36.0 * c[2]) / len(sample)

    bytes_p, bytes_b = per_ray(cp), per_ray(cb)
    # Dominant kernel: the step's one persistent launch over both batches
    # (refill above 0), its algorithmic bytes both batches'; with refill 0
    # the step is two launches and the bounce batch's dominates.
    if best[3] > 0:
        dom_bytes, dom_ms, dom_name = (bytes_p + bytes_b) * n, ms_f, "frame_closest_stream_kernel"
    else:
        dom_bytes, dom_ms, dom_name = bytes_b * n, ms_b, "closest_kernel"
    achieved = dom_bytes / (dom_ms * 1e-3) / 1e9
    traffic = measured_traffic(f"{dom_name}<{best[0]},{best[2]},{best[1]}>")

    # Scene build on the device (lr_scene_build): the flattened arrays in
    # their original order (recovered through the leaf order) to a traceable
    # scene - BVH build, reorder, collapse to widths 2/4/8, triangle packing -
    # against the host pipeline it replaces: build_bvh with its reorder on all
    # host cores, then the host collapse of the three widths.
    build_stats = None
    if args.build > 0:
        import ctypes
        from luce_render import _lib
        inv = np.empty_like(sc.order)
        inv[sc.order] = np.arange(len(sc.order), dtype=np.uint32)

        def flat_scene():
            f = render.CompiledScene()
            f.positions = sc.positions
            f.indices, f.faces = np.ascontiguousarray(sc.indices[inv]), np.ascontiguousarray(sc.faces[inv])
            f.shading, f.uvs = np.ascontiguousarray(sc.shading[inv]), np.ascontiguousarray(sc.uvs[inv])
            return f

        flat = flat_scene()
        dev_s = []
        for _ in range(args.build):
            t1 = time.perf_counter()
            g = render.GpuScene.build(device, flat, keep_tree=False)
            render.synchronize(device)
            dev_s.append(time.perf_counter() - t1)
            same = all(np.array_equal(g.layout(w), gpu.layout(w)) for w in (2, 4, 8))
            g.destroy()
        host = flat_scene()
        host_threads = min(16, os.cpu_count() or 1)
        t1 = time.perf_counter()
        render.build_bvh(host, threads=host_threads)
        host_build = time.perf_counter() - t1
        t1 = time.perf_counter()
        size = ctypes.c_size_t(0)
        for w in (2, 4, 8):
            _lib.check(_lib.lib().lr_bvh_collapse(_lib.ptr(host.nodes), len(host.nodes), w, None, 0,
                                                  ctypes.byref(size), None))
        host_collapse = time.perf_counter() - t1
        # Refit of the uploaded scene to moved points (same triangles), then
        # back: the first call also records the layouts' child slots.
        moved = (sc.positions * np.float32(1.001)).astype(np.float32)
        refit_s = []
        for pts in (moved, sc.positions, moved, sc.positions):
            t1 = time.perf_counter()
            gpu.refit(pts)
            refit_s.append(time.perf_counter() - t1)
        build_stats = {"triangles": sc.triangle_count(), "device_s": round(float(np.median(dev_s)), 4),
                       "refit_first_s": round(refit_s[0], 4), "refit_s": round(float(np.median(refit_s[1:])), 4),
                       "host_build_s": round(host_build, 3), "host_collapse_s": round(host_collapse, 3),
                       "host_threads": host_threads, "layouts_identical": bool(same),
                       "speedup": round((host_build + host_collapse) / float(np.median(dev_s)), 1)}
        log("scene build:", build_stats)

    # The full wavefront path tracer on the same scene: RGB, 12 bounces,
    # 256 lights in the light tree, light tracing at the reference's default.
    render_stats = None
    if args.render_spp > 0:
        from luce_render.loading import RenderSettings
        st = RenderSettings.of(WIDTH, HEIGHT, args.render_spp)
        st.spectral = False
        st.max_bounces = 12
        st.trace = tuple(best[:3])
        render_tune = None
        if args.render_tune and not args.config:
            # the integrator's own tracing kernels, tuned on one pass of ten
            # samples of this render (every variant renders the same image)
            from luce_render import autotune
            pick, table = autotune.tune_render(device, sc, gpu, replace(st, samples=min(10, args.render_spp)),
                                               widths=(4, 20, 24))
            st.trace = tuple(pick)
            render_tune = {",".join(map(str, k)): round(v, 3) for k, v in sorted(table.items(), key=lambda kv: kv[1])}
            log("render variants (ms a sample):", render_tune)
        renderer = render.Renderer.create(device, sc, st, gpu_scene=gpu)
        warm = render.Renderer.create(device, sc, RenderSettings(WIDTH, HEIGHT, 1, spectral=False, trace=st.trace),
                                      gpu_scene=gpu)
        warm.finish()
        warm.destroy()
        render.synchronize(device)
        timer = render.Timer()
        timer.begin(stream)
        renderer.finish()
        timer.end(stream)
        ms = timer.elapsed_ms()
        timer.destroy()
        stats = renderer.stats()
        img = renderer.read()
        renderer.destroy()
        rays = stats["closest_rays"] + stats["shadow_rays"]
        # The same render with the renderer's per-kernel timers (one pass in
        # flight, events between the stages): the closest-hit stage's own
        # rate inside the render (BVH walk, analytic lights, routing).
        prof = render.Renderer.create(device, sc, replace(st, profile=True), gpu_scene=gpu)
        prof.finish()
        kt = prof.kernel_times()
        prof.destroy()
        closest_ms = kt["intersect closest"]
        render_stats = {"spp": args.render_spp, "ms_per_sample": round(ms / args.render_spp, 3),
                        "samples_per_s": round(args.render_spp * WIDTH * HEIGHT / (ms * 1e-3), 1),
                        "mrays_per_s_in_render": round(rays / (ms * 1e-3) / 1e6, 1),
                        "closest_rays": stats["closest_rays"], "shadow_rays": stats["shadow_rays"],
                        "closest_stage_ms_per_sample": round(closest_ms / args.render_spp, 3),
                        "closest_mrays_per_s_in_render": round(stats["closest_rays"] / (closest_ms * 1e-3) / 1e6, 1),
                        "batch_over_render_closest": round(value / (stats["closest_rays"] / (closest_ms * 1e-3) / 1e6), 3),
                        "shadow_stage_ms_per_sample": round(kt["intersect shadow"] / args.render_spp, 3),
                        "shade_stage_ms_per_sample": round(kt["shade surface"] / args.render_spp, 3),
                        "finite": bool(np.all(np.isfinite(img))), "trace_config": list(st.trace),
                        "variants_ms_per_sample": render_tune}
        log("render:", render_stats)

    # The serving form: Poisson arrivals (40 requests/s) over the four serving
    # scenes, resolutions 256^2..1920x1080, 16..256 samples, shared launches
    # under a 200 GB budget; the schedule is a prefix of the 2,000-request one.
    serving_stats = None
    if args.serving > 0:
        from luce_render import serving
        served = serving.prepare_scenes(device, scenes.serving_sets())
        schedule = serving.poisson_schedule(count=args.serving, rate=40.0, scenes=len(served))
        rep = serving.run(device, served, schedule, server_kw={"budget_bytes": 200 << 30}, max_bounces=8,
                          light_tracing=0.5, log=log)

        def summary(rep, rate):
            return {"requests": rep.requests, "rate_per_s": round(rate, 3), "p50_s": round(rep.p50_s, 3),
                    "p99_s": round(rep.p99_s, 3), "p99_to_64spp_s": round(rep.p99_to_64spp_s, 3),
                    "mean_s": round(rep.mean_s, 3), "wall_s": round(rep.wall_s, 2),
                    "mrays_per_s": round(rep.rays_per_s / 1e6, 1), "peak_active": rep.peak_active}
        serving_stats = summary(rep, 40.0)
        # 40 requests a second overload one GPU from the first second on, so
        # the run drains a full queue: its requests over its wall time are
        # the GPU's capacity for this mix. The same mix offered at 3/4 of that
        # rate (a fresh schedule of its own) shows the latency of a load the
        # GPU sustains.
        capacity = rep.requests / rep.wall_s
        serving_stats["capacity_req_per_s"] = round(capacity, 3)
        if args.serving_sustained > 0:
            rate = 0.75 * capacity
            steady = serving.poisson_schedule(count=args.serving_sustained, rate=rate, scenes=len(served), seed=23)
            rep2 = serving.run(device, served, steady, server_kw={"budget_bytes": 200 << 30}, max_bounces=8,
                               light_tracing=0.5, log=log)
            serving_stats["sustained"] = summary(rep2, rate)
        log("serving:", serving_stats)
        for g, _ in served:
            g.destroy()

    # Path guiding: plain, the hash grid and the learned guide at 64 samples
    # against a long plain reference; MSE, wall time, and the time each guided
    # render would need to reach plain's 64-sample MSE (MSE falling as
    # 1 / samples), over plain's 64-sample time.
    guiding_stats = None
    if args.guiding:
        size, ref_spp = args.guiding.split(":")
        gw, gh = (int(x) for x in size.split("x"))
        from luce_render.loading import OpenPbr
        guiding_stats = {}
        for name, gset, w, h, rspp, kw in (
                ("glossy", scenes.glossy_set(soup=0), gw, gh, int(ref_spp), dict(max_bounces=8, light_tracing=0.5)),
                ("lamp_room", scenes.lamp_room_set(), 960, 720, 1024,
                 dict(max_bounces=2, light_tracing=0.0,
                      material=OpenPbr(base_color=(0.3, 0.3, 0.3), specular_weight=0.0)))):
            if name not in args.guide_scenes.split(","):
                continue
            guiding_stats[name] = guiding_leg(device, stream, render, gset, w, h, rspp, kw,
                                              methods=tuple(args.guide_methods.split(",")))
            log(f"guiding {name}:", guiding_stats[name])

    cpu = None
    cpu_detail = None
    if not args.no_cpu:
        host = host_cpu()
        threads = max(1, min(16, host["usable"]))
        m = 512
        sel = rng.choice(n, m, replace=False)
        ys2, xs2 = np.divmod(sel, WIDTH)
        d64 = frame.corner[None, :] + (xs2 + 0.5)[:, None] * frame.du[None, :] + (ys2 + 0.5)[:, None] * frame.dv[None, :]
        r64 = np.concatenate([np.tile(frame.position, (m, 1)), d64], 1)
        t1 = time.perf_counter()
        O.brute_force(sc.positions, sc.indices, sc.faces, rays=r64, threads=threads)
        secs = time.perf_counter() - t1
        cpu = {"value": m / secs / 1e6, "unit": "Mrays/s", "cores": threads, "kind": "port",
               "sample": f"{m} primary rays of the 1920x1080 frame, f64 brute force over all "
                         f"{sc.triangle_count()} triangles (tests/gpu/main.lucb:82-119), {secs:.1f} s on "
                         f"{threads} threads of {host['model']} ({host['nproc']} logical CPUs, {host['usable']} usable)"}
        log("cpu baseline:", cpu)
        cpu_detail = cpu_protocol(O, render, scenes, threads, full=args.cpu_full)
        log("cpu protocol:", cpu_detail)

    out = {
        "metric": "Mrays/s closest-hit",
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "closest-hit, 1920x1080 primary + bounce-1 batches over the instanced "
                        f"{sc.triangle_count()}-triangle scene (1,000 instances of 64 displaced icospheres), "
                        "both traced in one call (lr_trace_frame_and_closest_device)",
            "rays_per_step": total_rays,
            "trace_config": {"width": best[0], "group": best[1], "stack": best[2], "refill": best[3]},
            "autotune_ms": None if sweep is None else {",".join(map(str, k)): round(v, 3) for k, v in sweep.items()},
            "primary_mrays_s": round(n / (ms_p * 1e-3) / 1e6, 2),  # each batch in a call of its own
            "bounce_mrays_s": round(n / (ms_b * 1e-3) / 1e6, 2),
            "separate_calls_mrays_s": round(total_rays / ((ms_p + ms_b) * 1e-3) / 1e6, 2),
            "bytes_per_ray": {"primary": round(bytes_p, 1), "bounce": round(bytes_b, 1)},
            "bvh_nodes": info["binary_nodes"],
            "path_tracing": render_stats,
            "serving": serving_stats,
            "guiding": guiding_stats,
            "scene_build": build_stats,
            "wide_bytes": info["wide_bytes"],
            "cpu_protocol": cpu_detail,
        },
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": dom_name,
                     "kernel_ms": round(dom_ms, 4)},
        "cpu_baseline": cpu,
    }
    if args.copy_mib:
        nbytes = args.copy_mib << 20
        a, b = render.Buffer(device, nbytes), render.Buffer(device, nbytes)
        for _ in range(3):
            b.copy_from(a)
        render.synchronize(device)
        a.destroy()
        b.destroy()
        log(f"copy kernel: 3 dispatches of {nbytes} bytes read + {nbytes} written")
    print(json.dumps(out), flush=True)
    for b in (hits_p, hits_b, rays_b):
        b.destroy()
    gpu.destroy()
    device.destroy()
    return 0


if __name__ == "__main__":
    sys.exit(main())
