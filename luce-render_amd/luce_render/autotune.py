"""The traversal autotuner: times every (BVH width, workgroup size, LDS stack
depth, ray schedule) variant of the closest-hit kernel on a scene and a ray
batch and picks the fastest. The variants return identical hits (their
parity is tested), so the choice is by time alone.

    best, table = autotune.tune(device, gpu_scene, rays, n)
    gpu_scene.set_trace_config(*best)
"""
import json

from . import render

WIDTHS = (2, 4, 8, 20, 24)  # 20, 24: the quantized 4- and 8-wide layouts
GROUPS = (64, 128, 256)
STACKS = (16, 24, 32, 48)
REFILLS = (0, 32, 48)  # 0: one ray per thread; k: persistent waves refilled at k idle lanes


def tune(device, scene, rays, n, hits=None, widths=WIDTHS, groups=GROUPS, stacks=STACKS, refills=REFILLS, repeats=3,
         frame=None, split=None):
    """Times trace_closest_device over `rays` (a Buffer of n rays), and the
    probe of `frame` too when given, for each variant; returns the fastest
    (width, group, stack, refill) and {variant: ms per launch}. A dict
    `split` receives the two calls' own tables under "trace" and "probe"
    (the probe takes one ray per thread whatever the refill) and, with a
    frame, under "fused" the time of both in one call
    (trace_frame_and_closest_device)."""
    stream = render.device_stream(device)
    own = hits is None
    if own:
        hits = render.Buffer(device, max(n, 1) * 8)
    probe_hits = render.Buffer(device, frame.width * frame.height * 8) if frame is not None else None
    table = {}
    timer = render.Timer()
    try:
        for w in widths:
            for g in groups:
                for s in stacks:
                    for r in refills:
                        scene.set_trace_config(w, g, s, r)

                        def trace():
                            render.trace_closest_device(scene, rays, n, hits, stream)

                        def probe():
                            render.probe_device(scene, frame, probe_hits, stream)

                        def fused():
                            render.trace_frame_and_closest_device(scene, frame, probe_hits, rays, n, hits, stream)

                        ms = {}
                        for name, call in (("trace", trace), ("probe", probe), ("fused", fused)):
                            if name != "trace" and frame is None:
                                continue
                            if name == "fused" and split is None:
                                continue
                            call()  # first launch: spill area and code load
                            render.synchronize(device)
                            timer.begin(stream)
                            for _ in range(repeats):
                                call()
                            timer.end(stream)
                            ms[name] = timer.elapsed_ms() / repeats
                            if split is not None:
                                split.setdefault(name, {})[(w, g, s, r)] = ms[name]
                        table[(w, g, s, r)] = ms["trace"] + ms.get("probe", 0.0)
    finally:
        timer.destroy()
        if own:
            hits.destroy()
        if probe_hits is not None:
            probe_hits.destroy()
    best = min(table, key=table.get)
    scene.set_trace_config(*best)
    return best, table


def save(path, best, table, key=""):
    with open(path, "w") as f:
        json.dump({"key": key, "best": list(best), "table": {",".join(map(str, k)): ms for k, ms in table.items()}},
                  f, indent=1)


def load(path, key=""):
    """The cached choice for `key`, or None."""
    try:
        with open(path) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return None
    if d.get("key") != key:
        return None
    return tuple(d["best"])


RENDER_GROUPS = (64, 128, 256)


def tune_render(device, scene, gpu_scene, settings, widths=WIDTHS, groups=RENDER_GROUPS, stacks=STACKS, repeats=1):
    """The integrator's traversal variant: renders `settings` (a
    RenderSettings; its sample count is the timed work, a pass or two is
    enough) once per (layout, group, stack) and picks the fastest by wall
    time on the device's stream. The render's tracing kernels take one item
    per thread, so the ray schedule is not swept. Every variant returns the
    same hits, so every render is the same image; the choice is by time
    alone. Returns the fastest (width, group, stack) and {variant: ms a
    sample}."""
    import dataclasses
    stream = render.device_stream(device)
    table = {}
    timer = render.Timer()
    try:
        for w in widths:
            for g in groups:
                for s in stacks:
                    st = dataclasses.replace(settings, trace=(w, g, s))
                    r = render.Renderer.create(device, scene, st, gpu_scene=gpu_scene)
                    try:
                        r.finish()  # code load, buffers, tables
                        best = None
                        for _ in range(repeats):
                            r.restart(r.frame)  # the same view again: buffers kept, the film cleared
                            render.synchronize(device)
                            timer.begin(stream)
                            r.finish()
                            timer.end(stream)
                            ms = timer.elapsed_ms()
                            best = ms if best is None else min(best, ms)
                        table[(w, g, s)] = best / max(st.samples, 1)
                    finally:
                        r.destroy()
    finally:
        timer.destroy()
    best = min(table, key=table.get)
    return best, table
