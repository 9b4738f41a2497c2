"""The serving form: a request loop over one GPU with continuous batching.

Requests (a scene, a resolution, a sample target) arrive over time; the
server admits them in arrival order as their device memory fits its budget,
keeps their path state and films resident, and advances every admitted
request by one sample per step in shared kernel launches, so work that
arrives while others render joins the next step. Reports time to the
target sample count per request (p50/p99) and aggregate rays per second.

    server = Server(device, budget_bytes=200 << 30)
    rid = server.submit(scene_handle, compiled_scene, settings)
    while server.busy(): server.step()
"""
import ctypes
import dataclasses
import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib, render
from .loading import Loaded, RenderSettings
from .scenes import Xoshiro256


class Server:
    def __init__(self, device, budget_bytes=200 << 30, spectral=False, trace=(20, 64, 16), max_active=0):
        self.device = device
        cfg = _lib.ServerSettings(int(budget_bytes), 1 if spectral else 0, trace[0], trace[1], trace[2], max_active)
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().lr_server_create(device.handle, ctypes.byref(cfg), ctypes.byref(h)))
        self.handle = h
        self.spectral = spectral
        self.active = 0
        self.pending = 0
        self._keep = {}

    def submit(self, gpu_scene, scene, settings: RenderSettings, frame=None):
        """Queues a render; returns its id. `scene` is the compiled scene the
        loaded arrays are packed from, `gpu_scene` its upload."""
        if frame is None:
            frame = render.ray_frame(render.find_camera(scene, settings.camera), settings.width, settings.height)
        settings = dataclasses.replace(settings, width=frame.width, height=frame.height)
        loaded = Loaded(scene, settings, frame)
        ld = _lib.Loaded()
        ld.lights, ld.light_floats = _lib.ptr(loaded.lights), len(loaded.lights)
        ld.constants = _lib.ptr(loaded.constants)
        ld.materials, ld.material_floats = _lib.ptr(loaded.materials), len(loaded.materials)
        ld.emitters, ld.emitter_words = _lib.ptr(loaded.emitters), len(loaded.emitters)
        ld.light_paths, ld.through = loaded.light_paths, loaded.through
        ld.programs, ld.program_words = _lib.ptr(loaded.programs), len(loaded.programs)
        ld.textures, ld.texture_floats = _lib.ptr(loaded.textures), len(loaded.textures)
        ld.texture_table, ld.texture_count = _lib.ptr(loaded.texture_table), loaded.texture_count
        s = settings
        cs = _lib.RenderSettingsC(s.width, s.height, s.samples, s.max_bounces, 1 if s.spectral else 0, 0, s.seed,
                                  int(s.guiding), 0, 0, 0, 0, 0, 0, 0, int(s.pass_samples))
        packed = frame.pack()
        rid = ctypes.c_uint64()
        _lib.check(_lib.lib().lr_server_submit(self.handle, gpu_scene.handle, ctypes.byref(ld), ctypes.byref(cs),
                                               _lib.ptr(packed), ctypes.byref(rid)))
        self.pending += 1
        return rid.value

    def step(self):
        a, p = ctypes.c_uint32(), ctypes.c_uint32()
        _lib.check(_lib.lib().lr_server_step(self.handle, ctypes.byref(a), ctypes.byref(p)))
        self.active, self.pending = a.value, p.value
        return self.active, self.pending

    def busy(self):
        return self.active > 0 or self.pending > 0

    def status(self, rid):
        st = _lib.RequestStatus()
        _lib.check(_lib.lib().lr_server_status(self.handle, rid, ctypes.byref(st)))
        return {"state": st.state, "samples_done": st.samples_done, "samples": st.samples, "bytes": st.bytes}

    def read(self, rid, width, height):
        buf = np.empty(width * height * 4, np.float32)
        _lib.check(_lib.lib().lr_server_read(self.handle, rid, _lib.ptr(buf), buf.size))
        return buf.reshape(height, width, 4)

    def stats(self, rid):
        st = _lib.RenderStats()
        _lib.check(_lib.lib().lr_server_stats(self.handle, rid, ctypes.byref(st)))
        return {"closest_rays": st.closest_rays, "shadow_rays": st.shadow_rays, "samples_done": st.samples_done}

    def release(self, rid):
        _lib.check(_lib.lib().lr_server_release(self.handle, rid))

    def destroy(self):
        if self.handle:
            _lib.lib().lr_server_destroy(self.handle)
            self.handle = None


RESOLUTIONS = [(256, 256), (512, 512), (1024, 768), (1920, 1080)]


def prepare_scenes(device, gsets):
    """The served scenes on `device`, [(gpu_scene, compiled scene)]: a scene
    whose triangles overlap each other's boxes heavily (scene.box_overlap:
    the random-triangle clouds) is built over pieces of its triangles on the
    device (scene.split_bvh, the cuts scene.auto_split_depth picks: the same
    hits, a 1.8x to 3.9x faster render); the others are built from their
    flattened triangles on the device (GpuScene.build: the reference's
    tree)."""
    from . import scene as S
    out = []
    for g in gsets:
        sc = render.compile_scene(g, geometry=False)
        if S.auto_split_depth(sc) > 0:
            sc = render.compile_scene(g, split_depth=-1, device=device)
            out.append((render.GpuScene.upload(device, sc), sc))
        else:
            out.append((render.GpuScene.build(device, sc), sc))
    return out


@dataclass
class Arrival:
    time: float
    scene: int
    width: int
    height: int
    spp: int


def poisson_schedule(count=2000, rate=40.0, scenes=4, seed=11, spp_range=(16, 256), resolutions=RESOLUTIONS):
    """`count` requests with exponential inter-arrival times of mean 1/rate,
    a scene uniform over `scenes`, a resolution uniform over `resolutions`
    and a sample target log-uniform over `spp_range` (rounded)."""
    rng = Xoshiro256(seed)
    t = 0.0
    out = []
    lo, hi = math.log(spp_range[0]), math.log(spp_range[1])
    for _ in range(count):
        t += -math.log(1.0 - rng.uniform()) / rate
        scene = min(int(rng.uniform() * scenes), scenes - 1)
        w, h = resolutions[min(int(rng.uniform() * len(resolutions)), len(resolutions) - 1)]
        spp = int(round(math.exp(lo + rng.uniform() * (hi - lo))))
        out.append(Arrival(t, scene, w, h, spp))
    return out


@dataclass
class ServingReport:
    requests: int
    p50_s: float
    p99_s: float
    p99_to_64spp_s: float
    mean_s: float
    rays: int
    wall_s: float
    rays_per_s: float
    peak_active: int
    latencies: list = field(default_factory=list)


def run(device, scenes, schedule, server_kw=None, max_bounces=8, spectral=False, light_tracing=0.5, log=None,
        time_scale=1.0):
    """Drives `schedule` (arrival times in seconds, scaled by `time_scale`)
    through a Server over `scenes` [(gpu_scene, compiled scene)], in real
    time: a request is submitted once the wall clock passes its arrival.
    Returns a ServingReport; latency is arrival to final sample."""
    server = Server(device, spectral=spectral, **(server_kw or {}))
    start = time.perf_counter()
    submitted = {}
    done = {}
    reached = {}  # time each request with a target of 64 or more first had 64 samples
    open_ids = set()
    at = 0
    rays = 0
    peak = 0
    try:
        while at < len(schedule) or server.busy():
            now = time.perf_counter() - start
            while at < len(schedule) and schedule[at].time * time_scale <= now:
                a = schedule[at]
                gpu_scene, sc = scenes[a.scene]
                st = RenderSettings(a.width, a.height, a.spp, max_bounces=max_bounces, spectral=spectral,
                                    light_tracing=light_tracing)
                rid = server.submit(gpu_scene, sc, st)
                submitted[rid] = (a, a.time * time_scale)
                open_ids.add(rid)
                at += 1
            if not server.busy():
                # idle until the next arrival
                if at < len(schedule):
                    time.sleep(max(0.0, schedule[at].time * time_scale - (time.perf_counter() - start)))
                continue
            server.step()
            peak = max(peak, server.active)
            t_now = time.perf_counter() - start
            for rid in list(open_ids):
                st = server.status(rid)
                if rid not in reached and st["samples"] >= 64 and st["samples_done"] >= 64:
                    reached[rid] = t_now - submitted[rid][1]
                if st["state"] == 2:
                    s = server.stats(rid)
                    rays += s["closest_rays"] + s["shadow_rays"]
                    done[rid] = t_now - submitted[rid][1]
                    server.release(rid)
                    open_ids.discard(rid)
                    if log and len(done) % 50 == 0:
                        log(f"serving: {len(done)} of {len(schedule)} done, active {server.active}, "
                            f"pending {server.pending}")
        wall = time.perf_counter() - start
    finally:
        server.destroy()
    lat = np.array([done[r] for r in sorted(done)])
    to64 = np.array([reached[r] for r in sorted(reached)])
    return ServingReport(len(done), float(np.percentile(lat, 50)), float(np.percentile(lat, 99)),
                         float(np.percentile(to64, 99)) if len(to64) else float("nan"), float(lat.mean()), int(rays),
                         wall, rays / wall, peak, lat.tolist())
