"""luce-render's host API over the HIP library: Device, GpuScene.upload,
ray_frame / look_frame, probe (src/render/device.lucb, camera.lucb).

    from luce_render import render
    device = render.Device.open()
    scene = render.compile_scene(set)
    on_gpu = render.GpuScene.upload(device, scene)
    frame = render.ray_frame(scene.cameras[0], 256, 256)
    probes = render.probe(device, on_gpu, frame)
"""
import ctypes
import math
from dataclasses import dataclass, replace

import numpy as np

from . import _lib
from ._lib import INVALID, GpuUnavailable, LuceError, TraceConfig, SceneInfo
from .scene import (NODE_DTYPE, Camera, CompiledScene, GeometrySet, Instance, Light, Mesh, Stopping, build_bvh,
                    compile_scene, octahedral, decode_octahedral)

__all__ = ["Device", "GpuScene", "RayFrame", "ray_frame", "look_frame", "probe", "Probe", "compile_scene",
           "build_bvh", "GeometrySet", "Mesh", "Camera", "Light", "Instance", "Stopping", "CompiledScene",
           "LuceError", "GpuUnavailable", "octahedral", "decode_octahedral"]


class Device:
    """gpu.Device: one GPU and its stream."""

    def __init__(self, handle, ordinal):
        self.handle = handle
        self.ordinal = ordinal

    @staticmethod
    def open(ordinal=0):
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().lr_device_open(int(ordinal), ctypes.byref(h)))
        return Device(h, ordinal)

    def destroy(self):
        if self.handle:
            _lib.lib().lr_device_close(self.handle)
            self.handle = None


class GpuScene:
    """GpuScene (device.lucb:16-37): the scene's buffers on a device."""

    def __init__(self, handle, device, triangles):
        self.handle = handle
        self.device = device
        self.triangles = triangles

    @staticmethod
    def _arrays(scene, with_nodes):
        arrays = _lib.SceneArrays()
        arrays.positions = _lib.ptr(scene.positions)
        arrays.points = scene.point_count()
        arrays.indices = _lib.ptr(scene.indices)
        arrays.faces = _lib.ptr(scene.faces)
        arrays.shading = _lib.ptr(scene.shading)
        arrays.uvs = _lib.ptr(scene.uvs)
        arrays.triangles = scene.triangle_count()
        if with_nodes:
            arrays.nodes = _lib.ptr(scene.nodes)
            arrays.node_count = len(scene.nodes)
        sp = getattr(scene, "splats", None)
        if sp is not None and len(sp.records) and sp.nodes is not None:
            arrays.splat_records, arrays.splats = _lib.ptr(sp.records), len(sp.records)
            arrays.splat_harmonics, arrays.splat_words = _lib.ptr(sp.harmonics), len(sp.harmonics)
            arrays.splat_nodes, arrays.splat_node_count = _lib.ptr(sp.nodes), len(sp.nodes)
        return arrays

    @staticmethod
    def upload(device, scene: CompiledScene, accelerate=True):
        """GpuScene.upload (device.lucb:41-89); `accelerate` is accepted for the
        reference's signature (there is no hardware ray tracing here)."""
        arrays = GpuScene._arrays(scene, True)
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().lr_scene_upload(device.handle, ctypes.byref(arrays), ctypes.byref(h)))
        return GpuScene(h, device, scene.triangle_count())

    @staticmethod
    def build(device, scene: CompiledScene, keep_tree=True):
        """build_bvh and GpuScene.upload in one step on the device, for a scene
        compiled with geometry=False (triangles in flattening order). With
        `keep_tree` the scene receives its nodes and leaf order and its
        per-triangle arrays are reordered, as build_bvh leaves them."""
        n = scene.triangle_count()
        arrays = GpuScene._arrays(scene, False)
        h = ctypes.c_void_p()
        nodes = np.zeros(max(1, 2 * n), NODE_DTYPE) if keep_tree else None
        order = np.zeros(max(1, n), np.uint32) if keep_tree else None
        count = ctypes.c_size_t(0)
        _lib.check(_lib.lib().lr_scene_build(device.handle, ctypes.byref(arrays), ctypes.byref(h),
                                             _lib.ptr(nodes) if keep_tree else None, ctypes.byref(count),
                                             _lib.ptr(order) if keep_tree else None))
        if keep_tree:
            scene.nodes = nodes[:count.value].copy()
            scene.order = order[:n].copy()
            if n:
                scene.indices = np.ascontiguousarray(scene.indices[scene.order])
                scene.faces = np.ascontiguousarray(scene.faces[scene.order])
                scene.shading = np.ascontiguousarray(scene.shading[scene.order])
                scene.uvs = np.ascontiguousarray(scene.uvs[scene.order])
        return GpuScene(h, device, n)

    def refit(self, positions):
        """Moves the scene's points (same count, same triangles) and refits
        its boxes on the device; the tree's topology stays."""
        positions = np.ascontiguousarray(positions, np.float32).reshape(-1, 3)
        _lib.check(_lib.lib().lr_scene_refit(self.handle, _lib.ptr(positions), len(positions)))

    def layout(self, width):
        """The width-`width` traversal nodes as uploaded (raw bytes)."""
        size = ctypes.c_size_t(0)
        _lib.check(_lib.lib().lr_scene_read_layout(self.handle, width, None, 0, ctypes.byref(size)))
        out = np.zeros(size.value, np.uint8)
        _lib.check(_lib.lib().lr_scene_read_layout(self.handle, width, _lib.ptr(out), size.value, ctypes.byref(size)))
        return out

    def set_trace_config(self, width=4, group=64, stack=24, refill=32):
        """BVH width, workgroup size, LDS stack entries per lane, and the ray
        schedule: refill 0 gives each thread one ray, refill k runs persistent
        waves that take new rays once k lanes are idle."""
        cfg = TraceConfig(width, group, stack, refill)
        _lib.check(_lib.lib().lr_scene_set_trace_config(self.handle, ctypes.byref(cfg)))

    def trace_config(self):
        cfg = TraceConfig()
        _lib.check(_lib.lib().lr_scene_get_trace_config(self.handle, ctypes.byref(cfg)))
        return (cfg.width, cfg.group, cfg.stack, cfg.refill)

    def resident_grid(self, rays, any_hit=False):
        """Workgroups a refilled launch of the current traversal variant takes
        over `rays` rays: that variant's resident workgroups, at most what the
        rays need."""
        blocks = ctypes.c_uint32()
        _lib.check(_lib.lib().lr_trace_resident_grid(self.handle, int(rays), 1 if any_hit else 0,
                                                     ctypes.byref(blocks)))
        return blocks.value

    def info(self):
        info = SceneInfo()
        _lib.check(_lib.lib().lr_scene_info_get(self.handle, ctypes.byref(info)))
        return {
            "triangles": info.triangles, "binary_nodes": info.binary_nodes,
            "wide_nodes": list(info.wide_nodes), "wide_bytes": list(info.wide_bytes),
            "triangle_bytes": info.triangle_bytes, "stack_bound": list(info.stack_bound),
        }

    def destroy(self):
        if self.handle:
            _lib.lib().lr_scene_destroy(self.handle)
            self.handle = None


@dataclass
class RayFrame:
    """RayFrame (camera.lucb:26-55), f64."""
    position: np.ndarray
    corner: np.ndarray
    du: np.ndarray
    dv: np.ndarray
    forward: np.ndarray
    orthographic: bool
    width: int
    height: int
    lens_radius: float = 0.0
    focus_distance: float = 1.0

    def pack(self):
        """The 16 floats the kernels read (camera.lucb:41-47)."""
        v = [*self.position, 1.0 if self.orthographic else 0.0, *self.corner, 0.0,
             *self.du, float(self.width), *self.dv, float(self.height)]
        return np.array(v, dtype=np.float64).astype(np.float32)

    def ray(self, x, y):
        """Origin and unnormalized direction through (x, y) (camera.lucb:51-55)."""
        film = self.corner + self.du * x + self.dv * y
        if self.orthographic:
            return self.position + film, self.forward.copy()
        return self.position.copy(), film


def _normalized(v):
    v = np.asarray(v, dtype=np.float64)
    n = math.sqrt(float(v @ v))
    if not n > 0.0:
        raise LuceError(INVALID, "a zero vector has no direction")
    return v / n


def ray_frame(camera: Camera, width, height):
    """ray_frame (camera.lucb:57-76): film fit horizontally, looking down -Z."""
    if not isinstance(camera, Camera):
        raise LuceError(INVALID, "a ray frame needs a camera")
    if width == 0 or height == 0:
        raise LuceError(INVALID, "an image needs pixels")
    m = np.asarray(camera.matrix, np.float64)
    right, up, back = _normalized(m[0, :3]), _normalized(m[1, :3]), _normalized(m[2, :3])
    forward = -back
    aspect = height / width
    ortho = camera.projection == "orthographic"
    half_width = 0.05 * camera.horizontal_aperture if ortho else 0.5 * camera.horizontal_aperture / camera.focal_length
    half_height = half_width * aspect
    corner = right * -half_width + up * half_height + (np.zeros(3) if ortho else forward)
    du = right * (2.0 * half_width / width)
    dv = up * (-2.0 * half_height / height)
    lens = 0.0005 * camera.focal_length / camera.f_stop if camera.f_stop > 0.0 and not ortho else 0.0
    return RayFrame(m[3, :3].copy(), corner, du, dv, forward, ortho, int(width), int(height), lens,
                    camera.focus_distance)


def look_frame(position, forward, up, tan_half_height, width, height, orthographic=False):
    """look_frame (camera.lucb:84-95)."""
    if width == 0 or height == 0:
        raise LuceError(INVALID, "an image needs pixels")
    if not (0.0 < tan_half_height < 1e6):
        raise LuceError(INVALID, "a view needs a field of view between 0 and 180 degrees, or a height")
    ahead = _normalized(forward)
    right = _normalized(np.cross(ahead, np.asarray(up, np.float64)))
    above = np.cross(right, ahead)
    half_height = tan_half_height
    half_width = half_height * width / height
    corner = right * -half_width + above * half_height + (np.zeros(3) if orthographic else ahead)
    du = right * (2.0 * half_width / width)
    dv = above * (-2.0 * half_height / height)
    return RayFrame(np.asarray(position, np.float64), corner, du, dv, ahead, bool(orthographic), int(width),
                    int(height), 0.0, 1.0)


@dataclass
class Probe:
    """One pixel of a probe (device.lucb:132-134): polygon (None: missed) and
    distance along the frame's unnormalized direction."""
    face: object
    distance: float


def probe_arrays(device, scene: GpuScene, frame: RayFrame):
    """The probe as arrays: faces u32 (0xffffffff on a miss) and distances f32,
    row by row from the top."""
    count = frame.width * frame.height
    faces = np.empty(count, np.uint32)
    dist = np.empty(count, np.float32)
    packed = frame.pack()
    _lib.check(_lib.lib().lr_probe(device.handle, scene.handle, _lib.ptr(packed), frame.width, frame.height,
                                   _lib.ptr(faces), _lib.ptr(dist), count))
    return faces, dist


def probe(device, scene: GpuScene, frame: RayFrame, out=None):
    """probe (device.lucb:138-160): one Probe per pixel into `out` (a list of
    width*height entries) or a new list."""
    count = frame.width * frame.height
    if out is not None and len(out) != count:
        raise LuceError(INVALID, "a probe needs one result per pixel")
    faces, dist = probe_arrays(device, scene, frame)
    result = out if out is not None else [None] * count
    for i in range(count):
        f = int(faces[i])
        result[i] = Probe(None if f == 0xFFFFFFFF else f, float(dist[i]))
    return result


class Buffer:
    """gpu.Buffer: device memory on a Device (allocate / create / read / upload)."""

    def __init__(self, device, nbytes):
        self.device = device
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        _lib.check(_lib.lib().lr_buffer_alloc(device.handle, self.nbytes, ctypes.byref(p)))
        self.ptr = p

    @staticmethod
    def create(device, array):
        array = np.ascontiguousarray(array)
        b = Buffer(device, array.nbytes)
        b.upload(array)
        return b

    def upload(self, array, offset=0):
        array = np.ascontiguousarray(array)
        if offset + array.nbytes > self.nbytes:
            raise LuceError(INVALID, "an upload past the buffer's end")
        _lib.check(_lib.lib().lr_buffer_write(self.device.handle, self.ptr.value + offset, _lib.ptr(array),
                                              array.nbytes))

    def fill(self, byte):
        """Sets every byte (an upload from the host)."""
        self.upload(np.full(self.nbytes, byte & 0xFF, np.uint8))

    def read(self, dtype, count=None, offset=0):
        dtype = np.dtype(dtype)
        if count is None:
            count = (self.nbytes - offset) // dtype.itemsize
        out = np.empty(count, dtype)
        if offset + out.nbytes > self.nbytes:
            raise LuceError(INVALID, "a read past the buffer's end")
        _lib.check(_lib.lib().lr_buffer_read(self.device.handle, _lib.ptr(out), self.ptr.value + offset, out.nbytes))
        return out

    def copy_from(self, other, nbytes=None):
        """A device-side copy of `other`'s first nbytes (the copy kernel)."""
        nbytes = min(self.nbytes, other.nbytes) if nbytes is None else nbytes
        _lib.check(_lib.lib().lr_buffer_copy(self.device.handle, self.ptr, other.ptr, int(nbytes)))

    def destroy(self):
        if self.ptr:
            _lib.lib().lr_buffer_free(self.device.handle, self.ptr)
            self.ptr = None


def device_stream(device):
    return _lib.lib().lr_device_stream(device.handle)


def synchronize(device):
    _lib.check(_lib.lib().lr_device_synchronize(device.handle))


def probe_device(scene: GpuScene, frame: RayFrame, hits: Buffer, stream=None):
    """The probe into a device buffer of width*height {polygon, t bits} pairs,
    on `stream` (default: the device's own stream, which Buffer reads and
    synchronize() wait on)."""
    packed = frame.pack()
    if stream is None:
        stream = device_stream(scene.device)
    if hits.nbytes < frame.width * frame.height * 8:
        raise LuceError(INVALID, "the probe's output buffer is too small")
    _lib.check(_lib.lib().lr_probe_device(scene.handle, _lib.ptr(packed), frame.width, frame.height, hits.ptr,
                                          stream))


def trace_closest_device(scene: GpuScene, rays: Buffer, n, hits: Buffer, stream=None):
    """Closest hits of n rays (8 floats each: origin, t_max, direction, -)."""
    if rays.nbytes < n * 32 or hits.nbytes < n * 8:
        raise LuceError(INVALID, "ray or hit buffer too small")
    if stream is None:
        stream = device_stream(scene.device)
    _lib.check(_lib.lib().lr_trace_closest_device(scene.handle, rays.ptr, int(n), hits.ptr, stream))


def trace_frame_and_closest_device(scene: GpuScene, frame: RayFrame, frame_hits: Buffer, rays: Buffer, n,
                                   hits: Buffer, stream=None):
    """probe_device of `frame` and trace_closest_device of n rays in one
    call: with the scene's refill above 0 one persistent launch traces both
    (one ramp-up, one drain); the same hits as the two calls."""
    packed = frame.pack()
    if stream is None:
        stream = device_stream(scene.device)
    if frame_hits.nbytes < frame.width * frame.height * 8 or rays.nbytes < n * 32 or hits.nbytes < n * 8:
        raise LuceError(INVALID, "a frame hit, ray or hit buffer too small")
    _lib.check(_lib.lib().lr_trace_frame_and_closest_device(scene.handle, _lib.ptr(packed), frame.width, frame.height,
                                                            frame_hits.ptr, rays.ptr, int(n), hits.ptr, stream))


def trace_closest_device_noalloc(scene: GpuScene, rays: Buffer, n, hits: Buffer, stream=None):
    """trace_closest_device without allocating device memory: the library's
    fixed scratch array holds the traversal's spill area and counters (calls
    on one device must not overlap)."""
    if rays.nbytes < n * 32 or hits.nbytes < n * 8:
        raise LuceError(INVALID, "ray or hit buffer too small")
    if stream is None:
        stream = device_stream(scene.device)
    _lib.check(_lib.lib().lr_trace_closest_device_noalloc(scene.handle, rays.ptr, int(n), hits.ptr, stream))


def trace_any_device(scene: GpuScene, rays: Buffer, n, occluded: Buffer, stream=None):
    """Any-hit occlusion of n rays (8 floats each) before their t_max."""
    if rays.nbytes < n * 32 or occluded.nbytes < n * 4:
        raise LuceError(INVALID, "ray or output buffer too small")
    if stream is None:
        stream = device_stream(scene.device)
    _lib.check(_lib.lib().lr_trace_any_device(scene.handle, rays.ptr, int(n), occluded.ptr, stream))


class Timer:
    """A pair of HIP events on a stream: elapsed milliseconds between marks."""

    def __init__(self):
        self.start = ctypes.c_void_p()
        self.stop = ctypes.c_void_p()
        _lib.check(_lib.lib().lr_event_create(ctypes.byref(self.start)))
        _lib.check(_lib.lib().lr_event_create(ctypes.byref(self.stop)))

    def begin(self, stream):
        _lib.check(_lib.lib().lr_event_record(self.start, stream))

    def end(self, stream):
        _lib.check(_lib.lib().lr_event_record(self.stop, stream))

    def elapsed_ms(self):
        ms = ctypes.c_float()
        _lib.check(_lib.lib().lr_event_elapsed(self.start, self.stop, ctypes.byref(ms)))
        return ms.value

    def destroy(self):
        _lib.lib().lr_event_destroy(self.start)
        _lib.lib().lr_event_destroy(self.stop)


# ------------------------------------------------------------------ renderer

from .loading import Loaded, OpenPbr, RenderSettings  # noqa: E402

KERNEL_NAMES = ["camera", "intersect closest", "shade surface", "shade miss", "intersect shadow", "light emission",
                "finish", "film convert", "adaptive", "later rounds"]


def find_camera(scene, index=-1):
    """find_camera (integrator.lucb:955-965): the settings' camera or the first."""
    if not scene.cameras:
        raise LuceError(INVALID, "the scene has no camera")
    if index < 0:
        return scene.cameras[0]
    if index >= len(scene.cameras):
        raise LuceError(INVALID, "no camera of that index")
    return scene.cameras[index]


class Renderer:
    """Renderer (integrator.lucb:259-903): create, advance, finish,
    samples_done, image, read, kernel_times, destroy."""

    def __init__(self):
        self.handle = None

    @staticmethod
    def create(device, scene, settings: RenderSettings, gpu_scene=None):
        """Renderer.create (integrator.lucb:343): through the settings' camera.
        `gpu_scene` reuses an uploaded scene (else the scene is uploaded)."""
        if settings.width == 0 or settings.height == 0 or settings.width > 16384 or settings.height > 16384:
            raise LuceError(INVALID, "an image is 1 to 16384 pixels on a side")
        frame = ray_frame(find_camera(scene, settings.camera), settings.width, settings.height)
        return Renderer.create_viewed(device, scene, settings, frame, gpu_scene)

    @staticmethod
    def create_viewed(device, scene, settings: RenderSettings, frame: RayFrame, gpu_scene=None):
        if settings.samples == 0 or settings.max_bounces == 0 or settings.max_bounces > 1024:
            raise LuceError(INVALID, "a render needs samples and 1 to 1024 bounces")
        r = Renderer()
        r.device = device
        r.scene = scene
        r.owned_scene = gpu_scene is None
        r.gpu_scene = gpu_scene if gpu_scene is not None else GpuScene.upload(device, scene)
        ld, cs = r._prepare(scene, replace(settings, width=frame.width, height=frame.height), frame)
        h = ctypes.c_void_p()
        try:
            _lib.check(_lib.lib().lr_renderer_create(device.handle, r.gpu_scene.handle, ctypes.byref(ld),
                                                     ctypes.byref(cs), _lib.ptr(r.frame16), ctypes.byref(h)))
        except Exception:
            if r.owned_scene:
                r.gpu_scene.destroy()
            raise
        r.handle = h
        return r

    def _prepare(self, scene, settings, frame):
        """The loaded arrays and settings the C ABI takes (kept alive on self)."""
        self.settings = settings
        self.frame = frame
        self.loaded = Loaded(scene, settings, frame)
        ld = _lib.Loaded()
        ld.lights, ld.light_floats = _lib.ptr(self.loaded.lights), len(self.loaded.lights)
        ld.constants = _lib.ptr(self.loaded.constants)
        ld.materials, ld.material_floats = _lib.ptr(self.loaded.materials), len(self.loaded.materials)
        ld.emitters, ld.emitter_words = _lib.ptr(self.loaded.emitters), len(self.loaded.emitters)
        ld.light_paths, ld.through = self.loaded.light_paths, self.loaded.through
        ld.programs, ld.program_words = _lib.ptr(self.loaded.programs), len(self.loaded.programs)
        ld.textures, ld.texture_floats = _lib.ptr(self.loaded.textures), len(self.loaded.textures)
        ld.texture_table, ld.texture_count = _lib.ptr(self.loaded.texture_table), self.loaded.texture_count
        s = settings
        cs = _lib.RenderSettingsC(s.width, s.height, s.samples, s.max_bounces, 1 if s.spectral else 0, s.pool, s.seed,
                                  int(s.guiding), 1 if s.profile else 0, 1 if s.use_graph else 0,
                                  *tuple(s.trace)[:3], s.train_samples, s.train_steps, int(s.pass_samples))
        self.frame16 = frame.pack()
        return ld, cs

    def _reset(self, scene, gpu_scene, settings, frame):
        old = (self.scene, self.settings, self.frame, self.loaded, self.frame16)
        ld, cs = self._prepare(scene, settings, frame)
        try:
            _lib.check(_lib.lib().lr_renderer_reset(self.handle, gpu_scene.handle if gpu_scene else None,
                                                    ctypes.byref(ld), ctypes.byref(cs), _lib.ptr(self.frame16)))
        except Exception:
            self.scene, self.settings, self.frame, self.loaded, self.frame16 = old
            raise
        self.scene = scene

    def restart(self, frame: RayFrame):
        """Renderer.restart (integrator.lucb:463): start over through `frame`
        (a moved camera, a resized view), keeping the projection."""
        if frame.width == 0 or frame.height == 0 or frame.width > 16384 or frame.height > 16384:
            raise LuceError(INVALID, "an image is 1 to 16384 pixels on a side")
        if frame.orthographic != self.frame.orthographic:
            raise LuceError(INVALID, "a restart keeps the view's projection")
        self._reset(self.scene, None, replace(self.settings, width=frame.width, height=frame.height), frame)

    def reload(self, scene, settings: RenderSettings, frame: RayFrame = None):
        """Renderer.reload (integrator.lucb:488): the lights, materials and
        constants of `scene` (the same triangles), `settings` but for the
        size and the pool; the film starts over, through `frame` when given."""
        if settings.samples == 0 or settings.max_bounces == 0 or settings.max_bounces > 1024:
            raise LuceError(INVALID, "a render needs samples and 1 to 1024 bounces")
        if scene.triangle_count() != self.scene.triangle_count():
            raise LuceError(INVALID, "a reload keeps the scene's geometry")
        frame = frame or self.frame
        chosen = replace(settings, width=frame.width, height=frame.height, pool=self.settings.pool)
        self._reset(scene, None, chosen, frame)

    def replace_scene(self, scene, settings: RenderSettings, frame: RayFrame):
        """Renderer.replace_scene (integrator.lucb:542): new geometry in place
        of this one's (uploaded here), keeping the pool."""
        if settings.samples == 0 or settings.max_bounces == 0 or settings.max_bounces > 1024:
            raise LuceError(INVALID, "a render needs samples and 1 to 1024 bounces")
        if settings.pool != self.settings.pool:
            raise LuceError(INVALID, "a scene replaced keeps the pool")
        uploaded = GpuScene.upload(self.device, scene)
        try:
            self._reset(scene, uploaded, replace(settings, width=frame.width, height=frame.height), frame)
        except Exception:
            uploaded.destroy()
            raise
        if self.owned_scene:
            self.gpu_scene.destroy()
        self.gpu_scene = uploaded
        self.owned_scene = True

    def advance(self):
        busy = ctypes.c_int(0)
        _lib.check(_lib.lib().lr_renderer_advance(self.handle, ctypes.byref(busy)))
        return bool(busy.value)

    def finish(self):
        _lib.check(_lib.lib().lr_renderer_finish(self.handle))

    def samples_done(self):
        return int(_lib.lib().lr_renderer_samples_done(self.handle))

    def image(self):
        im = _lib.Image()
        _lib.check(_lib.lib().lr_renderer_image(self.handle, ctypes.byref(im)))
        return im

    def read(self, out=None):
        """The image as rgba32f (height, width, 4), row by row from the top."""
        n = self.settings.width * self.settings.height * 4
        if out is not None and out.size != n:
            raise LuceError(INVALID, "reading the image needs four floats a pixel")
        buf = np.empty(n, np.float32) if out is None else out
        _lib.check(_lib.lib().lr_renderer_read(self.handle, _lib.ptr(buf), n))
        return buf.reshape(self.settings.height, self.settings.width, 4)

    def read_display(self, out=None):
        """The display image (Renderer.display, integrator.lucb:743) as rgba32f
        (height, width, 4): exposed, linear Rec.709, through the view."""
        n = self.settings.width * self.settings.height * 4
        if out is not None and out.size != n:
            raise LuceError(INVALID, "reading the display image needs four floats a pixel")
        buf = np.empty(n, np.float32) if out is None else out
        _lib.check(_lib.lib().lr_renderer_read_display(self.handle, _lib.ptr(buf), n))
        return buf.reshape(self.settings.height, self.settings.width, 4)

    def set_display(self, exposure: float, view: int = 0):
        """Renderer.set_display (integrator.lucb:750-756): `exposure` stops,
        `view` 0 standard or 1 neutral tone mapping; the next pass, or once
        every sample is in a conversion of its own, applies it."""
        _lib.check(_lib.lib().lr_renderer_set_display(self.handle, float(exposure), int(view)))

    def read_film(self):
        n = self.settings.width * self.settings.height * 16
        buf = np.empty(n, np.float32)
        _lib.check(_lib.lib().lr_renderer_read_film(self.handle, _lib.ptr(buf), n))
        return buf.reshape(4, self.settings.height, self.settings.width, 4)

    GUIDE_CELLS = 524288

    def read_guide(self):
        """The hash-grid guide (guiding 1): keys (cells,) u32, lobes (cells, 5,
        4) f32 (four lobes' mean and kappa, then their weights) and statistics
        (cells, 4, 5) f32 (weight, weighted direction, count per lobe)."""
        n = self.GUIDE_CELLS
        buf = np.empty(n * 164 // 4, np.uint32)
        _lib.check(_lib.lib().lr_renderer_read_guide(self.handle, _lib.ptr(buf), buf.nbytes))
        keys = buf[:n].copy()
        lobes = buf[n:n + n * 20].view(np.float32).reshape(n, 5, 4).copy()
        stats = buf[n + n * 20:].view(np.float32).reshape(n, 4, 5).copy()
        return keys, lobes, stats

    def kernel_times(self):
        out = np.zeros(10, np.float64)
        _lib.check(_lib.lib().lr_renderer_kernel_times(self.handle, _lib.ptr(out)))
        return dict(zip(KERNEL_NAMES, out.tolist()))

    def converged_pixels(self):
        """Renderer.converged_pixels (integrator.lucb:759)."""
        out = ctypes.c_size_t()
        _lib.check(_lib.lib().lr_renderer_converged_pixels(self.handle, ctypes.byref(out)))
        return out.value

    def stats(self):
        st = _lib.RenderStats()
        _lib.check(_lib.lib().lr_renderer_stats(self.handle, ctypes.byref(st)))
        return {"samples_done": st.samples_done, "light_samples": st.light_samples,
                "closest_rays": st.closest_rays, "shadow_rays": st.shadow_rays}

    def destroy(self):
        if self.handle:
            _lib.lib().lr_renderer_destroy(self.handle)
            self.handle = None
        if getattr(self, "owned_scene", False) and self.gpu_scene is not None:
            self.gpu_scene.destroy()
            self.gpu_scene = None


def render(device, scene, settings, gpu_scene=None):
    """Create, finish and read a render (tests/gpu/integrator.lucb:69-76)."""
    r = Renderer.create(device, scene, settings, gpu_scene)
    try:
        r.finish()
        if r.samples_done() != settings.samples:
            raise LuceError(INVALID, f"finished with {r.samples_done()} of {settings.samples} samples")
        return r.read().copy()
    finally:
        r.destroy()
