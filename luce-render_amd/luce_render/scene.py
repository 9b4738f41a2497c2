"""Scene description and compilation: luce-render's compile_scene
(src/render/scene.lucb:111-129, 227-295, 400-412) over a small scene model.

The reference takes a luce-geocore GeometrySet; this module's GeometrySet
holds the same things the compile reads: meshes (points, triangles, each
triangle's polygon within its mesh, optional per-corner normals, per-polygon
material names, optional per-corner uvs), instances of other sets under a
placement, and scene objects (cameras, lights). Compilation flattens visible
instances into one world-space triangle list (the reference realizes
instances, scene.lucb:24-25), offsets polygon ids by the polygons of earlier
meshes, encodes shading normals octahedrally, and builds the BVH, reordering
every per-triangle array into leaf order (bvh.lucb:198-217).
"""
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import _lib
from ._lib import INVALID, LIMIT, STOPPED, LuceError
from .splats import CompiledSplats

IDENTITY = np.eye(4)


@dataclass
class Mesh:
    """A triangle mesh: points (P,3), triangles (T,3) point indices, each
    triangle's polygon (T,) in 0..face_count-1, optional corner normals
    (T,3,3), optional polygon materials (face_count,) of names, optional
    corner uvs (T,3,2)."""
    points: np.ndarray
    triangles: np.ndarray
    triangle_faces: Optional[np.ndarray] = None
    face_count: Optional[int] = None
    normals: Optional[np.ndarray] = None
    face_materials: Optional[list] = None
    uvs: Optional[np.ndarray] = None

    def __post_init__(self):
        self.points = np.asarray(self.points, dtype=np.float64).reshape(-1, 3)
        self.triangles = np.asarray(self.triangles, dtype=np.int64).reshape(-1, 3)
        if self.triangle_faces is None:
            self.triangle_faces = np.arange(len(self.triangles), dtype=np.int64)
        self.triangle_faces = np.asarray(self.triangle_faces, dtype=np.int64)
        if self.face_count is None:
            self.face_count = int(self.triangle_faces.max()) + 1 if len(self.triangle_faces) else 0


@dataclass
class Camera:
    """A USD camera: `matrix` a 4x4 placement (columns x, y, z axes and
    translation, row-vector convention: rows 0..2 axes, row 3 translation)."""
    matrix: np.ndarray
    projection: str = "perspective"
    focal_length: float = 50.0
    horizontal_aperture: float = 36.0
    f_stop: float = 0.0
    focus_distance: float = 1.0


@dataclass
class Light:
    """A UsdLux light (lights.lucb:29-131): kind rect/disk/sphere/distant/dome,
    placed by `matrix` (rows: axes, translation). A dome may carry `texture`:
    a lat-long image (height x width x 3, or height x width, linear Rec.709,
    rows from the top) wrapped around the scene (environment.lucb); the
    reference reads it from OpenEXR, this harness takes the pixels."""
    kind: str
    matrix: np.ndarray = field(default_factory=lambda: np.eye(4))
    color: tuple = (1.0, 1.0, 1.0)
    intensity: float = 1.0
    exposure: float = 0.0
    normalize: bool = False
    width: float = 1.0
    height: float = 1.0
    radius: float = 0.5
    angle: float = 0.53
    visible: bool = True
    temperature: float = 0.0
    texture: Optional[np.ndarray] = None


@dataclass
class SplatCloud:
    """A point cloud with luce-geocore's splat attributes (splats.lucb:5-13):
    points (N,3) about `origin`; orient (N,4) xyzw, scale (N,3) and opacity
    (N,) make it a splat cloud (without them it is a plain point cloud, not
    rendered); colour Cd (N,3) linear (0.5 when absent); `sh` (N, items, 3)
    with items 0, 3, 8 or 15 (degrees 0 to 3); optional restorient (N,4), the
    orientation the SH were authored at; `color_space` "linear" or "srgb",
    the encoding the DC colour and SH are authored in."""
    points: np.ndarray
    orient: Optional[np.ndarray] = None
    scale: Optional[np.ndarray] = None
    opacity: Optional[np.ndarray] = None
    color: Optional[np.ndarray] = None
    sh: Optional[np.ndarray] = None
    restorient: Optional[np.ndarray] = None
    color_space: str = "linear"
    origin: tuple = (0.0, 0.0, 0.0)


@dataclass
class Instance:
    set: "GeometrySet"
    matrix: np.ndarray
    visible: bool = True


@dataclass
class GeometrySet:
    meshes: List[Mesh] = field(default_factory=list)
    instances: List[Instance] = field(default_factory=list)
    cameras: List[Camera] = field(default_factory=list)
    lights: List[Light] = field(default_factory=list)
    materials: dict = field(default_factory=dict)  # name -> material record
    clouds: List[SplatCloud] = field(default_factory=list)


class Stopping:
    """A request to stop a compile (scene.lucb:143-149)."""

    def __init__(self):
        self.asked = False

    def ask(self):
        self.asked = True


class CompiledScene:
    """The kernels' layout (scene.lucb:31-99): positions f32 (P,3), indices u32
    (T,3) in leaf order, faces u32 (T,), shading u32 (T,4), uvs f32 (T,6),
    nodes (N,) of the 32-byte BvhNode, order u32 (T,)."""

    def __init__(self):
        self.positions = np.zeros((0, 3), np.float32)
        self.indices = np.zeros((0, 3), np.uint32)
        self.faces = np.zeros(0, np.uint32)
        self.shading = np.zeros((0, 4), np.uint32)
        self.uvs = np.zeros((0, 6), np.float32)
        self.nodes = np.zeros(0, NODE_DTYPE)
        self.order = np.zeros(0, np.uint32)
        self.cameras = []
        self.lights = []
        self.materials = []      # material records in index order (index + 1 on triangles)
        self.material_names = []
        self.skipped = 0
        self.split_pieces = 0    # leaf slots of a build over pieces of triangles (0: whole triangles)
        self.splats = CompiledSplats()

    def triangle_count(self):
        return len(self.indices)

    def point_count(self):
        return len(self.positions)


NODE_DTYPE = np.dtype([("lo", "<f4", 3), ("first", "<u4"), ("hi", "<f4", 3), ("count", "<u4")])


def octahedral(n):
    """scene.lucb:400-412 over an (M,3) array of f64 directions: two snorm16
    octahedral coordinates, x in the low half."""
    n = np.asarray(n, dtype=np.float64).reshape(-1, 3)
    length = np.abs(n).sum(axis=1)
    zero = length <= 0.0
    n = np.where(zero[:, None], np.array([0.0, 1.0, 0.0]), n)
    length = np.where(zero, 1.0, length)
    x = n[:, 0] / length
    y = n[:, 1] / length
    neg = n[:, 2] < 0.0
    fx = (1.0 - np.abs(y)) * np.where(x >= 0.0, 1.0, -1.0)
    fy = (1.0 - np.abs(x)) * np.where(y >= 0.0, 1.0, -1.0)
    x = np.where(neg, fx, x)
    y = np.where(neg, fy, y)
    # math.round: half away from zero
    qx = np.trunc(np.clip(x, -1.0, 1.0) * 32767.0 + np.copysign(0.5, x)).astype(np.int64)
    qy = np.trunc(np.clip(y, -1.0, 1.0) * 32767.0 + np.copysign(0.5, y)).astype(np.int64)
    return ((qx & 0xFFFF) | ((qy & 0xFFFF) << 16)).astype(np.uint32)


def decode_octahedral(packed):
    """The kernels' decode (common.glsl:262-267), in f64."""
    packed = np.asarray(packed, dtype=np.uint32)
    x = np.maximum(-1.0, (packed & 0xFFFF).astype(np.uint16).view(np.int16) / 32767.0)
    y = np.maximum(-1.0, (packed >> 16).astype(np.uint16).view(np.int16) / 32767.0)
    z = 1.0 - np.abs(x) - np.abs(y)
    neg = z < 0.0
    nx = np.where(neg, (1.0 - np.abs(y)) * np.where(x >= 0.0, 1.0, -1.0), x)
    ny = np.where(neg, (1.0 - np.abs(x)) * np.where(y >= 0.0, 1.0, -1.0), y)
    v = np.stack([nx, ny, z], axis=-1)
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


def smooth_normals(points, triangles):
    """Area-weighted point normals of a mesh (the mesh's smooth normals when it
    authors none)."""
    p = points[triangles]
    face_n = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
    acc = np.zeros_like(points)
    for k in range(3):
        np.add.at(acc, triangles[:, k], face_n)
    length = np.linalg.norm(acc, axis=1, keepdims=True)
    acc = np.where(length > 0, acc / np.where(length > 0, length, 1.0), np.array([0.0, 1.0, 0.0]))
    return acc


class _Builder:
    def __init__(self, scene, stopping, splats=True):
        self.scene = scene
        self.stopping = stopping
        self.splats = splats
        self.positions = []
        self.indices = []
        self.faces = []
        self.shading = []
        self.uvs = []
        self.points = 0
        self.face_base = 0

    def check_stop(self):
        if self.stopping is not None and self.stopping.asked:
            raise LuceError(STOPPED, "the compile was stopped")

    def add_set(self, gset, matrix, depth):
        """scene.lucb:227-243: meshes, then visible instances recursively."""
        if depth > 64:
            raise LuceError(INVALID, "instances nest more than 64 deep")
        for mesh in gset.meshes:
            self.check_stop()
            self.add_mesh(mesh, matrix)
        if self.splats:
            for cloud in gset.clouds:
                self.check_stop()
                self.scene.splats.add_cloud(cloud, matrix)
        for inst in gset.instances:
            self.check_stop()
            if not inst.visible:
                continue
            # matrix.multiply(item.matrix()): the instance's placement inside
            # the parent's (row-vector convention: child first).
            self.add_set(inst.set, np.asarray(inst.matrix, np.float64) @ matrix, depth + 1)

    def add_mesh(self, mesh, matrix):
        """scene.lucb:246-295."""
        m = np.asarray(matrix, dtype=np.float64)
        placed = mesh.points @ m[:3, :3] + m[3, :3]
        base = self.points
        self.positions.append(placed.astype(np.float32))
        tri = mesh.triangles
        count = len(tri)
        self.indices.append((tri + base).astype(np.uint32))
        # Normals by the placement's cofactors (inverse transpose up to scale),
        # sign kept for mirroring placements (scene.lucb:358-369).
        a = m[:3, :3]
        det = np.linalg.det(a)
        sign = 1.0 if det >= 0.0 else -1.0
        singular = det == 0.0 or not np.isfinite(det)
        c = np.stack([np.cross(a[1], a[2]), np.cross(a[2], a[0]), np.cross(a[0], a[1])]) * sign
        if mesh.normals is not None:
            corner_n = np.asarray(mesh.normals, np.float64).reshape(count, 3, 3)
        else:
            # smooth normals depend on the mesh alone: computed once per mesh
            sn = getattr(mesh, "_smooth", None)
            if sn is None:
                sn = smooth_normals(mesh.points, tri)
                mesh._smooth = sn
            corner_n = sn[tri]
        flat = corner_n.reshape(-1, 3)
        world_n = np.tile([0.0, 1.0, 0.0], (len(flat), 1)) if singular else flat @ c
        enc = octahedral(world_n).reshape(count, 3)
        # Materials: index + 1 in the scene's list by name, else 0.
        mats = np.zeros(count, np.uint32)
        if mesh.face_materials is not None:
            lookup = np.array([self.material_index(name) for name in mesh.face_materials], np.uint32)
            mats = lookup[mesh.triangle_faces]
        sh = np.empty((count, 4), np.uint32)
        sh[:, :3] = enc
        sh[:, 3] = mats
        self.shading.append(sh)
        if mesh.uvs is not None:
            self.uvs.append(np.asarray(mesh.uvs, np.float64).reshape(count, 6).astype(np.float32))
        else:
            self.uvs.append(np.zeros((count, 6), np.float32))
        self.faces.append((mesh.triangle_faces + self.face_base).astype(np.uint32))
        self.face_base += int(mesh.face_count)
        self.points += len(mesh.points)

    def material_index(self, name):
        try:
            return self.scene.material_names.index(name) + 1
        except ValueError:
            return 0


def _flatten_objects(gset, matrix, scene, depth):
    """Cameras and lights placed in world space, through visible instances."""
    if depth > 64:
        raise LuceError(INVALID, "instances nest more than 64 deep")
    for cam in gset.cameras:
        c = Camera(**{**cam.__dict__})
        c.matrix = np.asarray(cam.matrix, np.float64) @ matrix
        scene.cameras.append(c)
    for light in gset.lights:
        l = Light(**{**light.__dict__})
        l.matrix = np.asarray(light.matrix, np.float64) @ matrix
        scene.lights.append(l)
    for inst in gset.instances:
        if inst.visible:
            _flatten_objects(inst.set, np.asarray(inst.matrix, np.float64) @ matrix, scene, depth + 1)


def _flatten_materials(gset, scene, depth=0):
    for name, record in gset.materials.items():
        if name not in scene.material_names:
            scene.material_names.append(name)
            scene.materials.append(record)
    for inst in gset.instances:
        _flatten_materials(inst.set, scene, depth + 1)


def compile_scene(gset, geometry=True, stopping=None, threads=0, split_depth=0, device=None):
    """compile_scene (scene.lucb:111-129): the set's triangles in world space
    with their BVH, its splats with theirs, and its cameras, lights and
    materials. With geometry=False: triangles and materials, no BVH, no
    splats. `split_depth` > 0 builds over pieces of triangles (split_bvh; not
    the reference's tree, the same hits); -1 picks by the triangles' box
    overlap (auto_split_depth); `device` builds that tree on its GPU."""
    scene = CompiledScene()
    _flatten_materials(gset, scene)
    b = _Builder(scene, stopping, splats=geometry)
    b.add_set(gset, IDENTITY, 0)
    if b.points > 4294967295:
        raise LuceError(LIMIT, "a scene holds at most 2^32 points")
    _flatten_objects(gset, IDENTITY, scene, 0)
    if b.positions:
        scene.positions = np.ascontiguousarray(np.concatenate(b.positions))
        scene.indices = np.ascontiguousarray(np.concatenate(b.indices))
        scene.faces = np.ascontiguousarray(np.concatenate(b.faces))
        scene.shading = np.ascontiguousarray(np.concatenate(b.shading))
        scene.uvs = np.ascontiguousarray(np.concatenate(b.uvs))
    if geometry:
        b.check_stop()
        if split_depth < 0:
            split_depth = auto_split_depth(scene)
        if split_depth > 0:
            split_bvh(scene, split_depth, threads=threads, device=device)
        else:
            build_bvh(scene, threads=threads)
        b.check_stop()
        scene.splats.finish(threads=threads)
    return scene


def box_overlap(scene):
    """How many triangle boxes cover a point of the scene's box on average:
    the triangles' box volumes summed over the scene box's volume (a few for
    ordinary meshes; tens and more for dense clouds of random triangles,
    where every ray tests hundreds of triangles)."""
    if scene.triangle_count() == 0:
        return 0.0
    tri = scene.positions[scene.indices].astype(np.float64)
    lo, hi = tri.min(1), tri.max(1)
    extent = hi.max(0) - lo.min(0)
    volume = float(np.prod(np.maximum(extent, 1e-30)))
    return float(np.prod(hi - lo, axis=1).sum() / volume)


def auto_split_depth(scene):
    """Cuts for split_bvh by box_overlap, as measured on clouds of random
    triangles in the Cornell box (1080p, 8 bounces, light tracing): overlap
    8 (500K triangles) renders 24.3 -> 13.3 ms a sample at 2 cuts (3 and 4
    are slower), overlap 82 (5M) 74.5 -> 19.2 ms at 2 and 18.9 at 3; below
    4, whole triangles."""
    overlap = box_overlap(scene)
    return 3 if overlap > 40.0 else 2 if overlap > 4.0 else 0


def _splittable(scene):
    """Triangles whose pieces may go to several leaves: those of an explicit,
    opaque, non-emissive material without a node graph (a duplicate of an
    emitter would be a second light, of a cut-out a second pass-through)."""
    from .loading import OpenPbr
    ok = np.zeros(len(scene.materials) + 1, np.uint8)
    for k, record in enumerate(scene.materials):
        m = OpenPbr.of(record)
        ok[k + 1] = m.emission_luminance <= 0.0 and m.geometry_opacity >= 1.0 and not m.shader_graph
    mats = scene.shading[:, 3].astype(np.int64)
    return np.ascontiguousarray(np.where(mats < len(ok), ok[np.minimum(mats, len(ok) - 1)], 0).astype(np.uint8))


def split_bvh(scene, depth, extent=None, threads=0, device=None):
    """A BVH over pieces of triangles (no reference counterpart; for scenes of
    heavily overlapping triangles, see box_overlap): every splittable triangle
    whose box is longer than `extent` (default: the median triangle's box
    extent times 2^(-depth/2)) is cut, at most `depth` cuts deep, by
    lr_bvh_split_triangles; the tree is built over the pieces' boxes with the
    reference's binned SAH (build_tree, bvh.lucb:85-133) on the host's cores
    or on `device`'s GPU (each piece as a degenerate triangle spanning its
    box: the same boxes, so the same tree); each leaf slot then holds its
    piece's triangle, so the per-triangle arrays grow by the duplicates and
    `order` maps slots to the original triangles. Hits are the same as over
    whole triangles."""
    n = scene.triangle_count()
    if n == 0 or depth <= 0:
        return build_bvh(scene, threads=threads, device=device)
    pos = np.ascontiguousarray(scene.positions, np.float32)
    idx = np.ascontiguousarray(scene.indices, np.uint32)
    if extent is None:
        tri = pos[idx]
        extent = float(np.median((tri.max(1) - tri.min(1)).max(1))) * 2.0 ** (-0.5 * depth)
    mask = _splittable(scene)
    count = _lib.ctypes.c_size_t(0)
    _lib.check(_lib.lib().lr_bvh_split_triangles(_lib.ptr(pos), len(pos), _lib.ptr(idx), n, _lib.ptr(mask),
                                                 float(extent), int(depth), int(threads), None, None, 0,
                                                 _lib.ctypes.byref(count)))
    m = count.value
    if m >= 2 ** 31:
        raise LuceError(LIMIT, "the pieces of a split build number at most 2^31")
    boxes = np.zeros((m, 6), np.float32)
    tris = np.zeros(m, np.uint32)
    _lib.check(_lib.lib().lr_bvh_split_triangles(_lib.ptr(pos), len(pos), _lib.ptr(idx), n, _lib.ptr(mask),
                                                 float(extent), int(depth), int(threads), _lib.ptr(boxes),
                                                 _lib.ptr(tris), m, _lib.ctypes.byref(count)))
    nodes = np.zeros(max(1, 2 * m), NODE_DTYPE)
    order = np.zeros(m, np.uint32)
    got = _lib.ctypes.c_size_t(0)
    if device is not None:
        corners = np.ascontiguousarray(boxes.reshape(m * 2, 3))
        spans = np.ascontiguousarray(np.stack([np.arange(m) * 2, np.arange(m) * 2 + 1, np.arange(m) * 2], 1)
                                     .astype(np.uint32))
        _lib.check(_lib.lib().lr_bvh_build_device(device.handle, _lib.ptr(corners), len(corners), _lib.ptr(spans), m,
                                                  _lib.ptr(nodes), _lib.ctypes.byref(got), _lib.ptr(order)))
    else:
        _lib.check(_lib.lib().lr_bvh_build_boxes(_lib.ptr(boxes), m, int(threads), _lib.ptr(nodes),
                                                 _lib.ctypes.byref(got), _lib.ptr(order)))
    scene.nodes = nodes[:got.value].copy()
    slots = tris[order]
    scene.order = slots
    scene.indices = np.ascontiguousarray(scene.indices[slots])
    scene.faces = np.ascontiguousarray(scene.faces[slots])
    scene.shading = np.ascontiguousarray(scene.shading[slots])
    scene.uvs = np.ascontiguousarray(scene.uvs[slots])
    scene.split_pieces = m
    return scene


def build_bvh(scene, threads=0, device=None):
    """build_bvh (bvh.lucb:66-81) through the native builder (on the host's
    cores, or on `device`'s GPU when one is given: the same tree), then
    reorder (bvh.lucb:198-217)."""
    n = scene.triangle_count()
    nodes = np.zeros(max(1, 2 * n), NODE_DTYPE)
    order = np.zeros(max(1, n), np.uint32)
    count = _lib.ctypes.c_size_t(0)
    if device is not None:
        _lib.check(_lib.lib().lr_bvh_build_device(device.handle, _lib.ptr(scene.positions), scene.point_count(),
                                                  _lib.ptr(scene.indices), n, _lib.ptr(nodes),
                                                  _lib.ctypes.byref(count), _lib.ptr(order)))
    else:
        _lib.check(_lib.lib().lr_bvh_build(_lib.ptr(scene.positions), scene.point_count(),
                                           _lib.ptr(scene.indices), n, int(threads), _lib.ptr(nodes),
                                           _lib.ctypes.byref(count), _lib.ptr(order)))
    scene.nodes = nodes[:count.value].copy()
    scene.order = order[:n].copy()
    if n:
        scene.indices = np.ascontiguousarray(scene.indices[scene.order])
        scene.faces = np.ascontiguousarray(scene.faces[scene.order])
        scene.shading = np.ascontiguousarray(scene.shading[scene.order])
        scene.uvs = np.ascontiguousarray(scene.uvs[scene.order])
    return scene
