"""What a renderer loads beside the geometry (src/render/loading.lucb:51-284):
light records (lights.lucb:29-131), material records (materials.lucb:17-119),
the light tree (light_tree.lucb:91-244) and the constants rows, packed in the
layouts the kernels read. f64 on the host, stored as f32."""
import math
from dataclasses import dataclass, field, replace

import numpy as np

from . import shader_graph as SG
from . import spectrum as S
from ._lib import INVALID, LuceError

LIGHT_STRIDE = 7
MATERIAL_STRIDE = 16
EMITTER_TRIANGLE = 0x80000000
LEAF_NODE = 0x80000000
NO_EMITTER = 0xFFFFFFFF


@dataclass
class OpenPbr:
    """OpenPBR Surface parameters with the specification's defaults
    (materials.lucb:17-55, OpenPbr.standard)."""
    base_weight: float = 1.0
    base_color: tuple = (0.8, 0.8, 0.8)
    base_metalness: float = 0.0
    base_diffuse_roughness: float = 0.0
    specular_weight: float = 1.0
    specular_color: tuple = (1.0, 1.0, 1.0)
    specular_roughness: float = 0.3
    specular_roughness_anisotropy: float = 0.0
    specular_ior: float = 1.5
    transmission_weight: float = 0.0
    transmission_color: tuple = (1.0, 1.0, 1.0)
    coat_weight: float = 0.0
    coat_color: tuple = (1.0, 1.0, 1.0)
    coat_roughness: float = 0.0
    coat_roughness_anisotropy: float = 0.0
    coat_ior: float = 1.6
    coat_darkening: float = 1.0
    emission_luminance: float = 0.0
    emission_color: tuple = (1.0, 1.0, 1.0)
    geometry_opacity: float = 1.0
    geometry_thin_walled: bool = False
    transmission_dispersion_scale: float = 0.0
    transmission_dispersion_abbe_number: float = 20.0
    # A tabulated reflectance for the base colour (64 bins over 380..720 nm):
    # fitted to a sigmoid spectrum through its ACEScg colour.
    base_spectrum: object = None
    # A node graph driving some of the parameters (shader_graph.lucb), and the
    # images its image nodes name: {name: pixels, height x width x 3 or 4,
    # rows from the top; colour images sRGB-encoded in 0..1, data as stored}.
    shader_graph: str = ""
    images: object = None

    @staticmethod
    def of(record):
        """A material record (a dict of OpenPBR parameter names) over the defaults."""
        if isinstance(record, OpenPbr):
            return record
        m = OpenPbr()
        for k, v in (record or {}).items():
            if hasattr(m, k):
                setattr(m, k, tuple(v) if isinstance(v, (list, np.ndarray)) and k != "base_spectrum" else v)
        return m


def cauchy_b(m):
    """materials.lucb:95-103."""
    if m.transmission_dispersion_scale <= 0.0 or m.transmission_dispersion_abbe_number <= 0.0 or m.transmission_weight <= 0.0:
        return 0.0
    f, c = 486.1, 656.3
    return (m.specular_ior - 1.0) * m.transmission_dispersion_scale / (
        m.transmission_dispersion_abbe_number * (1.0 / (f * f) - 1.0 / (c * c)))


def texture_texels(pixels, color):
    """An image node's texels (textures.lucb): RGBA f32, rows from the top;
    a colour image's sRGB-encoded values decoded to linear light (the
    reference's 8-bit sRGB textures decode before filtering), data as stored,
    alpha as stored (1 when absent)."""
    px = np.asarray(pixels, np.float64)
    if px.ndim == 2:
        px = np.repeat(px[:, :, None], 3, axis=2)
    if px.ndim != 3 or px.shape[2] < 3 or px.shape[0] < 1 or px.shape[1] < 1:
        raise LuceError(INVALID, "an image needs height x width x 3 (or 4) pixels")
    rgb = px[:, :, :3]
    if color:
        rgb = np.where(rgb <= 0.04045, rgb / 12.92, ((np.maximum(rgb, 0.0) + 0.055) / 1.055) ** 2.4)
    alpha = px[:, :, 3] if px.shape[2] > 3 else np.ones(px.shape[:2])
    return np.concatenate([rgb, alpha[:, :, None]], axis=2).astype(np.float32)


def pack_material(m, program=-1.0):
    """pack_material (materials.lucb:105-119): 16 float4."""
    out = []
    colors = [m.base_color, m.specular_color, m.coat_color,
              tuple(c * m.emission_luminance for c in m.emission_color), m.transmission_color]
    weights = [m.base_weight, m.specular_weight, m.coat_weight, m.emission_luminance, m.transmission_weight]
    for color, w in zip(colors, weights):
        fit = S.fit_spectrum([max(c, 0.0) for c in color])
        out += fit.row() + list(color) + [w]
    out += [m.base_metalness, m.base_diffuse_roughness, m.specular_roughness, m.specular_roughness_anisotropy,
            m.specular_ior, m.coat_roughness, m.coat_ior, m.coat_darkening,
            m.geometry_opacity, 1.0 if m.geometry_thin_walled else 0.0, m.coat_roughness_anisotropy, cauchy_b(m),
            float(program), 0.0, 0.0, 0.0] + [0.0] * 8
    return out


# Linear Rec.709 (sRGB primaries, D65) to ACEScg, Bradford-adapted
# (environment.lucb:21-25).
REC709_TO_ACESCG = np.array([[0.613097, 0.339523, 0.047379],
                             [0.070194, 0.916354, 0.013452],
                             [0.020616, 0.109570, 0.869815]])


@dataclass
class Environment:
    """A dome's image ready for importance sampling (environment.lucb:53-79):
    texels (ACEScg radiance and weight = luminance x sin theta, four floats a
    texel, rows from the top), each row's cumulative weights (ending at 1)
    then the rows', the weights' sum and the mean luminance over the sphere."""
    width: int
    height: int
    texels: np.ndarray
    cdf: np.ndarray
    total: float
    average: float


def environment_of(pixels):
    """environment_of (environment.lucb:111-156) from linear Rec.709 pixels
    (height x width x 3, or height x width for grey)."""
    px = np.asarray(pixels, np.float64)
    if px.ndim == 2:
        px = np.repeat(px[:, :, None], 3, axis=2)
    if px.ndim != 3 or px.shape[2] < 3 or px.shape[0] < 1 or px.shape[1] < 1:
        raise LuceError(INVALID, "an environment image needs R, G and B channels, or Y")
    height, width = px.shape[0], px.shape[1]
    rgb = np.maximum(0.0, px[:, :, :3] @ REC709_TO_ACESCG.T)
    lum = S.luminance(np.moveaxis(rgb, 2, 0))
    sin_theta = np.sin(np.pi * (np.arange(height) + 0.5) / height)
    weight = lum * sin_theta[:, None]
    texels = np.concatenate([rgb, weight[:, :, None]], axis=2).astype(np.float32).reshape(-1)
    rows = np.zeros(height)
    cdf = np.zeros(width * height + height, np.float32)
    for y in range(height):
        running = np.cumsum(weight[y])  # the reference adds in f64, then stores f32
        row = running[-1]
        rows[y] = row
        part = (running.astype(np.float32).astype(np.float64) / row if row > 0.0
                else (np.arange(width) + 1.0) / width)
        cdf[y * width:(y + 1) * width] = part
    total = float(rows.sum())
    running = np.cumsum(rows)
    cdf[width * height:] = running / total if total > 0.0 else (np.arange(height) + 1.0) / height
    solid_angle = float(sin_theta.sum()) * width
    average = float(weight.sum()) / solid_angle if solid_angle > 0.0 else 0.0
    return Environment(width, height, texels, cdf, total, average)


def pack_lights(lights, radius, environment=None):
    """pack_lights (lights.lucb:29-131): records, the distant lights' total
    power; and the index of the dome that `environment` belongs to (the first
    dome with a texture, -1 if none)."""
    out = []
    powers = []
    textured = -1
    to_acescg = S.film_matrix()
    for light in lights:
        m = np.asarray(light.matrix, np.float64)
        position = m[3, :3]
        emits = -m[2, :3] / np.linalg.norm(m[2, :3])
        u = v = np.zeros(3)
        size = area = 0.0
        cos_half = 1.0
        kind = 0.0
        if light.kind == "rect":
            u = np.array([0.5 * light.width, 0.0, 0.0]) @ m[:3, :3]
            v = np.array([0.0, 0.5 * light.height, 0.0]) @ m[:3, :3]
            area = 4.0 * np.linalg.norm(np.cross(u, v))
        elif light.kind == "disk":
            kind = 1.0
            u = np.array([light.radius, 0.0, 0.0]) @ m[:3, :3]
            v = np.array([0.0, light.radius, 0.0]) @ m[:3, :3]
            area = math.pi * np.linalg.norm(np.cross(u, v))
        elif light.kind == "sphere":
            kind = 2.0
            size = light.radius * np.cbrt(abs(np.linalg.det(m[:3, :3])))
            area = 4.0 * math.pi * size * size
        elif light.kind == "dome":
            kind = 4.0
            u = m[0, :3] / np.linalg.norm(m[0, :3])
            v = m[1, :3] / np.linalg.norm(m[1, :3])
            emits = m[2, :3] / np.linalg.norm(m[2, :3])
        else:
            kind = 3.0
            cos_half = math.cos(0.5 * light.angle * math.pi / 180.0)
        scale = light.intensity * 2.0 ** light.exposure
        if kind < 3.0 and light.normalize:
            if area <= 0.0:
                continue
            scale /= area
        if kind == 3.0 and cos_half < 1.0 - 1e-7:
            scale /= 2.0 * math.pi * (1.0 - cos_half)
        color = tuple(light.color)
        rgb = np.array(color) * scale
        fit = S.fit_spectrum(color)
        fit = S.SpectrumFit(fit.c0, fit.c1, fit.c2, fit.scale * scale)
        flags = 1.0 if light.visible else 0.0
        temperature = getattr(light, "temperature", 0.0) or 0.0
        if temperature > 0.0:
            tint = S.blackbody_xyz(temperature)
            fit = S.SpectrumFit(fit.c0, fit.c1, fit.c2, fit.scale / tint[1])
            tinted = to_acescg @ (tint / tint[1])
            rgb = rgb * tinted
            flags += 2.0
        power = -1.0
        if kind == 3.0:
            power = max(0.0, S.luminance(rgb)) * (2.0 * math.pi * (1.0 - cos_half) if cos_half < 1.0 - 1e-7 else 1.0) \
                * math.pi * radius * radius
        if kind == 4.0:
            mean = 1.0
            if environment is not None and textured < 0 and getattr(light, "texture", None) is not None:
                textured = len(powers)
                mean = environment.average
            power = max(0.0, S.luminance(rgb)) * mean * math.pi * math.pi * radius * radius
        powers.append(power)
        out += [*position, kind, *u, size, *v, cos_half, *emits, area, *fit.row(), *rgb, flags, temperature, 0.0, 0.0, 0.0]
    total = sum(p for p in powers if p >= 0.0)
    distant = sum(1 for p in powers if p >= 0.0)
    running = 0.0
    emitter = 0
    for index, power in enumerate(powers):
        row = (index * LIGHT_STRIDE + 6) * 4
        if power < 0.0:
            out[row + 3] = float(emitter)
            emitter += 1
            continue
        share = power / total if total > 0.0 else 1.0 / distant
        running += share
        out[row + 1] = share
        out[row + 2] = running
    return out, len(powers), total, textured


def _cone_union(a_axis, a_theta, b_axis, b_theta):
    """pbrt-v4's DirectionCone union (light_tree.lucb:41-56)."""
    theta_d = math.acos(max(-1.0, min(1.0, float(a_axis @ b_axis))))
    if min(theta_d + b_theta, math.pi) <= a_theta:
        return a_axis, a_theta
    if min(theta_d + a_theta, math.pi) <= b_theta:
        return b_axis, b_theta
    theta_o = 0.5 * (a_theta + theta_d + b_theta)
    if theta_o >= math.pi:
        return a_axis, math.pi
    rotate = theta_o - a_theta
    w = np.cross(a_axis, b_axis)
    wl = np.linalg.norm(w)
    if wl < 1e-12:
        return a_axis, math.pi
    k = w / wl
    turned = a_axis * math.cos(rotate) + np.cross(k, a_axis) * math.sin(rotate) + k * (k @ a_axis) * (1.0 - math.cos(rotate))
    n = np.linalg.norm(turned)
    return (turned / n if n > 0 else a_axis), theta_o


def _select(keys, order, k):
    """Quickselect with a three-way partition (light_tree.lucb:217-239)."""
    low, high = 0, len(order)
    while high - low > 1:
        pivot = keys[order[(low + high) // 2]]
        less, at, more = low, low, high
        while at < more:
            key = keys[order[at]]
            if key < pivot:
                order[at], order[less] = order[less], order[at]
                less += 1
                at += 1
            elif key > pivot:
                more -= 1
                order[at], order[more] = order[more], order[at]
            else:
                at += 1
        if k < less:
            high = less
        elif k >= more:
            low = more
        else:
            return


def build_light_tree(scene, lights, light_count, materials):
    """build_light_tree (light_tree.lucb:91-199): nodes (12 f32 each), links
    (2 u32 per node), records (4 u32 per emitter, then each triangle's
    emitter), emitter count, root energy."""
    emitters = []  # (lo, hi, axis, theta_o, theta_e, energy, code)
    for index in range(light_count):
        row = lights[index * LIGHT_STRIDE * 4:(index + 1) * LIGHT_STRIDE * 4]
        kind = row[3]
        if kind >= 3.0:
            continue
        position = np.array(row[0:3])
        u, v = np.array(row[4:7]), np.array(row[8:11])
        normal = np.array(row[12:15])
        area = row[15]
        rgb = row[20:23]
        extent = np.abs(u) + np.abs(v)
        theta_o = 0.0
        if kind == 2.0:
            extent = np.full(3, row[7])
            theta_o = math.pi
        energy = max(0.0, S.luminance(rgb)) * area * math.pi
        emitters.append((position - extent, position + extent, normal, theta_o, 0.5 * math.pi, energy, index))
    emissive = any(m.emission_luminance > 0.0 for m in materials)
    n_tri = scene.triangle_count()
    if emissive and n_tri:
        p = scene.positions.astype(np.float64)
        tri = p[scene.indices]
        mat = scene.shading[:, 3]
        for t in range(n_tri):
            mi = int(mat[t])
            if mi >= len(materials) or materials[mi].emission_luminance <= 0.0:
                continue
            a, b, c = tri[t]
            cr = np.cross(b - a, c - a)
            area = 0.5 * np.linalg.norm(cr)
            if area <= 0.0:
                continue
            m = materials[mi]
            energy = max(0.0, S.luminance([x * m.emission_luminance for x in m.emission_color])) * area * 2.0 * math.pi
            emitters.append((tri[t].min(0), tri[t].max(0), cr * (0.5 / area), math.pi, 0.5 * math.pi, energy,
                             EMITTER_TRIANGLE | t))
    count = len(emitters)
    records = np.zeros(count * 4 + n_tri, np.uint32)
    records[count * 4:] = NO_EMITTER
    for index, e in enumerate(emitters):
        records[index * 4] = e[6]
        if e[6] & EMITTER_TRIANGLE:
            records[count * 4 + (e[6] & ~EMITTER_TRIANGLE)] = index
    nodes = []
    links = []
    energy_root = 0.0
    if count:
        order = list(range(count))
        centers = [0.5 * (e[0] + e[1]) for e in emitters]
        nodes.append([0.0] * 12)
        links.append([0, 0])
        stack = [(0, 0, count, 0, 0)]
        while stack:
            node, first, n, trail, depth = stack.pop()
            rng = order[first:first + n]
            f0 = emitters[rng[0]]
            lo, hi, axis, theta_o, theta_e = f0[0].copy(), f0[1].copy(), f0[2], f0[3], f0[4]
            energy = 0.0
            for index in rng:
                e = emitters[index]
                lo = np.minimum(lo, e[0])
                hi = np.maximum(hi, e[1])
                axis, theta_o = _cone_union(axis, theta_o, e[2], e[3])
                theta_e = max(theta_e, e[4])
                energy += e[5]
            nodes[node] = [*lo, energy, *hi, math.cos(theta_o), *axis, math.cos(theta_e)]
            if node == 0:
                energy_root = energy
            if n == 1:
                links[node] = [rng[0], LEAF_NODE]
                records[rng[0] * 4 + 1] = trail
                records[rng[0] * 4 + 2] = depth
                continue
            size = hi - lo
            axis_index = 0
            if size[1] > size[0] and size[1] >= size[2]:
                axis_index = 1
            elif size[2] > size[0] and size[2] > size[1]:
                axis_index = 2
            half = n // 2
            keys = [c[axis_index] for c in centers]
            sub = order[first:first + n]
            _select(keys, sub, half)
            order[first:first + n] = sub
            child = len(links)
            nodes += [[0.0] * 12, [0.0] * 12]
            links += [[0, 0], [0, 0]]
            links[node] = [child, 0]
            stack.append((child + 1, first + half, n - half, trail | (1 << depth), depth + 1))
            stack.append((child, first, half, trail, depth + 1))
    return (np.array(nodes, np.float64).reshape(-1), np.array(links, np.uint32).reshape(-1), records, count,
            energy_root)


@dataclass
class RenderSettings:
    """RenderSettings.of (integrator.lucb:42-85) with the reference's
    defaults: 12 bounces, spectral, seed 1, light tracing 0.5."""
    width: int
    height: int
    samples: int
    max_bounces: int = 12
    spectral: bool = True
    camera: int = -1
    background: tuple = (0.0, 0.0, 0.0)
    material: OpenPbr = field(default_factory=OpenPbr)
    pool: int = 0
    seed: int = 1
    diffuse_bounces: int = 1024
    glossy_bounces: int = 1024
    transmission_bounces: int = 1024
    clamp_indirect: float = 0.0
    noise_threshold: float = 0.0
    adaptive_minimum: int = 16
    profile: bool = False
    guiding: int = 0  # 0 off, 1 the hash grid of vMF mixtures, 2 the learned guide
    train_samples: int = 0  # learned guide: samples that train it (0: 8)
    train_steps: int = 0  # learned guide: Adam steps for each of them, after their pass (0: 8)
    # (the training samples go in passes of a quarter of them)
    # The guide's share of the one-sample mixture (loading.lucb:281: 0.5): of
    # all BSDF samples for the grid, of the diffuse lobe's for the learned guide.
    guide_share: float = 0.5
    light_tracing: float = 0.5
    use_graph: bool = True
    trace: tuple = (0, 0, 0)  # traversal width, group, stack (0: the scene's)
    # Samples one pass may carry when a sample is one band (0: as many as fit
    # in about 2^26 path slots, at most 16; 1: the reference's one a pass).
    pass_samples: int = 0

    @staticmethod
    def of(width, height, samples):
        return RenderSettings(width, height, samples)


def scene_radius(scene):
    """loading.lucb:244-249."""
    if scene.point_count() == 0 or len(scene.nodes) == 0:
        return 1.0
    root = scene.nodes[0]
    d = (root["hi"].astype(np.float64) - root["lo"].astype(np.float64))
    return max(1.0, 0.5 * float(np.linalg.norm(d)))


def _casts_caustics(m):
    if m.transmission_weight > 0.0:
        return True
    if m.base_metalness > 0.0 and m.specular_roughness < 0.3:
        return True
    if m.coat_weight > 0.0 and m.coat_roughness < 0.3:
        return True
    return m.specular_weight > 0.0 and m.specular_roughness < 0.1


def _caster_bounds(scene, library):
    mats = scene.shading[:, 3] if scene.triangle_count() else np.zeros(0, np.uint32)
    casters = np.array([_casts_caustics(m) for m in library], bool)
    sel = np.zeros(len(mats), bool)
    ok = mats < len(library)
    sel[ok] = casters[mats[ok]]
    if not sel.any():
        return np.zeros(3), 1.0, False
    pts = scene.positions.astype(np.float64)[scene.indices[sel].reshape(-1)]
    lo, hi = pts.min(0), pts.max(0)
    return 0.5 * (lo + hi), max(1e-3, 0.5 * float(np.linalg.norm(hi - lo))), True


class Loaded:
    """The packed arrays a renderer needs (Loaded.create, loading.lucb:51-161,
    and view_rows :279-284)."""

    def __init__(self, scene, settings, frame):
        radius = scene_radius(scene)
        # The first dome with an image lights by it (loading.lucb:55-61).
        environment = None
        for light in scene.lights:
            if light.kind == "dome" and getattr(light, "texture", None) is not None:
                environment = environment_of(light.texture)
                break
        lights, light_count, distant_power, dome = pack_lights(scene.lights, radius, environment)
        library = [settings.material] + [OpenPbr.of(r) for r in scene.materials]
        nodes, links, records, emitter_count, tree_energy = build_light_tree(scene, lights, light_count, library)
        node_offset = len(lights) // 4
        packed = list(lights) + list(nodes)
        # The dome's texels (a float4 each), then its cumulative weights
        # (loading.lucb:74-84).
        dome_row = [-1.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0]
        if dome >= 0:
            while len(packed) % 4:
                packed.append(0.0)
            texel_base = len(packed) // 4
            packed += environment.texels.tolist()
            cdf_base = len(packed)
            packed += environment.cdf.tolist()
            while len(packed) % 4:
                packed.append(0.0)
            dome_row = [float(dome), float(texel_base), float(environment.width), float(environment.height),
                        float(cdf_base), environment.total, 0.0, 0.0]
        emitters = np.concatenate([links, records]).astype(np.uint32)
        total = distant_power + tree_energy
        if emitter_count > 0:
            p_distant = distant_power / total if total > 0.0 else 0.0
        else:
            p_distant = 1.0 if distant_power > 0.0 else 0.0
        rows = np.zeros(128, np.float64)
        fm = S.film_matrix()
        integrals = S.cmf_integrals()
        bg = S.fit_spectrum(settings.background)
        rows[0:32] = [light_count, settings.max_bounces, settings.clamp_indirect, 0.0,
                      *bg.row(), *settings.background, 0.0,
                      *fm[0], 0.0, *fm[1], 0.0, *fm[2], 0.0,
                      *integrals, 0.0,
                      settings.diffuse_bounces, settings.glossy_bounces, settings.transmission_bounces, 0.0]
        rows[32:36] = [len(links) // 2, node_offset, p_distant, emitter_count]
        rows[40:44] = [len(links), len(links) + emitter_count * 4, 0.0, 0.0]
        rows[44] = settings.noise_threshold
        rows[45] = settings.adaptive_minimum
        basis = S.basis_weights()
        for r in range(3):
            rows[72 + r * 4:72 + r * 4 + 3] = basis[r]
        rows[64:72] = dome_row
        # The splats (loading.lucb:112-118): their count and alpha cull
        # (K_SPLAT_TREE z, w); the device finds their arrays in the scene.
        splats = getattr(scene, "splats", None)
        rows[106] = float(len(splats.records)) if splats is not None and splats.nodes is not None else 0.0
        rows[107] = 0.00392156862745098
        used = set(int(m) for m in np.unique(scene.shading[:, 3])) if scene.triangle_count() else set()
        # Node graphs: every material's program in one word array, each at
        # its entry (-1: none); their images in one texel array with a table
        # (first texel, width, height) per texture slot (textures.lucb).
        programs, texels, table, slots = [], [], [], {}

        def program_of(m):
            if not m.shader_graph:
                return -1.0

            def resolve(name, color):
                images = m.images or {}
                if name not in images:
                    raise LuceError(INVALID, f"a node graph reads image {name!r} its material does not carry")
                key = (id(images[name]), bool(color))
                if key not in slots:
                    t = texture_texels(images[name], color)
                    slots[key] = len(table)
                    table.append([sum(len(x) for x in texels) // 4, t.shape[1], t.shape[0], 0])
                    texels.append(t.reshape(-1))
                return slots[key]

            words = SG.compile_graph(m.shader_graph, resolve)
            if words is None:
                return -1.0
            entry = len(programs)
            programs.extend(words)
            return float(entry)

        materials = []
        cutouts = False
        for index, m in enumerate(library):
            materials += pack_material(m, program_of(m))
            cutouts = cutouts or (index in used and m.geometry_opacity < 1.0)
        rows[3] = 1.0 if cutouts else 0.0
        distant = any(lights[i * LIGHT_STRIDE * 4 + 3] == 4.0 or
                      (lights[i * LIGHT_STRIDE * 4 + 3] == 3.0 and lights[i * LIGHT_STRIDE * 4 + 11] < 1.0 - 1e-7)
                      for i in range(light_count))
        focus, focus_radius, afar = np.zeros(3), 1.0, False
        if distant and settings.light_tracing > 0.0 and not frame.orthographic:
            focus, focus_radius, afar = _caster_bounds(scene, library)
        afar_power = distant_power * (focus_radius / radius) ** 2 if afar else 0.0
        afar_share = afar_power / (afar_power + tree_energy) if afar_power + tree_energy > 0.0 else 0.0
        rows[84:92] = [afar_share, focus_radius, 0.0, 0.0, *focus, 0.0]
        self.tracing = settings.light_tracing > 0.0 and not frame.orthographic and (emitter_count > 0 or afar)
        pixels = frame.width * frame.height
        pool = settings.pool if settings.pool > 0 else 2097152
        band = min(pool, pixels)
        self.light_paths = 0
        if self.tracing:
            wanted = int(max(1.0, round(settings.light_tracing * pixels)))
            self.light_paths = min(wanted, band)
        # view rows 12..15 (floats 48..63)
        du, dv = np.linalg.norm(frame.du), np.linalg.norm(frame.dv)
        pixel_area = du * dv
        rows[48:64] = [1.0 if settings.guiding else 0.0, 8.0 * du, float(settings.guide_share), 16.0, radius * 1e-4,
                       0.0, 0.0, 0.0,
                       float(self.light_paths), 1.0 / pixel_area if pixel_area > 0.0 else 0.0, frame.lens_radius,
                       frame.focus_distance, *frame.forward, 0.0]
        self.radius = radius
        self.lights = np.array(packed if packed else [0.0] * 4, np.float32)
        self.constants = rows.astype(np.float32)
        self.materials = np.array(materials, np.float32)
        self.programs = np.array(programs if programs else [0], np.uint32)
        self.textures = np.concatenate(texels).astype(np.float32) if texels else np.zeros(4, np.float32)
        self.texture_table = np.array(table if table else [[0, 0, 0, 0]], np.uint32).reshape(-1)
        self.texture_count = len(table)
        self.emitters = emitters if len(emitters) else np.zeros(2, np.uint32)
        self.light_count = light_count
        self.through = 8 if cutouts else 0
        self.emitter_count = emitter_count
        self.library = library
