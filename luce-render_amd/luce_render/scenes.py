"""Seeded synthetic scenes of the measured workloads.

  cornell_probe_set      Cornell box (5 unit walls, 2 rotated boxes; 34
                         triangles, 17 polygons) plus 10,000 random triangles
                         (centres U[-0.4,0.4]^3, vertices N(0, 0.03) per axis,
                         seed 1); pinhole camera down -Z; probed at 256x256.
  cornell_render_set     the same box with Lambert walls (0.73 white, 0.65
                         red/green bands) and a 0.25x0.25 rect light at
                         y = 0.999, radiance 17; rendered at 512x512.
  furnace_set            an albedo-0.8 sphere under a constant environment.
  instanced_set          64 displaced icospheres (frequency 39: 30,420
                         triangles each, ~1.95M unique) placed as 1,000 rigid
                         instances in a 20^3 volume (30.4M triangles
                         realized) with 256 rect/disk/sphere lights of
                         log-uniform intensity in [1, 1000]; 1920x1080.

Random numbers: SplitMix64-seeded xoshiro256** (Blackman and Vigna), f64
uniforms from the top 53 bits, normals by Box-Muller (cosine branch).
"""
import math

import numpy as np

from .scene import Camera, GeometrySet, Instance, Light, Mesh, SplatCloud

MASK = (1 << 64) - 1


class Xoshiro256:
    """xoshiro256** seeded by SplitMix64."""

    def __init__(self, seed):
        s = seed & MASK
        self.s = []
        for _ in range(4):
            s = (s + 0x9E3779B97F4A7C15) & MASK
            z = s
            z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
            z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
            self.s.append(z ^ (z >> 31))

    def next(self):
        s = self.s
        result = (((s[1] * 5) & MASK) << 7 | ((s[1] * 5) & MASK) >> 57) & MASK
        result = (result * 9) & MASK
        t = (s[1] << 17) & MASK
        s[2] ^= s[0]
        s[3] ^= s[1]
        s[1] ^= s[2]
        s[0] ^= s[3]
        s[2] ^= t
        s[3] = ((s[3] << 45) | (s[3] >> 19)) & MASK
        return result

    def uniform(self):
        return (self.next() >> 11) * (1.0 / 9007199254740992.0)

    def normal(self):
        u1 = 1.0 - self.uniform()
        u2 = self.uniform()
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)


def placement(rotation=None, translation=(0.0, 0.0, 0.0), scale=1.0):
    """A 4x4 placement, rows: x, y, z axes then translation."""
    m = np.eye(4)
    r = np.eye(3) if rotation is None else np.asarray(rotation, np.float64)
    m[:3, :3] = r * scale
    m[3, :3] = translation
    return m


def rotation_y(degrees):
    a = math.radians(degrees)
    c, s = math.cos(a), math.sin(a)
    # rows are the images of the x, y, z axes
    return np.array([[c, 0.0, -s], [0.0, 1.0, 0.0], [s, 0.0, c]])


def quad_mesh(corners, normal, material=None):
    """One polygon of two triangles over four corners (counter-clockwise seen
    from `normal`)."""
    pts = np.asarray(corners, np.float64)
    tris = np.array([[0, 1, 2], [0, 2, 3]])
    n = np.tile(np.asarray(normal, np.float64), (2, 3, 1))
    return Mesh(pts, tris, np.array([0, 0]), 1, normals=n,
                face_materials=None if material is None else [material])


def box_mesh(size, material=None):
    """An axis-aligned box centred at the origin: 6 polygons, 12 triangles,
    flat outward normals."""
    hx, hy, hz = (0.5 * s for s in size)
    faces = [
        ([[hx, -hy, -hz], [hx, hy, -hz], [hx, hy, hz], [hx, -hy, hz]], [1, 0, 0]),
        ([[-hx, -hy, hz], [-hx, hy, hz], [-hx, hy, -hz], [-hx, -hy, -hz]], [-1, 0, 0]),
        ([[-hx, hy, -hz], [-hx, hy, hz], [hx, hy, hz], [hx, hy, -hz]], [0, 1, 0]),
        ([[-hx, -hy, hz], [-hx, -hy, -hz], [hx, -hy, -hz], [hx, -hy, hz]], [0, -1, 0]),
        ([[-hx, -hy, hz], [hx, -hy, hz], [hx, hy, hz], [-hx, hy, hz]], [0, 0, 1]),
        ([[hx, -hy, -hz], [-hx, -hy, -hz], [-hx, hy, -hz], [hx, hy, -hz]], [0, 0, -1]),
    ]
    pts, tris, tf, normals = [], [], [], []
    for f, (c, n) in enumerate(faces):
        b = len(pts)
        pts.extend(c)
        tris.extend([[b, b + 1, b + 2], [b, b + 2, b + 3]])
        tf.extend([f, f])
        normals.extend([[n, n, n], [n, n, n]])
    mats = None if material is None else [material] * 6
    return Mesh(np.array(pts), np.array(tris), np.array(tf), 6, normals=np.array(normals, np.float64),
                face_materials=mats)


def _cornell_geometry(white="white", red="red", green="green"):
    """The box [-1,1]^3 open toward +z: floor, ceiling, back, left (red),
    right (green), with two rotated boxes on the floor."""
    walls = GeometrySet()
    walls.meshes.append(quad_mesh([[-1, -1, 1], [1, -1, 1], [1, -1, -1], [-1, -1, -1]], [0, 1, 0], white))
    walls.meshes.append(quad_mesh([[-1, 1, -1], [1, 1, -1], [1, 1, 1], [-1, 1, 1]], [0, -1, 0], white))
    walls.meshes.append(quad_mesh([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1]], [0, 0, 1], white))
    walls.meshes.append(quad_mesh([[-1, -1, 1], [-1, -1, -1], [-1, 1, -1], [-1, 1, 1]], [1, 0, 0], red))
    walls.meshes.append(quad_mesh([[1, -1, -1], [1, -1, 1], [1, 1, 1], [1, 1, -1]], [-1, 0, 0], green))
    short = GeometrySet(meshes=[box_mesh((0.6, 0.6, 0.6), white)])
    tall = GeometrySet(meshes=[box_mesh((0.6, 1.2, 0.6), white)])
    walls.instances.append(Instance(short, placement(rotation_y(-17.0), (0.33, -0.7, 0.25))))
    walls.instances.append(Instance(tall, placement(rotation_y(20.0), (-0.33, -0.4, -0.3))))
    return walls


def cornell_camera(focal_length=30.0):
    """A pinhole near (0, 0, 3) looking down -Z, 36 mm aperture. It sits a
    little off the box's axis so that no pixel-centre ray runs exactly along
    the edge where two walls meet."""
    return Camera(placement(translation=(0.0123, 0.0217, 3.0)), "perspective", focal_length, 36.0)


def random_triangles(count=10000, seed=1, spread=0.4, sigma=0.03):
    """`count` triangles: centre U[-spread, spread]^3, each vertex the centre
    plus N(0, sigma^2) per axis; one polygon each."""
    rng = Xoshiro256(seed)
    pts = np.empty((count * 3, 3))
    for t in range(count):
        c = [(2.0 * rng.uniform() - 1.0) * spread for _ in range(3)]
        for k in range(3):
            pts[t * 3 + k] = [c[a] + sigma * rng.normal() for a in range(3)]
    tris = np.arange(count * 3).reshape(count, 3)
    return Mesh(pts, tris, np.arange(count), count)


def random_triangles_fast(count, seed=1, spread=0.4, sigma=0.03):
    """Large random-triangle clouds (hundreds of thousands to millions):
    the same distribution as random_triangles, drawn with numpy's PCG64."""
    rng = np.random.default_rng(seed)
    c = rng.uniform(-spread, spread, (count, 1, 3))
    pts = (c + sigma * rng.standard_normal((count, 3, 3))).reshape(-1, 3)
    tris = np.arange(count * 3).reshape(count, 3)
    return Mesh(pts, tris, np.arange(count), count)


def cornell_probe_set(random_count=10000, seed=1):
    gset = _cornell_geometry()
    if random_count:
        gset.meshes.append(random_triangles(random_count, seed))
    gset.cameras.append(cornell_camera())
    return gset


# ------------------------------------------------------------ instanced scene

def icosphere(frequency):
    """A geodesic sphere: each icosahedron face split into frequency^2
    triangles on a barycentric grid, points projected to the unit sphere and
    shared along edges (20 f^2 triangles, 10 f^2 + 2 points)."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], np.float64)
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
             (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    f = frequency
    keys = {}
    points = []

    def point(a, b, c, i, j):
        # grid point i along a->b, j along a->c; edge points keyed by their
        # integer position on the edge so both faces share them
        k = f - i - j
        weights = {a: k, b: i, c: j}
        key = tuple(sorted((idx, w) for idx, w in weights.items() if w))
        if key not in keys:
            p = (v[a] * k + v[b] * i + v[c] * j) / f
            keys[key] = len(points)
            points.append(p / np.linalg.norm(p))
        return keys[key]

    tris = []
    for a, b, c in faces:
        for i in range(f):
            for j in range(f - i):
                p0 = point(a, b, c, i, j)
                p1 = point(a, b, c, i + 1, j)
                p2 = point(a, b, c, i, j + 1)
                tris.append((p0, p1, p2))
                if i + j < f - 1:
                    p3 = point(a, b, c, i + 1, j + 1)
                    tris.append((p1, p3, p2))
    return np.array(points), np.array(tris, np.int64)


def _hash3(ix, iy, iz, seed):
    h = (ix.astype(np.int64) * 73856093) ^ (iy.astype(np.int64) * 19349663) ^ (iz.astype(np.int64) * 83492791) ^ seed
    h = h & 0xFFFFFFFF
    h = (h ^ (h >> 16)) * 0x45D9F3B & 0xFFFFFFFF
    h = (h ^ (h >> 16)) * 0x45D9F3B & 0xFFFFFFFF
    h = h ^ (h >> 16)
    return (h & 0xFFFFFF) / float(0xFFFFFF) * 2.0 - 1.0


def value_noise(p, seed, octaves=3, frequency=2.0):
    """Seeded 3-D value noise (smoothstep trilinear lattice values in [-1, 1]),
    `octaves` octaves of doubling frequency and halving weight, normalized to
    [-1, 1]."""
    total = np.zeros(len(p))
    norm = 0.0
    amp = 1.0
    for o in range(octaves):
        q = p * frequency * (2 ** o)
        i = np.floor(q)
        fr = q - i
        s = fr * fr * (3.0 - 2.0 * fr)
        acc = np.zeros(len(p))
        for dx in (0, 1):
            for dy in (0, 1):
                for dz in (0, 1):
                    w = (s[:, 0] if dx else 1 - s[:, 0]) * (s[:, 1] if dy else 1 - s[:, 1]) * \
                        (s[:, 2] if dz else 1 - s[:, 2])
                    acc += w * _hash3(i[:, 0] + dx, i[:, 1] + dy, i[:, 2] + dz, seed + 7919 * o)
        total += amp * acc
        norm += amp
        amp *= 0.5
    return total / norm


def random_rotation(rng):
    """A uniform random rotation (Shoemake's quaternion)."""
    u1, u2, u3 = rng.uniform(), rng.uniform(), rng.uniform()
    q = (math.sqrt(1 - u1) * math.sin(2 * math.pi * u2), math.sqrt(1 - u1) * math.cos(2 * math.pi * u2),
         math.sqrt(u1) * math.sin(2 * math.pi * u3), math.sqrt(u1) * math.cos(2 * math.pi * u3))
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y + z * w), 2 * (x * z - y * w)],
                     [2 * (x * y - z * w), 1 - 2 * (x * x + z * z), 2 * (y * z + x * w)],
                     [2 * (x * z + y * w), 2 * (y * z - x * w), 1 - 2 * (x * x + y * y)]])


def instanced_set(meshes=64, instances=1000, frequency=39, radius=0.6, extent=20.0, lights=256, seed=3,
                  amplitude=0.05):
    """The large instanced scene: `meshes` displaced icospheres placed as
    `instances` rigid instances in an `extent`^3 volume, `lights` lights."""
    rng = Xoshiro256(seed)
    base_points, base_tris = icosphere(frequency)
    protos = []
    for m in range(meshes):
        d = value_noise(base_points * 1.0, seed * 1000 + m, 3, 2.0)
        pts = base_points * (radius * (1.0 + amplitude * d))[:, None]
        protos.append(GeometrySet(meshes=[Mesh(pts, base_tris, np.arange(len(base_tris)), len(base_tris),
                                               face_materials=["surface"] * len(base_tris))],
                                  materials={}))
    gset = GeometrySet(materials={"surface": {"base_color": (0.7, 0.7, 0.7)}})
    half = extent / 2.0
    for i in range(instances):
        rot = random_rotation(rng)
        t = [(2.0 * rng.uniform() - 1.0) * half for _ in range(3)]
        gset.instances.append(Instance(protos[i % meshes], placement(rot, t)))
    kinds = ["rect", "disk", "sphere"]
    for i in range(lights):
        kind = kinds[i % 3]
        t = [(2.0 * rng.uniform() - 1.0) * half for _ in range(3)]
        intensity = math.exp(rng.uniform() * math.log(1000.0))
        gset.lights.append(Light(kind, placement(random_rotation(rng), t), (1.0, 1.0, 1.0), intensity,
                                 width=0.3, height=0.3, radius=0.15))
    # Camera outside the volume looking down -Z at its centre.
    gset.cameras.append(Camera(placement(translation=(0.0, 0.0, half + 14.0)), "perspective", 30.0, 36.0))
    return gset


# ------------------------------------------------------------- render scenes

BAND_BINS = 64
BAND_LAMBDA = np.linspace(380.0, 720.0, BAND_BINS)


def band_spectrum(low, high, value=0.65):
    """A 64-bin 380..720 nm reflectance: `value` inside [low, high] nm, 0 outside."""
    return np.where((BAND_LAMBDA >= low) & (BAND_LAMBDA <= high), value, 0.0)


def spectrum_to_acescg(reflectance):
    """A tabulated reflectance (64 bins over 380..720 nm, linear between bins,
    0 outside) as the ACEScg colour of its XYZ under E, through the same
    colour-matching fit and film matrix the renderer uses. The spectral build
    renders its sigmoid fit (fit_spectrum), the RGB build this colour."""
    from . import spectrum as S
    lam = np.arange(S.LAMBDA_MIN, S.LAMBDA_MAX + 0.5, 1.0)
    r = np.interp(lam, BAND_LAMBDA, reflectance, left=0.0, right=0.0)
    xyz = (S.cmf(lam) * r[:, None]).sum(axis=0) / S.cmf_integrals()
    return tuple(float(c) for c in S.film_matrix() @ xyz)


def lambert(color):
    """A Lambertian OpenPBR material record: no specular, roughness 0."""
    return {"base_color": tuple(color), "specular_weight": 0.0, "base_diffuse_roughness": 0.0}


def cornell_render_set(random_count=0, light_radiance=17.0):
    """The Cornell render workload: the probe's box with Lambert walls (0.73
    white; 0.65 red band 600..720 nm on the left, 0.65 green band 500..570 nm
    on the right) and a 0.25 x 0.25 rect light at y = 0.999 facing down,
    radiance 17 (flat spectrum)."""
    gset = _cornell_geometry()
    gset.materials = {
        "white": lambert((0.73, 0.73, 0.73)),
        "red": lambert(spectrum_to_acescg(band_spectrum(600.0, 720.0))),
        "green": lambert(spectrum_to_acescg(band_spectrum(500.0, 570.0))),
    }
    if random_count:
        gset.meshes.append(random_triangles(random_count, 1) if random_count <= 20000
                           else random_triangles_fast(random_count, 1))
        gset.materials["random"] = lambert((0.5, 0.5, 0.5))
        gset.meshes[-1].face_materials = ["random"] * random_count
    # The light's local -Z is its emission direction: rows x, y, z with z = +y
    # world so that it shines down.
    m = np.eye(4)
    m[0, :3] = (1.0, 0.0, 0.0)
    m[1, :3] = (0.0, 0.0, 1.0)
    m[2, :3] = (0.0, 1.0, 0.0)
    m[3, :3] = (0.0, 0.999, 0.0)
    gset.lights.append(Light("rect", m, (1.0, 1.0, 1.0), light_radiance, width=0.25, height=0.25))
    gset.cameras.append(cornell_camera())
    return gset


def furnace_set(albedo=0.8, frequency=16, roughness=0.0):
    """A white furnace: an albedo-`albedo` diffuse sphere (radius 1 at the
    origin; EON roughness `roughness`, 0 being Lambert) under a constant
    environment (the render's background), seen from (0, 0, 4)."""
    pts, tris = icosphere(frequency)
    gset = GeometrySet(meshes=[Mesh(pts, tris, np.arange(len(tris)), len(tris), face_materials=["ball"] * len(tris))])
    ball = lambert((albedo, albedo, albedo))
    ball["base_diffuse_roughness"] = float(roughness)
    gset.materials = {"ball": ball}
    gset.cameras.append(Camera(placement(translation=(0.0, 0.0, 4.0)), "perspective", 50.0, 36.0))
    return gset


def sky_texture(width=64, height=32, seed=11):
    """A small synthetic lat-long environment (linear Rec.709, rows from the
    top): a blue-to-white sky gradient over a dark ground, a warm sun patch and
    seeded cloud noise, so radiance varies strongly across texels."""
    rng = np.random.default_rng(seed)
    v = (np.arange(height) + 0.5) / height
    u = (np.arange(width) + 0.5) / width
    sky = np.where(v[:, None] < 0.5, 1.0, 0.05)
    img = np.zeros((height, width, 3))
    img[:, :, 0] = sky * (0.4 + 0.6 * v[:, None])
    img[:, :, 1] = sky * (0.6 + 0.4 * v[:, None])
    img[:, :, 2] = sky * 1.0
    img *= 0.5 + rng.random((height, width, 1))
    sun_v, sun_u = int(height * 0.25), int(width * 0.6)
    img[sun_v - 1:sun_v + 1, sun_u - 1:sun_u + 1] = (40.0, 30.0, 20.0)
    return img


def dome_set(texture=None, albedo=0.8, frequency=16, intensity=1.0, ground=True, ball=None):
    """A textured dome (environment.lucb) over an albedo-`albedo` diffuse
    sphere (radius 1 at the origin; `ball` an OpenPBR dict to use instead)
    and, with `ground`, a diffuse floor at y = -1; seen from (0, 0, 4) so the
    environment fills the background."""
    pts, tris = icosphere(frequency)
    meshes = [Mesh(pts, tris, np.arange(len(tris)), len(tris), face_materials=["ball"] * len(tris))]
    if ground:
        floor = np.array([[-6.0, -1.0, -6.0], [6.0, -1.0, -6.0], [6.0, -1.0, 6.0], [-6.0, -1.0, 6.0]])
        meshes.append(Mesh(floor, np.array([[0, 2, 1], [0, 3, 2]]), np.array([0, 0]), 1,
                           face_materials=["floor"]))
    gset = GeometrySet(meshes=meshes)
    gset.materials = {"ball": lambert((albedo, albedo, albedo)) if ball is None else ball,
                      "floor": lambert((0.5, 0.5, 0.5))}
    tex = sky_texture() if texture is None else texture
    gset.lights.append(Light("dome", placement(), (1.0, 1.0, 1.0), intensity, texture=tex))
    gset.cameras.append(Camera(placement(translation=(0.0, 0.0, 4.0)), "perspective", 50.0, 36.0))
    return gset


GRAPH_FLOOR = """
node s openpbr_surface
node c checker
input c scale 4
input c color1 0.8 0.7 0.6
input c color2 0.1 0.1 0.12
link s base_color c color
"""

GRAPH_BALL = """
node s openpbr_surface
node n noise
input n scale 3
input n detail 3.5
node r ramp
link r fac n fac
input r color1 0.1 0.3 0.7
input r color2 0.9 0.8 0.2
input r position1 0.3
input r position2 0.7
link s base_color r color
node b bump
input b strength 0.6
input b distance 0.5
link b height n fac
link s geometry_normal b normal
node m map_range
link m value n fac
input m to_min 0.1
input m to_max 0.6
node k clamp
link k value m value
input k min 0.2
input k max 0.5
link s specular_roughness k value
node f fresnel
input f ior 1.45
link s coat_weight f fac
"""

GRAPH_WALL = """
node s openpbr_surface
node i image wall
node t image bumps data
node nm normal_map
link nm color t color
input nm strength 0.8
link s base_color i color
link s geometry_normal nm normal
node w math multiply
link w a i alpha
input w b 0.5
link s base_metalness w value
node x separate
link x color i color
node y combine
link y r x b
link y g x r
link y b x g
node z mix
input z fac 0.3
link z a y color
input z b 1 1 1
link s specular_color z color
node lw layer_weight
input lw blend 0.3
node pw math power
link pw a lw facing
input pw b 2
link s specular_weight pw value
"""


def graph_images(seed=13):
    """The wall's images: an 8 x 6 colour image with alpha (sRGB-encoded,
    rows from the top) and a 4 x 4 tangent-space normal map (data)."""
    rng = np.random.default_rng(seed)
    wall = rng.random((6, 8, 4))
    wall[:, :, 3] = 0.25 + 0.75 * rng.random((6, 8))
    tilt = rng.normal(0.0, 0.35, (4, 4, 2))
    n = np.concatenate([tilt, np.ones((4, 4, 1))], axis=2)
    n /= np.linalg.norm(n, axis=2, keepdims=True)
    return {"wall": wall, "bumps": 0.5 + 0.5 * n}


def graph_set(frequency=24):
    """Materials driven by node graphs (shader_graph.lucb) under the textured
    dome: a checker floor by its uvs; a ball whose noise drives its colour
    (a ramp), bump, roughness (map range and clamp) and, by Fresnel, its coat;
    a back wall whose image and normal-map textures drive colour, normal and
    metalness, with separate/combine/mix specular colour and a layer weight."""
    pts, tris = icosphere(frequency)
    meshes = [Mesh(pts, tris, np.arange(len(tris)), len(tris), face_materials=["ball"] * len(tris))]
    floor = np.array([[-6.0, -1.0, -6.0], [6.0, -1.0, -6.0], [6.0, -1.0, 6.0], [-6.0, -1.0, 6.0]])
    corner_uv = np.array([[0.0, 0.0], [3.0, 0.0], [3.0, 3.0], [0.0, 3.0]])
    ftris = np.array([[0, 2, 1], [0, 3, 2]])
    meshes.append(Mesh(floor, ftris, np.array([0, 0]), 1, face_materials=["floor"], uvs=corner_uv[ftris]))
    wall = np.array([[-3.0, -1.0, -2.5], [3.0, -1.0, -2.5], [3.0, 2.5, -2.5], [-3.0, 2.5, -2.5]])
    wtris = np.array([[0, 1, 2], [0, 2, 3]])
    wall_uv = np.array([[0.0, 1.0], [2.0, 1.0], [2.0, 0.0], [0.0, 0.0]])
    meshes.append(Mesh(wall, wtris, np.array([0, 0]), 1, face_materials=["wall"], uvs=wall_uv[wtris]))
    gset = GeometrySet(meshes=meshes)
    gset.materials = {"ball": {"base_color": (0.5, 0.5, 0.5), "coat_color": (0.9, 0.95, 1.0),
                               "shader_graph": GRAPH_BALL},
                      "floor": {"base_color": (0.5, 0.5, 0.5), "shader_graph": GRAPH_FLOOR},
                      "wall": {"base_color": (0.5, 0.5, 0.5), "specular_roughness": 0.35,
                               "shader_graph": GRAPH_WALL, "images": graph_images()}}
    gset.lights.append(Light("dome", placement(), (1.0, 1.0, 1.0), 1.0, texture=sky_texture()))
    gset.cameras.append(Camera(placement(translation=(0.0, 0.3, 4.5)), "perspective", 40.0, 36.0))
    return gset


def grid_mesh(size=20.0, divisions=8, material=None, uv_scale=0.05):
    """A size x size floor at y = 0 in divisions^2 squares (two triangles
    each, counter-clockwise from above, normals up), corner uvs
    uv_scale * (x, z)."""
    t = np.linspace(-0.5 * size, 0.5 * size, divisions + 1)
    xs, zs = np.meshgrid(t, t, indexing="xy")
    pts = np.stack([xs.reshape(-1), np.zeros(xs.size), zs.reshape(-1)], 1)
    tris = []
    for j in range(divisions):
        for i in range(divisions):
            a = j * (divisions + 1) + i
            b, c, d = a + 1, a + divisions + 2, a + divisions + 1
            tris += [[a, d, c], [a, c, b]]
    tris = np.array(tris)
    n = np.zeros((len(tris), 3, 3))
    n[:, :, 1] = 1.0
    uvs = (pts[:, [0, 2]] * uv_scale)[tris]
    return Mesh(pts, tris, np.arange(len(tris)) // 2, divisions * divisions, normals=n,
                face_materials=None if material is None else [material] * (divisions * divisions), uvs=uvs)


def _rotation_x(radians):
    c, s = math.cos(radians), math.sin(radians)
    return np.array([[1.0, 0.0, 0.0], [0.0, c, s], [0.0, -s, c]])


def graph_floor_set(material):
    """The reference's floor window (tests/gpu/integrator.lucb:754-769): a
    20 x 20 floor (8 x 8 squares, uvs 0.05 (x, z)) of `material` under a 1 x 1
    rect light at y = 1 facing down (radiance 4), seen from (0, 0.5, 3)
    pitched down toward the light's foot through a 200 mm lens."""
    gset = GeometrySet(meshes=[grid_mesh(material="floor")])
    gset.materials = {"floor": material}
    m = placement(rotation=_rotation_x(-0.5 * math.pi), translation=(0.0, 1.0, 0.0))
    gset.lights.append(Light("rect", m, (1.0, 1.0, 1.0), 4.0, width=1.0, height=1.0))
    pitch = -math.atan2(0.5, 3.0)
    gset.cameras.append(Camera(placement(rotation=_rotation_x(pitch), translation=(0.0, 0.5, 3.0)), "perspective",
                               200.0, 36.0))
    return gset


def viewed_graph_set(material, frequency=16):
    """The reference's viewed-graph scene (tests/gpu/integrator.lucb:701-730):
    a radius-0.5 sphere at the origin and a 20 x 20 floor at y = 0, both of
    `material`, under a 1 x 1 rect light at y = 2 facing down (radiance 8),
    seen from (0, 0.6, 3) pitched down 0.1 rad through a 50 mm lens."""
    pts, tris = icosphere(frequency)
    ball = Mesh(np.asarray(pts) * 0.5, tris, np.arange(len(tris)), len(tris), face_materials=["viewed"] * len(tris))
    gset = GeometrySet(meshes=[ball, grid_mesh(material="viewed")])
    gset.materials = {"viewed": material}
    m = placement(rotation=_rotation_x(-0.5 * math.pi), translation=(0.0, 2.0, 0.0))
    gset.lights.append(Light("rect", m, (1.0, 1.0, 1.0), 8.0, width=1.0, height=1.0))
    gset.cameras.append(Camera(placement(rotation=_rotation_x(-0.1), translation=(0.0, 0.6, 3.0)), "perspective", 50.0,
                               36.0))
    return gset


def _about(axis, angle):
    a = np.asarray(axis, np.float64)
    a = a / np.linalg.norm(a)
    return np.array([*(a * math.sin(0.5 * angle)), math.cos(0.5 * angle)])


def splat_description(centers, scales, opacities, colors, orients=None, sh=None):
    """Splats as a dict of arrays (centre, orient xyzw, scale, opacity, linear
    colour, SH items (n, items, 3))."""
    n = len(centers)
    d = {"center": np.asarray(centers, np.float64).reshape(n, 3),
         "orient": np.tile([0.0, 0.0, 0.0, 1.0], (n, 1)) if orients is None else np.asarray(orients, np.float64),
         "scale": np.asarray(scales, np.float64).reshape(n, 3),
         "opacity": np.asarray(opacities, np.float64).reshape(n),
         "color": np.asarray(colors, np.float64).reshape(n, 3)}
    if sh is not None:
        d["sh"] = np.asarray(sh, np.float64).reshape(n, -1, 3)
    return d


def splat_cloud(splats, color_space="linear"):
    return SplatCloud(points=splats["center"], orient=splats["orient"], scale=splats["scale"],
                      opacity=splats["opacity"], color=splats["color"], sh=splats.get("sh"), color_space=color_space)


def splat_primary_splats():
    """The reference's six overlapping, rotated, anisotropic splats with SH
    degree 1 (tests/gpu/splats.lucb:146-159): one hides behind the quad."""
    sh = np.zeros((6, 3, 3))
    sh[0] = np.array([0.1, 0.0, -0.05, 0.0, 0.08, 0.0, -0.1, 0.0, 0.1]).reshape(3, 3)
    return splat_description(
        centers=[(-0.4, 0.1, 0.0), (0.3, -0.1, -0.5), (0.0, 0.0, 0.6), (0.9, 0.6, 0.2), (-0.8, -0.7, -1.0),
                 (0.5, -0.5, -2.6)],
        scales=[(0.5, 0.25, 0.3), (0.35, 0.6, 0.2), (0.2, 0.2, 0.2), (0.3, 0.1, 0.3), (0.4, 0.4, 0.4), (0.5, 0.5, 0.5)],
        opacities=[0.8, 0.7, 0.5, 0.99, 0.35, 0.95],
        colors=[(0.9, 0.2, 0.1), (0.1, 0.8, 0.2), (0.2, 0.3, 1.0), (1.0, 1.0, 0.3), (0.6, 0.6, 0.6), (1.0, 0.0, 1.0)],
        orients=[_about((0.0, 0.0, 1.0), 0.6), _about((0.2822, 0.9406, 0.1881), 0.9), [0.0, 0.0, 0.0, 1.0],
                 _about((1.0, 0.0, 0.0), 1.1), [0.0, 0.0, 0.0, 1.0], [0.0, 0.0, 0.0, 1.0]],
        sh=sh)


SPLAT_QUAD_Z = -2.0
SPLAT_QUAD_HALF = 1.5
SPLAT_QUAD_EMISSION = (0.2, 0.4, 0.9)


def splat_primary_set(splats=None):
    """tests/gpu/splats.lucb:143-214: the six splats in front of an emissive
    quad (z = -2, 3 x 3, emission (0.2, 0.4, 0.9), no reflection) seen from
    (0, 0, 5) through a 50 mm lens."""
    splats = splat_primary_splats() if splats is None else splats
    h, z = SPLAT_QUAD_HALF, SPLAT_QUAD_Z
    quad = quad_mesh([[-h, -h, z], [h, -h, z], [h, h, z], [-h, h, z]], [0, 0, 1], "quad")
    gset = GeometrySet(meshes=[quad], clouds=[splat_cloud(splats)])
    gset.materials = {"quad": {"base_weight": 0.0, "specular_weight": 0.0, "emission_luminance": 1.0,
                               "emission_color": SPLAT_QUAD_EMISSION}}
    gset.cameras.append(Camera(placement(translation=(0.0, 0.0, 5.0)), "perspective", 50.0, 36.0))
    return gset


def splat_shadow_set(light_tracing_view=False):
    """tests/gpu/splats.lucb:226-288: a white floor (8 x 8, albedo 0.8, no
    specular) under the sun straight down (radiance 3, 0.01 degrees) with a
    black splat of opacity 1 and sigma 0.25 one unit above it, seen from
    (0, 2, 4) pitched down 0.6 rad through a 35 mm lens."""
    splats = splat_description(centers=[(0.0, 1.0, 0.0)], scales=[(0.25, 0.25, 0.25)], opacities=[1.0],
                               colors=[(0.0, 0.0, 0.0)])
    gset = GeometrySet(meshes=[grid_mesh(size=8.0, divisions=8, material="floor")], clouds=[splat_cloud(splats)])
    gset.materials = {"floor": {"base_color": (0.8, 0.8, 0.8), "specular_weight": 0.0}}
    gset.lights.append(Light("distant", placement(rotation=_rotation_x(-0.5 * math.pi)), (1.0, 1.0, 1.0), 3.0,
                             angle=0.01))
    gset.cameras.append(Camera(placement(rotation=_rotation_x(-0.6), translation=(0.0, 2.0, 4.0)), "perspective",
                               35.0, 36.0))
    return gset, splats


def splat_ball(count=3000, seed=17):
    """Random splats in a unit ball, many overlapping, a few large: random
    orientations, scales 0.02..0.12 (one in fifty 0.3..0.45), opacity
    0.2..0.9, colours in 0..1 and small degree-1 SH."""
    rng = np.random.default_rng(seed)
    u = rng.normal(size=(count, 3))
    centers = u / np.linalg.norm(u, axis=1, keepdims=True) * rng.random((count, 1)) ** (1.0 / 3.0)
    scales = rng.uniform(0.02, 0.12, (count, 3))
    big = rng.random(count) < 0.02
    scales[big] = rng.uniform(0.3, 0.45, (int(big.sum()), 3))
    q = rng.normal(size=(count, 4))
    return splat_description(centers, scales, rng.uniform(0.2, 0.9, count), rng.random((count, 3)),
                             orients=q / np.linalg.norm(q, axis=1, keepdims=True),
                             sh=rng.normal(0.0, 0.1, (count, 3, 3)))


def splat_ball_set(splats=None):
    """A ball of splats and no triangles, against a (0.1, 0.2, 0.3)
    background, seen from (0, 0, 4) through a 50 mm lens."""
    splats = splat_ball() if splats is None else splats
    gset = GeometrySet(clouds=[splat_cloud(splats)])
    gset.cameras.append(Camera(placement(translation=(0.0, 0.0, 4.0)), "perspective", 50.0, 36.0))
    return gset, splats


def many_lights_set(count=40, seed=5):
    """The Cornell render box lit by `count` small analytic lights instead of
    its rect light: rect, disk and sphere lights in turn (0.12 wide, radius
    0.06) at random places and orientations inside the box, radiance
    log-uniform in [1, 100], all visible to the camera. Exercises the closest
    hit's loop over many lights and its culling."""
    gset = cornell_render_set()
    gset.lights = []
    rng = Xoshiro256(seed)
    kinds = ["rect", "disk", "sphere"]
    for i in range(count):
        t = [(2.0 * rng.uniform() - 1.0) * 0.8 for _ in range(3)]
        intensity = math.exp(rng.uniform() * math.log(100.0))
        gset.lights.append(Light(kinds[i % 3], placement(random_rotation(rng), t), (1.0, 1.0, 1.0), intensity,
                                 width=0.12, height=0.12, radius=0.06))
    return gset


def glossy_set(seed=7, soup=0, light_radiance=30.0):
    """The learned-guiding workload: the Cornell box and its two boxes (the
    plain render's geometry; `soup` random triangles of the probe's kind may
    be added), every surface a random
    OpenPBR material (base colour U[0.2, 0.9]^3, metalness 1 with chance
    0.3, specular roughness U[0.05, 0.3]; the triangles in 16 material
    groups), lit only by 8 sphere lights of radius 0.02 just under the
    ceiling, each hidden from the room below by a 0.16-wide occluder: the
    room is lit by the bright patches the lights throw on the ceiling."""
    rng = Xoshiro256(seed)

    def material():
        return {"base_color": tuple(0.2 + 0.7 * rng.uniform() for _ in range(3)),
                "base_metalness": 1.0 if rng.uniform() < 0.3 else 0.0,
                "specular_roughness": 0.05 + 0.25 * rng.uniform()}

    names = ["floor", "ceiling", "back", "left", "right", "short", "tall"]
    gset = GeometrySet()
    gset.meshes.append(quad_mesh([[-1, -1, 1], [1, -1, 1], [1, -1, -1], [-1, -1, -1]], [0, 1, 0], "floor"))
    gset.meshes.append(quad_mesh([[-1, 1, -1], [1, 1, -1], [1, 1, 1], [-1, 1, 1]], [0, -1, 0], "ceiling"))
    gset.meshes.append(quad_mesh([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1]], [0, 0, 1], "back"))
    gset.meshes.append(quad_mesh([[-1, -1, 1], [-1, -1, -1], [-1, 1, -1], [-1, 1, 1]], [1, 0, 0], "left"))
    gset.meshes.append(quad_mesh([[1, -1, -1], [1, -1, 1], [1, 1, 1], [1, 1, -1]], [-1, 0, 0], "right"))
    short = GeometrySet(meshes=[box_mesh((0.6, 0.6, 0.6), "short")])
    tall = GeometrySet(meshes=[box_mesh((0.6, 1.2, 0.6), "tall")])
    gset.instances.append(Instance(short, placement(rotation_y(-17.0), (0.33, -0.7, 0.25))))
    gset.instances.append(Instance(tall, placement(rotation_y(20.0), (-0.33, -0.4, -0.3))))
    gset.materials = {name: material() for name in names}
    if soup:
        tri = random_triangles(soup, 1) if soup <= 20000 else random_triangles_fast(soup, 1)
        tri.face_materials = [f"soup{t % 16}" for t in range(soup)]
        gset.meshes.append(tri)
        for k in range(16):
            gset.materials[f"soup{k}"] = material()
    # lights below the ceiling, each with a downward-facing occluder under it
    h = 0.08
    for i, (x, z) in enumerate([(x, z) for x in (-0.6, -0.2, 0.2, 0.6) for z in (-0.4, 0.2)]):
        gset.lights.append(Light("sphere", placement(translation=(x, 0.9, z)), (1.0, 1.0, 1.0), light_radiance,
                                 radius=0.02))
        gset.meshes.append(quad_mesh([[x - h, 0.86, z - h], [x + h, 0.86, z - h], [x + h, 0.86, z + h],
                                      [x - h, 0.86, z + h]], [0, -1, 0], "ceiling"))
    gset.cameras.append(cornell_camera())
    return gset


def lamp_room_set():
    """The guiding workload of tests/gpu/integrator.lucb:362-389: a closed
    6 x 3 x 6 room (floor y = 0, ceiling y = 3, walls at x, z = +-3) lit only
    by a 0.2 x 0.2 rect lamp at (0, 0.25, 0) shining up (radiance 200) inside
    a shade of four 0.3-wide, 1-tall panels (y 0.2..1.2), so everything the
    camera sees is lit by the small spot the lamp throws on the ceiling. The
    surfaces carry no material (the render's default one applies). Camera at
    (2.5, 1.5, 2.5) looking toward the room's centre, focal length 24."""
    meshes = []
    h = 3.0
    # floor, ceiling, then the walls; normals face the room
    meshes.append(quad_mesh([[-3, 0, 3], [3, 0, 3], [3, 0, -3], [-3, 0, -3]], [0, 1, 0]))
    meshes.append(quad_mesh([[-3, h, -3], [3, h, -3], [3, h, 3], [-3, h, 3]], [0, -1, 0]))
    meshes.append(quad_mesh([[3, 0, 3], [3, h, 3], [3, h, -3], [3, 0, -3]], [-1, 0, 0]))
    meshes.append(quad_mesh([[-3, 0, -3], [-3, h, -3], [-3, h, 3], [-3, 0, 3]], [1, 0, 0]))
    meshes.append(quad_mesh([[-3, 0, 3], [-3, h, 3], [3, h, 3], [3, 0, 3]], [0, 0, -1]))
    meshes.append(quad_mesh([[3, 0, -3], [3, h, -3], [-3, h, -3], [-3, 0, -3]], [0, 0, 1]))
    lo, hi, w = 0.2, 1.2, 0.15
    meshes.append(quad_mesh([[w, lo, -w], [w, hi, -w], [w, hi, w], [w, lo, w]], [1, 0, 0]))
    meshes.append(quad_mesh([[-w, lo, w], [-w, hi, w], [-w, hi, -w], [-w, lo, -w]], [-1, 0, 0]))
    meshes.append(quad_mesh([[-w, lo, w], [w, lo, w], [w, hi, w], [-w, hi, w]], [0, 0, 1]))
    meshes.append(quad_mesh([[w, lo, -w], [-w, lo, -w], [-w, hi, -w], [w, hi, -w]], [0, 0, -1]))
    gset = GeometrySet(meshes=meshes)
    # The light's local -Z is its emission direction: z = -y world, so it
    # shines up.
    m = np.eye(4)
    m[0, :3] = (1.0, 0.0, 0.0)
    m[1, :3] = (0.0, 0.0, 1.0)
    m[2, :3] = (0.0, -1.0, 0.0)
    m[3, :3] = (0.0, 0.25, 0.0)
    gset.lights.append(Light("rect", m, (1.0, 1.0, 1.0), 200.0, width=0.2, height=0.2))
    pitch, yaw = -0.15, 0.785
    cx, sx, cy, sy = math.cos(pitch), math.sin(pitch), math.cos(yaw), math.sin(yaw)
    rx = np.array([[1.0, 0.0, 0.0], [0.0, cx, sx], [0.0, -sx, cx]])
    ry = np.array([[cy, 0.0, -sy], [0.0, 1.0, 0.0], [sy, 0.0, cy]])
    gset.cameras.append(Camera(placement(rx @ ry, (2.5, 1.5, 2.5)), "perspective", 24.0, 36.0))
    return gset


def serving_sets():
    """The serving workload's four scenes: the Cornell box with 10,000
    random triangles (the probe's geometry), the plain Cornell render, and
    the Cornell render with 500,000 and 5,000,000 random triangles."""
    return [cornell_render_set(random_count=10000), cornell_render_set(),
            cornell_render_set(random_count=500000), cornell_render_set(random_count=5000000)]
