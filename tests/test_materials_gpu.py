"""The reference's closed-form integrator checks (tests/gpu/integrator.lucb:
113-328) on the GPU: OpenPBR furnaces (rough metal, coat, glass, dispersive
glass), cut-out opacity, emission, assigned materials, rect-light and
emissive-quad irradiance, and cut-out shadows."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def look_at(eye, target, up=(0.0, 1.0, 0.0)):
    """A camera placement at `eye` whose -Z looks at `target`."""
    eye, target, up = (np.asarray(v, np.float64) for v in (eye, target, up))
    back = eye - target
    back /= np.linalg.norm(back)
    right = np.cross(up, back)
    right /= np.linalg.norm(right)
    upv = np.cross(back, right)
    m = np.eye(4)
    m[0, :3], m[1, :3], m[2, :3], m[3, :3] = right, upv, back, eye
    return m


def sphere_set(frequency=24, centre=(0.0, 0.0, 0.0), material=None):
    from luce_render import render, scenes
    pts, tris = scenes.icosphere(frequency)
    mats = None if material is None else [material] * len(tris)
    mesh = render.Mesh(pts + np.asarray(centre), tris, face_materials=mats)
    return render.GeometrySet(meshes=[mesh])


def grid_mesh(size, n, y=0.0, material=None):
    from luce_render import render
    xs = np.linspace(-size / 2, size / 2, n + 1)
    px, pz = np.meshgrid(xs, xs, indexing="ij")
    pts = np.stack([px.ravel(), np.full(px.size, y), pz.ravel()], 1)
    tris, faces = [], []
    for i in range(n):
        for j in range(n):
            a, b, c, d = i * (n + 1) + j, (i + 1) * (n + 1) + j, (i + 1) * (n + 1) + j + 1, i * (n + 1) + j + 1
            tris += [[a, c, b], [a, d, c]]  # counter-clockwise seen from +y
            faces += [i * n + j] * 2
    mats = None if material is None else [material] * (n * n)
    return render.Mesh(pts, np.array(tris), np.array(faces), n * n, face_materials=mats)


def run(device, gset, width, height, samples, **kw):
    from luce_render import render
    from luce_render.loading import OpenPbr, RenderSettings
    sc = render.compile_scene(gset)
    st = RenderSettings.of(width, height, samples)
    material = kw.pop("material", None)
    if material:
        st.material = OpenPbr(**material)
    for k, v in kw.items():
        setattr(st, k, v)
    say(f"render {width}x{height} spp {samples} spectral {st.spectral} lt {st.light_tracing}")
    return render.render(device, sc, st)


def window(img, x0, y0, size):
    return img[y0:y0 + size, x0:x0 + size, :3].astype(np.float64).mean(axis=(0, 1))


def furnace_camera(gset):
    from luce_render import render
    gset.cameras.append(render.Camera(look_at((0, 0, 4), (0, 0, 0)), "perspective", 50.0, 36.0))
    return gset


@pytest.mark.parametrize("spectral,roughness", [(True, 0.5), (False, 1.0)])
def test_metal_furnace(device, spectral, roughness):
    """integrator.lucb:113-133: white metal (F0 = 1) of any roughness reflects
    everything once Kulla–Conty compensation restores multiple scattering."""
    img = run(device, furnace_camera(sphere_set()), 64, 48, 64, spectral=spectral, max_bounces=64,
              background=(1.0, 1.0, 1.0), material=dict(base_color=(1.0, 1.0, 1.0), base_metalness=1.0,
                                                        specular_roughness=roughness))
    c = window(img, 24, 16, 16)
    say("metal furnace", spectral, roughness, c)
    assert abs(c[1] - 1.0) <= 0.02


@pytest.mark.parametrize("spectral,kind", [(True, "coat"), (False, "glass"), (True, "glass"), (True, "dispersive")])
def test_layer_furnace(device, spectral, kind):
    """integrator.lucb:139-166: a white base under a clear coat, and closed
    spheres of smooth glass (and dispersive glass) stay white under a white sky."""
    material = dict(base_color=(1.0, 1.0, 1.0))
    samples = 64
    if kind == "coat":
        material.update(coat_weight=1.0, coat_roughness=0.3)
    else:
        material.update(transmission_weight=1.0, specular_roughness=0.0)
        if kind == "dispersive":
            material.update(transmission_dispersion_scale=1.0, transmission_dispersion_abbe_number=10.0)
            samples = 1024
    img = run(device, furnace_camera(sphere_set(32)), 64, 48, samples, spectral=spectral, max_bounces=64,
              background=(1.0, 1.0, 1.0), material=material)
    c = window(img, 24, 16, 16)
    say(kind, "furnace", spectral, c)
    assert abs(c[1] - 1.0) <= 0.02


def test_cutout_emitter(device):
    """integrator.lucb:170-190: a 3-nit sphere at opacity one half shows 2.25."""
    from luce_render import render
    g = sphere_set()
    g.cameras.append(render.Camera(look_at((0, 0, 4), (0, 0, 0)), "perspective", 100.0, 36.0))
    img = run(device, g, 32, 32, 256, spectral=False, material=dict(emission_luminance=3.0, geometry_opacity=0.5,
                                                                     base_color=(0.0, 0.0, 0.0), specular_weight=0.0))
    c = window(img, 12, 12, 8)
    say("cut-out", c)
    assert abs(c[1] - 2.25) <= 0.05


@pytest.mark.parametrize("spectral", [False, True])
def test_emission(device, spectral):
    """integrator.lucb:194-212: emission luminance 3 of colour (1, .5, .25)."""
    from luce_render import render
    g = sphere_set()
    g.cameras.append(render.Camera(look_at((0, 0, 4), (0, 0, 0)), "perspective", 100.0, 36.0))
    img = run(device, g, 32, 32, 16, spectral=spectral,
              material=dict(emission_luminance=3.0, emission_color=(1.0, 0.5, 0.25)))
    c = window(img, 12, 12, 8)
    say("emission", spectral, c)
    assert abs(c[0] - 3.0) <= 0.03 and abs(c[1] - 1.5) <= 0.02 and abs(c[2] - 0.75) <= 0.02


def test_assigned_materials(device):
    """integrator.lucb:216-245: per-face materials by name; an unknown name
    falls back to the default material."""
    from luce_render import render, scenes
    pts, tris = scenes.icosphere(24)
    left = render.Mesh(pts + [-1.2, 0, 0], tris, face_materials=["materials/red"] * len(tris))
    right = render.Mesh(pts + [1.2, 0, 0], tris, face_materials=["materials/missing"] * len(tris))
    g = render.GeometrySet(meshes=[left, right],
                           materials={"materials/red": {"emission_luminance": 2.0, "emission_color": (1.0, 0.0, 0.0)}})
    g.cameras.append(render.Camera(look_at((0, 0, 6), (0, 0, 0)), "perspective", 35.0, 36.0))
    img = run(device, g, 48, 32, 4, spectral=False)
    glowing, dark = window(img, 12, 14, 4), window(img, 32, 14, 4)
    say("assigned", glowing, dark)
    assert abs(glowing[0] - 2.0) <= 0.03 and glowing[1] <= 0.01 and dark[0] <= 0.05


def corner_irradiance(x, y, h):
    """Irradiance under the corner of an x by y rectangle of radiance 1 at height h."""
    a, b = math.sqrt(x * x + h * h), math.sqrt(y * y + h * h)
    return 0.5 * (x / a * math.atan(y / a) + y / b * math.atan(x / b))


def _floor_scene(with_rect=False, emissive_quad=False, veil_opacity=None):
    from luce_render import render, scenes
    g = render.GeometrySet(meshes=[grid_mesh(20.0, 8)])
    if emissive_quad:
        g.meshes.append(grid_mesh(1.0, 1, y=1.0, material="materials/glow"))
        g.materials["materials/glow"] = {"emission_luminance": 4.0, "base_weight": 0.0, "specular_weight": 0.0}
    if veil_opacity is not None:
        g.meshes.append(grid_mesh(3.0, 1, y=0.5, material="materials/veil"))
        g.materials["materials/veil"] = {"geometry_opacity": veil_opacity, "base_weight": 0.0, "specular_weight": 0.0}
    if with_rect:
        m = np.eye(4)
        m[1, :3], m[2, :3], m[3, :3] = (0, 0, 1), (0, 1, 0), (0, 1, 0)  # local -Z points down
        g.lights.append(render.Light("rect", m, (1.0, 1.0, 1.0), 4.0, width=1.0, height=1.0))
    return g


@pytest.mark.parametrize("spectral,lt", [(True, 0.0), (False, 0.0), (True, 1.0), (False, 0.5)])
def test_rect_light_irradiance(device, spectral, lt):
    """integrator.lucb:252-289: the floor below a 1x1 rect light of radiance 4
    at height 1 shows albedo/pi x 4 x 4 corner_irradiance(.5, .5, 1) within 2%."""
    from luce_render import render
    g = _floor_scene(with_rect=True)
    eye_y = 0.5
    g.cameras.append(render.Camera(look_at((0, eye_y, 3), (0, 0, 0)), "perspective", 200.0, 36.0))
    samples = 256 if lt == 0.0 else 4096
    img = run(device, g, 32, 32, samples, spectral=spectral, max_bounces=1, light_tracing=lt,
              material=dict(base_color=(0.5, 0.5, 0.5), specular_weight=0.0))
    measured = window(img, 14, 14, 4)[1]
    expected = 0.5 / math.pi * 4.0 * 4.0 * corner_irradiance(0.5, 0.5, 1.0)
    say("rect light", spectral, lt, measured, expected)
    assert abs(measured / expected - 1.0) <= 0.02


@pytest.mark.parametrize("spectral,lt", [(True, 0.0), (False, 0.0), (True, 1.0)])
def test_emissive_quad_irradiance(device, spectral, lt):
    """integrator.lucb:251-289: the same square as an emissive mesh (light
    tree, two-sided) lights the floor as the rect light does."""
    from luce_render import render
    g = _floor_scene(emissive_quad=True)
    g.cameras.append(render.Camera(look_at((0, 0.5, 3), (0, 0, 0)), "perspective", 200.0, 36.0))
    samples = 256 if lt == 0.0 else 4096
    img = run(device, g, 32, 32, samples, spectral=spectral, max_bounces=1, light_tracing=lt,
              material=dict(base_color=(0.5, 0.5, 0.5), specular_weight=0.0))
    measured = window(img, 14, 14, 4)[1]
    expected = 0.5 / math.pi * 4.0 * 4.0 * corner_irradiance(0.5, 0.5, 1.0)
    say("emissive quad", spectral, lt, measured, expected)
    assert abs(measured / expected - 1.0) <= 0.02


def test_cutout_shadow(device):
    """integrator.lucb:295-328: a veil of opacity one half passes half of what
    a fully transparent one passes (0.5 +- 0.015)."""
    from luce_render import render
    means = []
    for opacity in (0.5, 0.0):
        g = _floor_scene(with_rect=True, veil_opacity=opacity)
        g.cameras.append(render.Camera(look_at((0, 0.3, 3), (0, 0, 0)), "perspective", 200.0, 36.0))
        img = run(device, g, 32, 32, 256, spectral=False, max_bounces=1,
                  material=dict(base_color=(0.5, 0.5, 0.5), specular_weight=0.0))
        means.append(window(img, 14, 14, 4)[1])
    say("cut-out shadow", means)
    assert abs(means[0] / means[1] - 0.5) <= 0.015
