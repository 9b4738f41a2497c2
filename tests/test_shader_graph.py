"""Node-graph materials on the CPU (src/render/shader_graph.lucb,
shaders/shader_program.glsl): the compiler's programs against the word format
worked out by hand, its errors, and the oracle's interpreter against closed
forms: the reference's own test graphs (tests/gpu/integrator.lucb:594-730)
that must come out a 0.5 grey with a flat normal, bumps of a height rising 0.5
a unit that must tilt the normal to (-0.5, 1, 0), Fresnel at normal incidence,
Perlin noise's zeros at lattice points, bilinear texture lookups at texel
centres and between them."""
import math

import numpy as np
import pytest

from luce_render import loading, scenes
from luce_render import shader_graph as SG
from luce_render._lib import LuceError
from oracle import oracle as O

TAG = SG.SLOT_TAG


def f32(x):
    return int(np.array(x, np.float32).view(np.uint32))


def test_checker_graph_words():
    """A checker driving base_color: the implicit uv (slots 0-2), the checker
    (its vector the uv's slots, then color1, color2, scale as constants;
    outputs color 3-5 and fac 6), the output of register 0, width 3, bit 0."""
    words = SG.compile_graph(scenes.GRAPH_FLOOR)
    want = [SG.OP_UV, 0, 0, 0, 1, 2,
            SG.OP_CHECKER, TAG | 0, TAG | 1, TAG | 2, f32(0.8), f32(0.7), f32(0.6), f32(0.1), f32(0.1), f32(0.12), f32(4.0),
            3, 4, 5, 6,
            SG.OP_OUTPUT, 0, 3, 0, TAG | 3, TAG | 4, TAG | 5,
            SG.OP_END]
    assert words == want


def test_value_spreads_to_a_color_and_bump_compiles_three_heights():
    """A value linked to a color input is spread by a color op; a bump's
    height compiles at the point and at two offsets (uv ops with axis 1 and
    2, their footprint the bump's filter width operand)."""
    words = SG.compile_graph("node v value; input v value 0.25; node s openpbr_surface; link s base_color v value")
    assert words == [SG.OP_VALUE, f32(0.25), 0, SG.OP_COLOR, TAG | 0, TAG | 0, TAG | 0, 1, 2, 3,
                     SG.OP_OUTPUT, 0, 3, 0, TAG | 1, TAG | 2, TAG | 3, SG.OP_END]
    words = SG.compile_graph("node u uv; node x separate; link x color u uv; node b bump; link b height x r; "
                             "node s openpbr_surface; link s geometry_normal b normal")
    w = f32(0.1)  # the filter width's default: the offsets' footprint
    one = f32(1.0)
    assert words == [SG.OP_UV, 0, 0, 0, 1, 2, SG.OP_SEPARATE, TAG | 0, TAG | 1, TAG | 2, 3, 4, 5,
                     SG.OP_UV, 1, w, 6, 7, 8, SG.OP_SEPARATE, TAG | 6, TAG | 7, TAG | 8, 9, 10, 11,
                     SG.OP_UV, 2, w, 12, 13, 14, SG.OP_SEPARATE, TAG | 12, TAG | 13, TAG | 14, 15, 16, 17,
                     SG.OP_BUMP, one, one, w, TAG | 3, TAG | 9, TAG | 15, 0, 0, 0, 18, 19, 20,
                     SG.OP_OUTPUT, 11, 3, 9, TAG | 18, TAG | 19, TAG | 20, SG.OP_END]


@pytest.mark.parametrize("text, message", [
    ("node a warp; node s openpbr_surface; link s base_color a color", "unknown kind"),
    ("node a math; node b math; link a a b value; link b a a value; node s openpbr_surface; "
     "link s base_weight a value", "loops"),
    ("node b bump; node c bump; link b height c normal; node s openpbr_surface; link s geometry_normal b normal",
     "another Bump"),
    ("node i image x.png; node s openpbr_surface; link s base_color i color", "image"),
    ("node s openpbr_surface; link s base_color nowhere color", "lacks"),
    ("node c color; node s openpbr_surface; link s base_color c shade", "does not have"),
    ("node m math; input m a x; node s openpbr_surface; link s base_weight m value", "parse"),
])
def test_graph_errors(text, message):
    with pytest.raises(LuceError) as e:
        SG.compile_graph(text)
    assert message in str(e.value)


def test_graphs_that_drive_nothing_compile_to_none():
    assert SG.compile_graph("") is None
    assert SG.compile_graph("node c checker") is None
    assert SG.compile_graph("node c checker; node s openpbr_surface; input s base_color 1 0 0") is None


def test_more_than_64_values_is_an_error():
    chain = ["node c0 color"] + [f"node c{i} color; link c{i} color c{i - 1} color" for i in range(1, 30)]
    text = "; ".join(chain) + "; node s openpbr_surface; link s base_color c29 color"
    with pytest.raises(LuceError) as e:
        SG.compile_graph(text)
    assert "64" in str(e.value)


# ------------------------------------------------------------ the oracle's VM

FRAME = np.zeros(16, np.float32)
FRAME[:3] = (0.0, 0.5, 3.0)
FRAME[8:11] = (0.001, 0.0, 0.0)  # a pixel 1e-3 wide at unit distance


def _floor_triangle():
    """A floor triangle (normal +y) with uvs 0.05 (x, z), as grid_mesh maps."""
    corners = np.array([[-2.0, 0.0, -2.0], [-2.0, 0.0, 2.0], [2.0, 0.0, 2.0]])
    return corners, (corners[:, [0, 2]] * 0.05).reshape(-1)


def _run(text, x=(0.3, 0.0, -0.4), n=(0.0, 1.0, 0.0), view=(0.0, 0.6, 0.8), images=None, uv=None):
    material = loading.OpenPbr.of({"shader_graph": text, "images": images})
    texels, table, slots = [], [], {}

    def resolve(name, color):
        if (name, color) not in slots:
            t = loading.texture_texels(images[name], color)
            slots[(name, color)] = len(table)
            table.append([sum(len(a) for a in texels) // 4, t.shape[1], t.shape[0], 0])
            texels.append(t.reshape(-1))
        return slots[(name, color)]

    words = SG.compile_graph(text, resolve)
    materials = np.array(loading.pack_material(material, 0.0), np.float32)
    corners, uvs = _floor_triangle()
    if uv is None:
        uv = np.array(x)[[0, 2]] * 0.05
    return O.run_program(materials, words, 0, x, n, view, uv, corners=corners, uvs=uvs, frame=FRAME,
                         textures=np.concatenate(texels) if texels else None,
                         texture_table=np.array(table, np.uint32).reshape(-1) if table else None,
                         texture_count=len(table))


GRAPH_NODES = ("node n noise; input n scale 3; node r ramp; link r fac n fac; input r color1 0.5; input r color2 0.5; "
               "node c checker; input c color1 0.5; input c color2 0.5; node a math multiply; input a a 0.25; "
               "input a b 2; node b map_range; input b value 0.25; input b from_max 0.5; node d clamp; "
               "input d value 0.7; input d max 0.5; node k combine; link k r a value; link k g b value; "
               "link k b d value; node m mix; link m fac c fac; link m a r color; link m b k color; node s separate; "
               "link s color m color; node x combine; link x r s r; link x g s g; link x b s b; node nm normal_map; "
               "node surface openpbr_surface; link surface base_color x color; link surface geometry_normal nm normal")


@pytest.mark.parametrize("seed", range(6))
def test_every_node_kind_comes_out_grey(seed):
    """tests/gpu/integrator.lucb:639-660: noise into a ramp of one grey, a
    checker choosing between that and (multiply, map range, clamp) combined,
    split and combined again, a normal map of the flat colour: base colour
    0.5 grey and the unchanged normal at any point."""
    rng = np.random.default_rng(seed)
    x = (rng.uniform(-2, 2), 0.0, rng.uniform(-2, 2))
    mask, viewed, values = _run(GRAPH_NODES, x=x)
    assert mask == (1 << 0) | (1 << 9) and not viewed
    assert np.allclose(values[0:3], 0.5, atol=1e-12)
    assert np.allclose(values[11:14], (0.0, 1.0, 0.0), atol=1e-12)


TILTED = np.array([-0.5, 1.0, 0.0]) / math.sqrt(1.25)


@pytest.mark.parametrize("graph", [
    "node p position; node s separate; link s color p position; node h math multiply; link h a s r; input h b 0.5; "
    "node b bump; link b height h value; node surface openpbr_surface; link surface geometry_normal b normal",
    "node u uv; node s separate; link s color u uv; node h math multiply; link h a s r; input h b 10; node b bump; "
    "link b height h value; node surface openpbr_surface; link surface geometry_normal b normal"])
def test_bump_of_a_slope_tilts_the_normal(graph):
    """tests/gpu/integrator.lucb:662-699: a height rising 0.5 a unit along x,
    by position and by uv (u = 0.05 x), tilts the normal to (-0.5, 1, 0)."""
    for x in [(0.3, 0.0, -0.4), (-1.5, 0.0, 1.2)]:
        mask, _, values = _run(graph, x=x)
        assert mask == 1 << 9
        assert np.allclose(values[11:14], TILTED, atol=1e-5), values[11:14]


def test_bump_of_a_constant_height_leaves_the_normal():
    _, _, values = _run("node b bump; input b height 0.3; node surface openpbr_surface; "
                        "link surface geometry_normal b normal")
    assert np.allclose(values[11:14], (0.0, 1.0, 0.0), atol=1e-12)


def test_fresnel_and_layer_weight():
    """Fresnel at normal incidence is ((ior - 1) / (ior + 1))^2; Layer
    Weight's facing at blend 0.5 is 1 - |cos|; both mark the program viewed."""
    text = ("node f fresnel; input f ior 1.5; node w layer_weight; input w blend 0.5; node c combine; "
            "link c r f fac; link c g w facing; link c b w fresnel; node s openpbr_surface; link s base_color c color")
    mask, viewed, values = _run(text, view=(0.0, 1.0, 0.0))
    assert viewed and mask == 1
    assert abs(values[0] - 0.04) < 1e-6
    assert abs(values[1] - 0.0) < 1e-12
    c = 0.6
    _, _, values = _run(text, view=(0.0, c, 0.8))
    assert abs(values[1] - (1.0 - c)) < 1e-6


def test_checker_cells():
    """The checker's cell parity at scale 4 on uv = 0.05 (x, z)."""
    for x in [(0.3, 0.0, -0.4), (5.1, 0.0, 2.2), (-7.3, 0.0, 9.9)]:
        _, _, values = _run(scenes.GRAPH_FLOOR, x=x)
        u, v = np.float32(x[0] * 0.05), np.float32(x[2] * 0.05)
        odd = (math.floor(u * 4 + 1e-5) + math.floor(v * 4 + 1e-5)) & 1
        want = (0.8, 0.7, 0.6) if odd else (0.1, 0.1, 0.12)
        assert np.allclose(values[0:3], want, atol=1e-7)


def test_math_operations():
    table = {"add": lambda a, b: a + b, "subtract": lambda a, b: a - b, "multiply": lambda a, b: a * b,
             "divide": lambda a, b: a / b if b else 0.0, "minimum": min, "maximum": max,
             "absolute": lambda a, b: abs(a), "sine": lambda a, b: math.sin(a), "cosine": lambda a, b: math.cos(a),
             "floor": lambda a, b: math.floor(a), "fraction": lambda a, b: a - math.floor(a)}
    pairs = [(0.75, 0.5), (-1.25, 2.0), (3.0, 0.0), (-2.0, 3.0)]
    for name, fn in table.items():
        for a, b in pairs:
            text = f"node m math {name}; input m a {a}; input m b {b}; node s openpbr_surface; link s base_weight m value"
            _, _, values = _run(text)
            assert abs(values[3] - fn(a, b)) < 1e-12, (name, a, b)
    # power: negative bases with whole exponents keep the sign of odd powers
    for a, b, want in [(2.0, 3.0, 8.0), (-2.0, 3.0, -8.0), (-2.0, 2.0, 4.0), (-2.0, 0.5, 0.0), (0.25, 0.5, 0.5)]:
        text = f"node m math power; input m a {a}; input m b {b}; node s openpbr_surface; link s base_weight m value"
        _, _, values = _run(text)
        assert abs(values[3] - want) < 1e-12, (a, b)


def test_noise_is_zero_on_the_lattice_and_bounded():
    rng = np.random.default_rng(3)
    for c in rng.integers(-50, 50, (20, 3)):
        assert O.gradient_noise(*[float(v) for v in c]) == 0.0
    vals = [O.gradient_noise(*rng.uniform(-20, 20, 3)) for _ in range(2000)]
    assert max(abs(v) for v in vals) <= 1.0 and np.std(vals) > 0.1
    # continuous across a lattice plane
    a, b = O.gradient_noise(2.0 - 1e-9, 0.3, 0.7), O.gradient_noise(2.0 + 1e-9, 0.3, 0.7)
    assert abs(a - b) < 1e-7


def test_image_bilinear_at_centres_between_and_across_the_seam():
    """A 4 x 2 data image: at a texel's centre its value, halfway between two
    its neighbours' mean, and at u = 0 the mean of the last and first columns
    (repeat)."""
    px = np.arange(4 * 2 * 4, dtype=np.float64).reshape(2, 4, 4) / 40.0
    images = {"t": px}
    text = ("node i image t data; node s openpbr_surface; link s base_color i color; link s base_weight i alpha")
    for uv, want in [((0.125, 0.25), px[0, 0]), ((0.375, 0.75), px[1, 1]), ((0.25, 0.25), (px[0, 0] + px[0, 1]) / 2),
                     ((0.0, 0.25), (px[0, 3] + px[0, 0]) / 2), ((1.125, -0.75), px[0, 0])]:
        _, _, values = _run(text, images=images, uv=np.array(uv))
        assert np.allclose(values[0:4], want, atol=1e-6), (uv, values[0:4], want)


def test_colour_images_decode_srgb_then_go_to_acescg():
    """tests/gpu/integrator.lucb:594-637: 188s in an sRGB image are
    ((188/255 + 0.055) / 1.055)^2.4 linear; a grey stays grey in ACEScg (the
    matrix's rows sum to one within 1e-6)."""
    grey = np.full((4, 4, 4), 188.0 / 255.0)
    grey[:, :, 3] = 1.0
    _, _, values = _run("node g image grey; node s openpbr_surface; link s base_color g color", images={"grey": grey})
    albedo = ((188.0 / 255.0 + 0.055) / 1.055) ** 2.4
    assert np.allclose(values[0:3], albedo, rtol=3e-6)


def test_scene_programs_pack():
    """graph_set: each material's entry (row 13 x) starts its program, which
    ends in OP_END; textures hold 8 x 6 and 4 x 4 texels in order."""
    from luce_render import render
    from luce_render.loading import Loaded, RenderSettings
    sc = render.compile_scene(scenes.graph_set(frequency=4))
    st = RenderSettings.of(32, 24, 1)
    ld = Loaded(sc, st, render.ray_frame(sc.cameras[0], 32, 24))
    entries = ld.materials.reshape(-1, 16, 4)[:, 13, 0]
    assert entries[0] == -1.0 and sorted(entries[1:]) == sorted(set(entries[1:]))
    assert ld.programs[-1] == SG.OP_END
    assert list(ld.texture_table) == [0, 8, 6, 0, 48, 4, 4, 0] and ld.texture_count == 2
    assert len(ld.textures) == (48 + 16) * 4
