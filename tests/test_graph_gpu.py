"""GPU parity of node-graph materials (shaders/shader_program.glsl, run by the
shading kernels before a surface is read): the f64 oracle's camera paths pixel
by pixel on a scene of checker, noise, bump, ramp, image, normal-map, math,
mix, separate/combine, Fresnel and Layer Weight nodes, and the reference's own
integrator tests of graphs (tests/gpu/integrator.lucb:594-730): a textured
floor lights as its grey, a graph of every node kind as a 0.5 grey, bumps of
a slope as the tilted normal given outright, and light tracing through a
graph that reads the view leaves the image mean unchanged."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _render(device, gset, **kw):
    from luce_render import render
    from luce_render.loading import RenderSettings
    sc = render.compile_scene(gset)
    st = RenderSettings.of(kw.pop("width"), kw.pop("height"), kw.pop("samples"))
    for k, v in kw.items():
        setattr(st, k, v)
    say(f"render {st.width}x{st.height} spp {st.samples} spectral {st.spectral} lt {st.light_tracing}")
    return sc, st, render.render(device, sc, st)


@pytest.mark.parametrize("spectral", [False, True])
def test_graph_scene_matches_f64_oracle(device, spectral):
    """graph_set against the f64 oracle at the same Sobol seeds, camera paths
    only. Tolerances as the Cornell comparison's, with 98% of pixels (not
    99%) within 1e-3: the checker's cell edges, the noise's lattice floors
    and the image texels flip on f32 rounding where a sample lands within
    rounding of them."""
    from luce_render import render, scenes
    from luce_render.loading import Loaded
    sc, st, img = _render(device, scenes.graph_set(), width=48, height=36, samples=16, spectral=spectral,
                          max_bounces=8, light_tracing=0.0)
    pixels = np.arange(48 * 36, dtype=np.uint32)
    frame = render.ray_frame(sc.cameras[0], st.width, st.height)
    loaded = Loaded(sc, st, frame)
    ref = O.render(sc, loaded, frame.pack(), st.width, st.height, st.samples, st.max_bounces, st.spectral, st.seed,
                   pixels)
    got = img.reshape(-1, 4)[pixels, :3].astype(np.float64)
    per_pixel = (np.abs(got - ref) / (np.abs(ref) + 1e-6)).max(axis=1)
    close = (per_pixel < 1e-3).mean()
    say(f"graph spectral={spectral}: mean gpu {got.mean(0)} oracle {ref.mean(0)}, within 1e-3: {close:.4f}")
    assert np.all(np.abs(got.mean(0) - ref.mean(0)) <= 0.005 * ref.mean(0))
    assert close >= 0.98


def _floor_window(device, material, spectral):
    """The reference's floor_window: the green mean of the 4 x 4 window at
    (14, 14) of a 32 x 32, 1024-sample, one-bounce render."""
    from luce_render import scenes
    _, _, img = _render(device, scenes.graph_floor_set(material), width=32, height=32, samples=1024,
                        spectral=spectral, max_bounces=1, light_tracing=0.0)
    return float(img[14:18, 14:18, 1].astype(np.float64).mean())


@pytest.mark.parametrize("spectral", [False, True])
def test_textured_floor_lights_as_its_grey(device, spectral):
    """tests/gpu/integrator.lucb:594-637: a 4 x 4 image of sRGB 188s (0.502
    linear) on the floor by its uvs lights as a material of that grey."""
    grey = np.full((4, 4, 4), 188.0 / 255.0)
    grey[:, :, 3] = 1.0
    albedo = ((188.0 / 255.0 + 0.055) / 1.055) ** 2.4
    graph = "node grey image grey; node surface openpbr_surface; link surface base_color grey color"
    a = _floor_window(device, {"specular_weight": 0.0, "shader_graph": graph, "images": {"grey": grey}}, spectral)
    b = _floor_window(device, {"specular_weight": 0.0, "base_color": (albedo, albedo, albedo)}, spectral)
    say(f"textured floor (spectral {spectral}): {a} against a grey material's {b}")
    assert abs(a / b - 1.0) < 0.005


@pytest.mark.parametrize("spectral", [False, True])
def test_every_node_kind_lights_as_grey(device, spectral):
    """tests/gpu/integrator.lucb:639-660."""
    from test_shader_graph import GRAPH_NODES
    a = _floor_window(device, {"specular_weight": 0.0, "shader_graph": GRAPH_NODES}, spectral)
    b = _floor_window(device, {"specular_weight": 0.0, "base_color": (0.5, 0.5, 0.5)}, spectral)
    say(f"graph nodes (spectral {spectral}): {a} against a grey material's {b}")
    assert abs(a / b - 1.0) < 0.005


def test_bump_of_a_slope_lights_as_the_tilted_normal(device):
    """tests/gpu/integrator.lucb:662-699: bumps of a height rising 0.5 a unit
    along x (by position, by uv) light as the normal (-0.5, 1, 0) given
    outright; a constant height leaves the normal alone (lights differently)."""
    graphs = [
        "node c color; input c color -0.4472136 0.8944272 0; node surface openpbr_surface; "
        "link surface geometry_normal c color",
        "node p position; node s separate; link s color p position; node h math multiply; link h a s r; "
        "input h b 0.5; node b bump; link b height h value; node surface openpbr_surface; "
        "link surface geometry_normal b normal",
        "node u uv; node s separate; link s color u uv; node h math multiply; link h a s r; input h b 10; "
        "node b bump; link b height h value; node surface openpbr_surface; link surface geometry_normal b normal",
        "node b bump; input b height 0.3; node surface openpbr_surface; link surface geometry_normal b normal"]
    means = [_floor_window(device, {"specular_weight": 0.0, "shader_graph": g}, False) for g in graphs]
    say(f"bump slope: tilted {means[0]}, by position {means[1]}, by uv {means[2]}, flat {means[3]}")
    assert abs(means[1] / means[0] - 1.0) < 0.005 and abs(means[2] / means[0] - 1.0) < 0.005
    assert abs(means[3] / means[0] - 1.0) > 0.02


@pytest.mark.parametrize("spectral", [False, True])
def test_viewed_graph_light_tracing_is_unbiased(device, spectral):
    """tests/gpu/integrator.lucb:701-730: a material whose graph reads the
    view (Layer Weight's facing picks its colour, Fresnel its roughness) on a
    sphere and the floor: light paths connect there but scatter no further,
    camera paths count no light-path density through it; the mean of the
    light-traced image equals the camera-only one within 0.5%."""
    from luce_render import scenes
    graph = ("node w layer_weight; input w blend 0.3; node m mix; link m fac w facing; input m a 0.8 0.7 0.6; "
             "input m b 0.1 0.2 0.4; node f fresnel; input f ior 1.45; node r map_range; link r value f fac; "
             "input r to_min 0.6; input r to_max 0.05; node surface openpbr_surface; link surface base_color m color; "
             "link surface specular_roughness r value")
    material = {"specular_weight": 1.0, "shader_graph": graph}
    means = []
    for way in range(2):
        _, _, img = _render(device, scenes.viewed_graph_set(material), width=48, height=32, samples=2048,
                            spectral=spectral, max_bounces=3, light_tracing=1.0 if way else 0.0, seed=way + 1)
        means.append(float(img[:, :, 1].astype(np.float64).sum()))
    say(f"viewed graph (spectral {spectral}): light paths give {means[1] / means[0]} of the camera paths' mean")
    assert abs(means[1] / means[0] - 1.0) < 0.005


def test_graph_request_served_beside_plain_requests(device):
    """The serving loop with a node-graph request in the mix: the shared
    launches take the program-running shading kernels; the graph request
    renders the same bits alone and mixed (camera paths), and the plain
    request matches its own alone render."""
    from luce_render import render, scenes, serving
    from luce_render.loading import RenderSettings
    sets = [render.compile_scene(scenes.graph_set(frequency=12)), render.compile_scene(scenes.cornell_render_set())]
    gpus = [render.GpuScene.upload(device, sc) for sc in sets]

    def serve(jobs):
        server = serving.Server(device, budget_bytes=8 << 30)
        try:
            ids = [server.submit(gpus[k], sets[k], RenderSettings(w, h, spp, max_bounces=6, spectral=False,
                                                                  light_tracing=0.0)) for k, w, h, spp in jobs]
            steps = 0
            while server.busy() or steps == 0:
                server.step()
                steps += 1
                assert steps < 1000
            out = [server.read(i, w, h).copy() for i, (k, w, h, spp) in zip(ids, jobs)]
            for i in ids:
                server.release(i)
            return out
        finally:
            server.destroy()

    graph_job, plain_job = (0, 64, 48, 8), (1, 48, 48, 6)
    (alone_graph,) = serve([graph_job])
    (alone_plain,) = serve([plain_job])
    mixed = serve([plain_job, graph_job])

    # The server copies a request's programs and textures at submission: the
    # caller's arrays are gone (collected and overwritten) before the step
    # that admits it, and it still renders the same bits.
    import gc
    server = serving.Server(device, budget_bytes=8 << 30)
    try:
        k, w, h, spp = graph_job
        rid = server.submit(gpus[k], sets[k], RenderSettings(w, h, spp, max_bounces=6, spectral=False,
                                                              light_tracing=0.0))
        gc.collect()
        churn = [np.full(1 << 16, 7.0e30, np.float32) for _ in range(64)]
        del churn
        gc.collect()
        while server.busy() or server.status(rid)["state"] != 2:
            server.step()
        late = server.read(rid, w, h).copy()
        server.release(rid)
    finally:
        server.destroy()
    for g in gpus:
        g.destroy()
    assert np.array_equal(late, alone_graph)
    assert np.array_equal(mixed[1], alone_graph)
    assert np.allclose(mixed[0], alone_plain, rtol=1e-6, atol=1e-7)
    assert float(alone_graph[:, :, :3].mean()) > 0.0
