"""GPU parity of the wavefront integrator: closed forms (white furnace), the
f64 oracle integrator pixel by pixel at fixed Sobol seeds, and the
reference's self-consistency checks (spectral vs RGB, light tracing vs camera
paths alone, recorded vs launched passes)."""
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
    say(f"render {st.width}x{st.height} spp {st.samples} spectral {st.spectral} lt {st.light_tracing} graph {st.use_graph}")
    img = render.render(device, sc, st)
    return sc, st, img


@pytest.mark.parametrize("spectral", [False, True])
def test_white_furnace(device, spectral):
    """tests/gpu/integrator.lucb:87-109 with albedo 0.8: a convex Lambert
    sphere under a constant environment L = 1 is exactly 0.8 L; the
    environment itself 1."""
    from luce_render import scenes
    _, _, img = _render(device, scenes.furnace_set(), width=64, height=48, samples=32, spectral=spectral,
                        background=(1.0, 1.0, 1.0), max_bounces=8)
    centre = img[16:32, 24:40, :3].mean(axis=(0, 1))
    corner = img[0:6, 0:6, :3].mean(axis=(0, 1))
    assert np.all(np.abs(centre - 0.8) < 0.01), centre
    assert np.all(np.abs(corner - 1.0) < 0.01), corner
    assert np.all(np.isfinite(img))


def _oracle_compare(sc, st, img, pixels):
    from luce_render import render
    from luce_render.loading import Loaded
    frame = render.ray_frame(sc.cameras[0], st.width, st.height)
    loaded = Loaded(sc, st, frame)
    ref = O.render(sc, loaded, frame.pack(), st.width, st.height, st.samples, st.max_bounces, st.spectral, st.seed,
                   pixels)
    got = img.reshape(-1, 4)[pixels, :3].astype(np.float64)
    return got, ref


@pytest.mark.parametrize("spectral", [False, True])
def test_cornell_matches_f64_oracle(device, spectral):
    """The Cornell workload against the f64 oracle at the same Sobol seeds,
    camera paths only: the same paths, so pixel values agree to f32 rounding
    except where rounding flips a decision (Russian roulette, a grazing hit).
    Tolerances: the image mean within 0.5%; 99% of pixels within 1e-3
    relative (plus 1e-6 absolute)."""
    from luce_render import scenes
    sc, st, img = _render(device, scenes.cornell_render_set(), width=96, height=96, samples=16, spectral=spectral,
                          max_bounces=8, light_tracing=0.0)
    pixels = np.arange(96 * 96, dtype=np.uint32)
    got, ref = _oracle_compare(sc, st, img, pixels)
    rel = np.abs(got - ref) / (np.abs(ref) + 1e-6)
    per_pixel = rel.max(axis=1)
    close = (per_pixel < 1e-3).mean()
    say(f"cornell spectral={spectral}: mean gpu {got.mean(0)} oracle {ref.mean(0)}, within 1e-3: {close:.4f}, "
        f"median rel {np.median(per_pixel):.2e}")
    assert np.all(np.abs(got.mean(0) - ref.mean(0)) <= 0.005 * ref.mean(0))
    assert close >= 0.99


@pytest.mark.parametrize("spectral", [False, True])
def test_many_lights_match_f64_oracle(device, spectral):
    """Forty rect, disk and sphere lights inside the Cornell box, seen
    directly and through bounces: the closest hit's culled light loop must
    find the same light (or triangle) as the oracle's plain loop over every
    light. Same tolerances as the Cornell comparison."""
    from luce_render import scenes
    sc, st, img = _render(device, scenes.many_lights_set(), width=64, height=64, samples=4, spectral=spectral,
                          max_bounces=4, light_tracing=0.0)
    pixels = np.arange(64 * 64, dtype=np.uint32)
    got, ref = _oracle_compare(sc, st, img, pixels)
    rel = np.abs(got - ref) / (np.abs(ref) + 1e-6)
    per_pixel = rel.max(axis=1)
    close = (per_pixel < 1e-3).mean()
    say(f"many lights spectral={spectral}: mean gpu {got.mean(0)} oracle {ref.mean(0)}, within 1e-3: {close:.4f}")
    assert np.all(np.abs(got.mean(0) - ref.mean(0)) <= 0.005 * ref.mean(0))
    assert close >= 0.99


def test_tied_lights_resolve_as_the_reference_loop(device):
    """Two rect lights in one horizontal plane whose footprints overlap, the
    lower index at the larger x: a camera ray through the overlap meets both
    at exactly the same distance (axis-aligned frames, so the plane distance
    has no rounding between them). The reference tests lights in index order
    with a strict `<` (intersect_closest.comp:58-66), so the lower index is
    the one seen; the closest hit's light tree meets the higher index first
    (its centre has the smaller Morton code) and must still end on the lower
    one. Camera rays only, pixel by pixel against the f64 oracle's loop; the
    lights' radiances differ tenfold, so a wrong pick shows."""
    import dataclasses
    from luce_render import scenes
    gset = scenes.cornell_render_set()
    base = gset.lights[0]
    lights = []
    for x, radiance in ((0.15, 5.0), (-0.15, 50.0)):
        m = np.array(base.matrix, dtype=np.float64).copy()
        m[3, :3] = (x, 0.6, 0.0)
        lights.append(dataclasses.replace(base, matrix=m, intensity=radiance, width=0.5, height=0.5))
    gset.lights = lights
    sc, st, img = _render(device, gset, width=96, height=96, samples=1, spectral=False, max_bounces=1,
                          light_tracing=0.0)
    pixels = np.arange(96 * 96, dtype=np.uint32)
    got, ref = _oracle_compare(sc, st, img, pixels)
    lum = ref.mean(axis=1)
    dim = np.abs(lum - 5.0) < 1e-3 * 5.0  # light 0 seen
    bright = np.abs(lum - 50.0) < 1e-3 * 50.0
    say(f"tied lights: light 0 on {dim.sum()} pixels, light 1 on {bright.sum()} pixels")
    # Light 0 alone spans x in [0.1, 0.4], light 1 alone [-0.4, -0.1], the overlap [-0.1, 0.1] goes to light 0.
    assert bright.sum() > 10 and dim.sum() > 1.5 * bright.sum()
    assert np.allclose(got[dim], ref[dim], rtol=1e-5)
    assert np.allclose(got[bright], ref[bright], rtol=1e-5)
    per_pixel = (np.abs(got - ref) / (np.abs(ref) + 1e-6)).max(axis=1)
    assert (per_pixel < 1e-3).mean() >= 0.99


@pytest.mark.parametrize("spectral", [False, True])
@pytest.mark.parametrize("scene", ["cornell", "many_lights", "dome"])
def test_light_traced_render_matches_f64_oracle(device, scene, spectral):
    """Light tracing at half a light path a pixel (the reference's default)
    against the oracle's light paths and three-technique MIS
    (light_emit.comp, light_paths.glsl, connect.glsl): camera paths with
    their light-tracing weights plus every light path's splats, per pixel.
    The dome's light paths start from afar (a disk facing its metal ball). The
    splats add by float atomics on the GPU, in path order in the oracle. Tolerances as the camera-path comparison's: image mean within
    0.5%, 99% of pixels within 1e-3; the light paths change the oracle's
    image (the splat plane is compared)."""
    from luce_render import scenes
    metal = {"base_color": [0.9, 0.8, 0.6], "base_metalness": 1.0, "specular_roughness": 0.05}
    gset = {"cornell": scenes.cornell_render_set, "many_lights": scenes.many_lights_set,
            "dome": lambda: scenes.dome_set(ball=metal)}[scene]()
    width, height = {"cornell": (96, 96), "many_lights": (64, 64), "dome": (48, 36)}[scene]
    sc, st, img = _render(device, gset, width=width, height=height, samples=16, spectral=spectral, max_bounces=8,
                          light_tracing=0.5)
    pixels = np.arange(width * height, dtype=np.uint32)
    got, ref = _oracle_compare(sc, st, img, pixels)
    import dataclasses
    _, camera_only = _oracle_compare(sc, dataclasses.replace(st, light_tracing=0.0), img, pixels)
    per_pixel = (np.abs(got - ref) / (np.abs(ref) + 1e-6)).max(axis=1)
    close = (per_pixel < 1e-3).mean()
    moved = (np.abs(ref - camera_only).max(axis=1) > 1e-4 * (np.abs(ref).max(axis=1) + 1e-6)).mean()
    say(f"{scene} light traced spectral={spectral}: mean gpu {got.mean(0)} oracle {ref.mean(0)}, within 1e-3: "
        f"{close:.4f}; light paths change {moved:.3f} of the oracle's pixels")
    assert moved > (0.2 if scene == "dome" else 0.5)  # the dome's light paths reach what its ball reflects
    assert np.all(np.abs(got.mean(0) - ref.mean(0)) <= 0.005 * ref.mean(0))
    assert close >= 0.99


@pytest.mark.parametrize("spectral", [False, True])
def test_furnace_matches_f64_oracle(device, spectral):
    from luce_render import scenes
    sc, st, img = _render(device, scenes.furnace_set(), width=48, height=36, samples=8, spectral=spectral,
                          background=(1.0, 1.0, 1.0), max_bounces=8, light_tracing=0.0)
    pixels = np.arange(48 * 36, dtype=np.uint32)
    got, ref = _oracle_compare(sc, st, img, pixels)
    per_pixel = (np.abs(got - ref) / (np.abs(ref) + 1e-6)).max(axis=1)
    assert (per_pixel < 1e-4).mean() >= 0.97
    assert np.abs(got.mean() - ref.mean()) < 1e-3


def test_spectral_agrees_with_rgb(device):
    """tests/gpu/integrator.lucb:773-798: the two builds agree on the Cornell
    mean within the spectral upsampling's error (3%)."""
    from luce_render import scenes
    means = []
    for spectral in (False, True):
        _, _, img = _render(device, scenes.cornell_render_set(), width=96, height=96, samples=64, spectral=spectral,
                            max_bounces=8, light_tracing=0.0)
        means.append(img[..., :3].mean(axis=(0, 1)))
    say("rgb", means[0], "spectral", means[1])
    assert np.all(np.abs(means[1] - means[0]) <= 0.03 * means[0])


@pytest.mark.parametrize("spectral", [False, True])
def test_light_tracing_is_unbiased(device, spectral):
    """tests/gpu/integrator.lucb:434-465: with light paths (0.5 a pixel, MIS
    over three techniques) the image mean matches camera paths alone within
    1%, and its variance does not grow."""
    from luce_render import scenes
    _, _, off = _render(device, scenes.cornell_render_set(), width=96, height=96, samples=128, spectral=spectral,
                        max_bounces=8, light_tracing=0.0)
    _, _, on = _render(device, scenes.cornell_render_set(), width=96, height=96, samples=128, spectral=spectral,
                       max_bounces=8, light_tracing=0.5)
    m_off, m_on = off[..., :3].mean(axis=(0, 1)), on[..., :3].mean(axis=(0, 1))
    say("camera paths", m_off, "with light paths", m_on)
    assert np.all(np.abs(m_on - m_off) <= 0.01 * m_off)
    assert np.all(np.isfinite(on))


def test_recorded_pass_equals_launched_pass(device):
    """The hipGraph replay of a pass gives the image the plain launches give,
    bit for bit (camera paths only: no float-atomic splats)."""
    from luce_render import scenes
    _, _, a = _render(device, scenes.cornell_render_set(), width=64, height=64, samples=8, spectral=True,
                      max_bounces=6, light_tracing=0.0, use_graph=True)
    _, _, b = _render(device, scenes.cornell_render_set(), width=64, height=64, samples=8, spectral=True,
                      max_bounces=6, light_tracing=0.0, use_graph=False)
    assert np.array_equal(a, b)


def test_traversal_variants_render_the_same_image(device):
    """The integrator's closest-hit and shadow kernels at any layout, group
    and LDS stack (48 entries included) give the same image bit for bit
    (camera paths only), which is what lets the render autotuner choose by
    time alone; its table holds every variant it swept and it returns the
    fastest."""
    from luce_render import autotune, render, scenes
    from luce_render.loading import RenderSettings
    sc = render.compile_scene(scenes.cornell_render_set())
    gpu = render.GpuScene.upload(device, sc)
    try:
        ref = None
        for trace in ((4, 64, 16), (20, 128, 48), (24, 256, 32), (2, 64, 48), (8, 256, 48)):
            st = RenderSettings(64, 64, 4, max_bounces=6, spectral=False, light_tracing=0.0, trace=trace)
            say(f"render variant {trace}")
            img = render.render(device, sc, st, gpu_scene=gpu)
            if ref is None:
                ref = img
            assert np.array_equal(img, ref), trace
        st = RenderSettings(64, 64, 2, max_bounces=6, spectral=False, light_tracing=0.0)
        best, table = autotune.tune_render(device, sc, gpu, st, widths=(4, 20), groups=(64, 128), stacks=(16, 48))
        say(f"render autotune: {best} of {table}")
        assert set(table) == {(w, g, s) for w in (4, 20) for g in (64, 128) for s in (16, 48)}
        assert table[best] == min(table.values())
    finally:
        gpu.destroy()


def test_bands_equal_one_pass(device):
    """A pool smaller than the image traces it in bands (integrator.lucb:
    788-796): the same image as one band."""
    from luce_render import scenes
    _, _, a = _render(device, scenes.cornell_render_set(), width=64, height=64, samples=4, spectral=False,
                      max_bounces=6, light_tracing=0.0, pool=0)
    _, _, b = _render(device, scenes.cornell_render_set(), width=64, height=64, samples=4, spectral=False,
                      max_bounces=6, light_tracing=0.0, pool=1000)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("spectral", [False, True])
def test_passes_of_several_samples_equal_passes_of_one(device, spectral):
    """A pass that carries several samples of a one-band image (camera paths
    of each sample, then their light paths, in one set of launches) renders
    what one pass a sample renders: camera paths bit for bit, light paths
    within 1e-4 of the image's peak (their splats add by float atomics). 19
    samples at 8 a pass (passes of 8, 8 and 3) and at the default (16, 3)."""
    from luce_render import render, scenes
    from luce_render.loading import RenderSettings
    sc = render.compile_scene(scenes.cornell_render_set())
    for lt in (0.0, 0.5):
        images = []
        for k in (1, 8, 0):
            st = RenderSettings(96, 96, 19, max_bounces=8, spectral=spectral, light_tracing=lt, pass_samples=k)
            say(f"passes of {k or 'default'} samples, spectral {spectral}, light tracing {lt}")
            r = render.Renderer.create(device, sc, st)
            try:
                r.finish()
                assert r.samples_done() == 19
                stats = r.stats()
                assert stats["light_samples"] == (19 if lt else 0)
                images.append(r.read().copy())
            finally:
                r.destroy()
        for img in images[1:]:
            if lt == 0.0:
                assert np.array_equal(img, images[0])
            else:
                rel = np.abs(img - images[0]).max() / np.abs(images[0]).max()
                say(f"  max difference {rel:.2e} of the peak")
                assert rel < 1e-4


def test_renderer_lifecycle(device):
    """advance keeps at most two passes in flight and never waits; finish
    completes every sample; kernel_times charges the kinds with profile on."""
    from luce_render import render, scenes
    from luce_render.loading import RenderSettings
    sc = render.compile_scene(scenes.cornell_render_set())
    st = RenderSettings.of(48, 48, 6)
    st.profile = True
    r = render.Renderer.create(device, sc, st)
    try:
        while r.advance():
            pass
        assert r.samples_done() == 6
        times = r.kernel_times()
        assert times["intersect closest"] > 0 and times["shade surface"] > 0 and times["camera"] > 0
        img = r.read()
        assert img.shape == (48, 48, 4) and np.all(np.isfinite(img))
    finally:
        r.destroy()
    with pytest.raises(render.LuceError) as e:
        render.Renderer.create(device, sc, RenderSettings.of(0, 10, 1))
    assert e.value.code == 400
    bad = RenderSettings.of(8, 8, 1)
    bad.max_bounces = 0
    with pytest.raises(render.LuceError):
        render.Renderer.create(device, sc, bad)


def test_adaptive_sampling_stops_converged_pixels(device):
    """adaptive.comp with a noise threshold: the furnace's constant background
    has no variance, so its pixels converge at the first check (16 samples)
    and stop sampling, while the image keeps its values (each pixel the mean
    of the samples it took): background 1, sphere 0.8 (within 0.01), and
    fewer rays than the same render without the threshold."""
    from luce_render import render, scenes
    from luce_render.loading import RenderSettings
    sc = render.compile_scene(scenes.furnace_set())
    out = {}
    for threshold in (0.0, 0.05):
        st = RenderSettings(64, 48, 64, max_bounces=8, spectral=False, background=(1.0, 1.0, 1.0),
                            noise_threshold=threshold, adaptive_minimum=16, light_tracing=0.0)
        r = render.Renderer.create(device, sc, st)
        try:
            r.finish()
            out[threshold] = (r.converged_pixels(), r.read().copy(), r.stats())
        finally:
            r.destroy()
    conv0, _, stats0 = out[0.0]
    conv, img, stats = out[0.05]
    say(f"adaptive: converged {conv} of {64 * 48}, rays {stats['closest_rays']} vs {stats0['closest_rays']}")
    assert conv0 == 0
    assert conv > 64 * 48 // 4
    assert stats["closest_rays"] < stats0["closest_rays"]
    assert np.all(np.abs(img[0:6, 0:6, :3] - 1.0) < 0.01)
    assert np.all(np.abs(img[16:32, 24:40, :3].mean(axis=(0, 1)) - 0.8) < 0.01)


def test_adaptive_verdict_matches_the_oracle_on_a_noisy_and_smooth_image(device):
    """adaptive.comp on the Cornell box (smooth lit walls, noisy shadows and
    edges) after its one check at 16 samples: each pixel's own verdict (the
    scratch plane) and its converged flag after the 3 x 3 rule (the half
    plane's w) equal the oracle's restatement computed from the same film,
    pixel for pixel; the threshold is the median error, so about half the
    pixels pass alone and the neighbourhood rule stops fewer of them."""
    from luce_render import render, scenes
    from luce_render.loading import RenderSettings
    sc = render.compile_scene(scenes.cornell_render_set())
    films = {}
    for threshold in (0.0, None):
        if threshold is None:
            threshold = float(np.median(O.adaptive_error(films[0.0][0], films[0.0][1])))
        st = RenderSettings(64, 48, 16, max_bounces=6, spectral=False, light_tracing=0.0,
                            noise_threshold=threshold, adaptive_minimum=16)
        r = render.Renderer.create(device, sc, st)
        try:
            r.finish()
            films[threshold] = r.read_film().copy()
            conv_count = r.converged_pixels()
        finally:
            r.destroy()
    (_, first), (thr, film) = films.items()
    assert np.array_equal(film[0], first[0]) and np.array_equal(film[1][..., :3], first[1][..., :3])
    own, conv = O.adaptive_verdict(film[0], film[1], thr)
    err = O.adaptive_error(film[0], film[1])
    clear = np.abs(err - np.float32(thr)) > 1e-5 * thr
    got_own = film[2][..., 0] > 0.5
    got_conv = film[1][..., 3] > 0.5
    say(f"adaptive at threshold {thr:.4g}: own {own.sum()}, converged {conv.sum()} (GPU {got_conv.sum()}, "
        f"counted {conv_count}) of {own.size}")
    assert 0.3 * own.size < own.sum() < 0.7 * own.size and conv.sum() < own.sum()
    assert np.array_equal(got_own[clear], own[clear])
    assert np.array_equal(got_conv[clear], conv[clear]) and conv_count == got_conv.sum()


@pytest.mark.parametrize("spectral", [False, True])
def test_textured_dome_matches_f64_oracle(device, spectral):
    """A textured dome (environment.lucb, lights.glsl:89-175) over a Lambert
    sphere and floor: background lookups, light samples by the texel CDFs and
    their MIS pdfs against the f64 oracle at the same Sobol seeds. Same
    tolerances as the Cornell comparison (a sample that lands on a texel
    border within f32 rounding of atan2/acos may take the other texel)."""
    from luce_render import scenes
    sc, st, img = _render(device, scenes.dome_set(), width=48, height=36, samples=16, spectral=spectral,
                          max_bounces=8, light_tracing=0.0)
    pixels = np.arange(48 * 36, dtype=np.uint32)
    got, ref = _oracle_compare(sc, st, img, pixels)
    per_pixel = (np.abs(got - ref) / (np.abs(ref) + 1e-6)).max(axis=1)
    close = (per_pixel < 1e-3).mean()
    say(f"dome spectral={spectral}: mean gpu {got.mean(0)} oracle {ref.mean(0)}, within 1e-3: {close:.4f}")
    assert np.all(np.abs(got.mean(0) - ref.mean(0)) <= 0.005 * ref.mean(0))
    assert close >= 0.99


def test_textured_dome_light_tracing_is_unbiased(device):
    """Light paths that start from the textured dome (light_emit.comp:33-60:
    a direction by the dome's CDFs, a point on the disk facing the caustic
    casters, here a smooth metal ball over the floor) and their MIS leave the
    image mean unchanged: camera-only and light-traced renders agree within
    1%, and the light-traced one differs (light paths ran)."""
    from luce_render import scenes
    metal = {"base_color": [0.9, 0.8, 0.6], "base_metalness": 1.0, "specular_roughness": 0.05}
    means = []
    for lt in (0.0, 0.5):
        _, _, img = _render(device, scenes.dome_set(ball=metal), width=64, height=48, samples=256, spectral=False,
                            max_bounces=8, light_tracing=lt)
        means.append(img[:, :, :3].mean(axis=(0, 1)))
    say(f"dome camera-only {means[0]} light-traced {means[1]}")
    assert not np.array_equal(means[0], means[1])
    assert np.all(np.abs(means[1] - means[0]) <= 0.01 * means[0])


def test_set_display_converts_the_film_again(device):
    """Renderer.set_display (integrator.lucb:750-756, 656-669): once every
    sample is in, a change of exposure or view converts the film again without
    rendering (the image stays); one stop doubles the display below the clamp
    and saturates the rest; the neutral view changes it; set during the render,
    the next pass applies it (the same display as set afterwards)."""
    from luce_render import render, scenes
    from luce_render.loading import RenderSettings
    sc = render.compile_scene(scenes.cornell_render_set())
    st = RenderSettings(48, 40, 4, max_bounces=5, spectral=False, light_tracing=0.0)
    frame = render.ray_frame(sc.cameras[0], 48, 40)
    r = render.Renderer.create_viewed(device, sc, st, frame)
    try:
        r.set_display(0.0, 0)
        r.finish()
        image = r.read().copy()
        base = r.read_display()[:, :, :3].copy()
        r.set_display(1.0, 0)
        r.finish()
        assert r.samples_done() == 4 and np.array_equal(r.read(), image)
        bright = r.read_display()[:, :, :3].copy()
        r.set_display(0.0, 1)
        r.finish()
        neutral = r.read_display()[:, :, :3].copy()
        with pytest.raises(render.LuceError):
            r.set_display(0.0, 2)
    finally:
        r.destroy()
    low = base < 0.49
    say(f"display: {low.sum()} values below the clamp of {base.size}")
    assert low.sum() > base.size // 4 and (~low).sum() > 0
    assert np.allclose(bright[low], 2.0 * base[low], rtol=1e-6, atol=0.0)
    assert np.all(bright[~low] >= 0.98)
    assert not np.array_equal(neutral, base) and neutral.min() >= 0.0 and neutral.max() <= 1.0
    r = render.Renderer.create_viewed(device, sc, st, frame)
    try:
        r.set_display(0.0, 0)
        r.advance()
        r.set_display(1.0, 0)
        r.finish()
        during = r.read_display()[:, :, :3].copy()
    finally:
        r.destroy()
    assert np.array_equal(during, bright)
