"""Path guiding (the vMF mixture grid, shaders/guide.glsl) on the GPU: it must
keep a furnace white, and at equal samples it must bring a hard-to-light
room closer to a long reference than BSDF sampling alone
(tests/gpu/integrator.lucb:87-109 and :390-428)."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.mark.parametrize("guiding", [1, 2])
@pytest.mark.parametrize("spectral", [False, True])
def test_guided_furnace_stays_white(device, spectral, guiding):
    """A white EON sphere of roughness 1 under a constant environment
    reflects everything; guided sampling (one-sample MIS with the BSDF) must
    not change that. Tolerances as the reference's: image mean within 0.01,
    the sphere's centre within 0.02."""
    from luce_render import render, scenes
    from luce_render.loading import RenderSettings
    sc = render.compile_scene(scenes.furnace_set(albedo=1.0, roughness=1.0))
    st = RenderSettings(64, 48, 64, max_bounces=64, spectral=spectral, background=(1.0, 1.0, 1.0), guiding=guiding)
    say(f"guided furnace spectral={spectral} guiding={guiding}")
    img = render.render(device, sc, st)
    mean = img[:, :, :3].mean(axis=(0, 1))
    centre = img[16:32, 24:40, :3].mean(axis=(0, 1))
    say(f"  mean {mean}, centre {centre}")
    assert np.all(np.abs(mean - 1.0) <= 0.01), mean
    assert abs(centre[1] - 1.0) <= 0.02, centre
    assert np.all(np.isfinite(img))


@pytest.mark.parametrize("guiding", [1, 2])
def test_guided_room_is_closer_to_reference(device, guiding):
    """The lamp room at 256x192, two bounces, grey default material without
    specular, no light tracing: a 4096-sample unguided reference (seed 1),
    then four seeds at 256 samples each way. Guiding (the hash grid, or the
    learned guide trained on the first 8 samples) learns where the ceiling
    spot is; its summed squared error over the four seeds must be lower than
    unguided sampling's."""
    from luce_render import render, scenes
    from luce_render.loading import OpenPbr, RenderSettings
    sc = render.compile_scene(scenes.lamp_room_set())
    material = OpenPbr(base_color=(0.3, 0.3, 0.3), specular_weight=0.0)

    def run(samples, seed, way):
        st = RenderSettings(256, 192, samples, max_bounces=2, spectral=False, material=material, guiding=way,
                            light_tracing=0.0, seed=seed)
        return render.render(device, sc, st)[:, :, 1].astype(np.float64)

    say("lamp room reference, 4096 samples")
    ref = run(4096, 1, 0)
    assert ref.mean() > 0.0 and np.all(np.isfinite(ref))
    errors = [0.0, 0.0]
    for seed in range(2, 6):
        for way in (0, 1):
            say(f"lamp room seed {seed} guiding {guiding if way else 0}")
            errors[way] += float(((run(256, seed, guiding if way else 0) - ref) ** 2).sum())
    say(f"lamp room squared error: guided ({guiding}) {errors[1]:.4g}, unguided {errors[0]:.4g} "
        f"({errors[0] / max(errors[1], 1e-30):.2f}x better)")
    assert errors[1] < errors[0]


def test_learned_render_is_repeatable(device):
    """A learned-guide render trains its network during the render; two
    renders with one seed must give the same image bit for bit (the batch's
    weight sum is added in a fixed order, and every gradient wave writes its
    own row: no float atomics decide the network)."""
    from luce_render import render, scenes
    from luce_render.loading import OpenPbr, RenderSettings
    sc = render.compile_scene(scenes.lamp_room_set())
    st = RenderSettings(192, 144, 24, max_bounces=2, spectral=False, guiding=2, light_tracing=0.0, seed=3,
                        material=OpenPbr(base_color=(0.3, 0.3, 0.3), specular_weight=0.0))
    first = render.render(device, sc, st)
    for _ in range(2):
        again = render.render(device, sc, st)
        assert np.array_equal(first, again), float(np.abs(first - again).max())
    assert np.all(np.isfinite(first)) and first[:, :, :3].mean() > 0.0


def _guided_cornell(device, spectral, samples):
    """The Cornell box at 96x96, camera paths, 8 bounces, grid-guided: the
    GPU image and guide grid, and the oracle's (image, grid state)."""
    from luce_render import render, scenes
    from luce_render.loading import Loaded, RenderSettings
    sc = render.compile_scene(scenes.cornell_render_set())
    st = RenderSettings(96, 96, samples, max_bounces=8, spectral=spectral, light_tracing=0.0, guiding=1,
                        pass_samples=1)  # the reference's schedule: the grid learns after each sample
    r = render.Renderer.create(device, sc, st)
    try:
        r.finish()
        img = r.read().reshape(-1, 4)[:, :3].astype(np.float64)
        grid = r.read_guide()
    finally:
        r.destroy()
    frame = render.ray_frame(sc.cameras[0], 96, 96)
    pixels = np.arange(96 * 96, dtype=np.uint32)
    state = {}
    ref = O.render(sc, Loaded(sc, st, frame), frame.pack(), 96, 96, samples, 8, spectral, st.seed, pixels,
                   guided=True, guide_state=state)
    plain = O.render(sc, Loaded(sc, RenderSettings(96, 96, samples, max_bounces=8, spectral=spectral,
                                                   light_tracing=0.0), frame),
                     frame.pack(), 96, 96, samples, 8, spectral, st.seed, pixels)
    return img, grid, ref, (state["keys"], state["lobes"], state["stats"]), plain


@pytest.mark.parametrize("spectral", [False, True])
def test_guide_grid_after_one_pass_matches_the_oracle(device, spectral):
    """After the first sample the hash grid (guide.glsl:121-139 trained by
    every camera path's first three vertices, guide_fit.comp:16-35 fitted
    and decayed) equals the oracle's cell for cell: the same cells hold
    samples, each lobe the same number of them; weights and weighted
    directions within 1e-4 (the GPU adds them by float atomics in another
    order); fitted means within 1e-4 and kappas within 1e-3 relative for
    lobes with at least 1% of their cell's weight."""
    _, (gk, gl, gs), _, (ok, ol, os_), _ = _guided_cornell(device, spectral, 1)
    gcells = {int(k): i for i, k in enumerate(gk) if k and gs[i, :, 4].sum() > 0}
    ocells = {int(k): i for i, k in enumerate(ok) if k and os_[i, :, 4].sum() > 0}
    say(f"guide after one pass (spectral {spectral}): {len(gcells)} trained cells on the GPU, {len(ocells)} in the oracle")
    assert len(ocells) > 100 and set(gcells) == set(ocells)
    checked = 0
    for key, a in gcells.items():
        b = ocells[key]
        assert np.array_equal(gs[a, :, 4], os_[b, :, 4]), key
        assert np.allclose(gs[a, :, :4], os_[b, :, :4], rtol=1e-4, atol=1e-6), key
        share = gl[a, 4] / gl[a, 4].sum()
        for k in range(4):
            if share[k] < 0.01:
                continue
            assert np.allclose(gl[a, k, :3], ol[b, k, :3], atol=1e-4), (key, k)
            assert abs(gl[a, k, 3] - ol[b, k, 3]) <= 1e-3 * ol[b, k, 3], (key, k)
            checked += 1
        assert np.allclose(gl[a, 4], ol[b, 4], rtol=1e-4, atol=1e-6)
    assert checked > 200


@pytest.mark.parametrize("spectral", [False, True])
def test_grid_guided_cornell_matches_the_oracle(device, spectral):
    """The guided render per pixel against the oracle at 4 samples (the grid
    guides samples 1 to 3): mean within 0.5%, 99% of pixels within 1e-3.
    Later samples drift apart: a lobe fitted to few samples is sharp (kappa
    up to 1000), so the stray last bits of the atomics' sums decide which
    lobe a later sample joins, and the grids of the two then differ in shape
    (both unbiased: at 16 samples the means agree within 0.5%)."""
    img, _, ref, _, plain = _guided_cornell(device, spectral, 4)
    per_pixel = (np.abs(img - ref) / (np.abs(ref) + 1e-6)).max(axis=1)
    close = (per_pixel < 1e-3).mean()
    changed = (np.abs(ref - plain).max(axis=1) > 1e-6).mean()
    say(f"guided Cornell (spectral {spectral}) at 4 samples: mean gpu {img.mean(0)} oracle {ref.mean(0)}; "
        f"within 1e-3: {close:.4f}; guiding changed {changed:.3f} of the oracle's pixels")
    assert changed > 0.5  # the guide is in use
    assert np.all(np.abs(img.mean(0) - ref.mean(0)) <= 0.005 * ref.mean(0))
    assert close >= 0.99
    img, _, ref, _, _ = _guided_cornell(device, spectral, 16)
    say(f"  at 16 samples: mean gpu {img.mean(0)} oracle {ref.mean(0)}")
    assert np.all(np.abs(img.mean(0) - ref.mean(0)) <= 0.005 * ref.mean(0))


def _glossy(device, width, height):
    from luce_render import render, scenes
    sc = render.compile_scene(scenes.glossy_set())
    gpu = render.GpuScene.upload(device, sc)

    def run(samples, seed, guiding):
        from luce_render.loading import RenderSettings
        st = RenderSettings(width, height, samples, spectral=False, guiding=guiding, seed=seed, max_bounces=8,
                            light_tracing=0.5)
        say(f"glossy {width}x{height} {samples} spp guiding {guiding} seed {seed}")
        return render.render(device, sc, st, gpu_scene=gpu)[:, :, :3].astype(np.float64)
    return gpu, run


def test_glossy_guides_are_unbiased(device):
    """The learned-guiding workload (the Cornell geometry, random OpenPBR
    materials of roughness 0.05-0.3, a third of them metal, eight small
    sphere lights hidden behind occluders) at 240x136: grid-guided and
    learned-guided renders of 512 samples keep a 2048-sample plain
    reference's image mean within 1% and each quadrant's within 3%."""
    gpu, run = _glossy(device, 240, 136)
    try:
        ref = run(2048, 1, 0)
        for mode in (1, 2):
            img = run(512, 2, mode)
            whole = img.mean() / ref.mean()
            quads = [img[y:y + 68, x:x + 120].mean() / ref[y:y + 68, x:x + 120].mean() for y in (0, 68) for x in (0, 120)]
            say(f"  guiding {mode}: mean ratio {whole:.4f}, quadrants {np.round(quads, 4)}")
            assert abs(whole - 1.0) < 0.01
            assert np.all(np.abs(np.array(quads) - 1.0) < 0.03)
    finally:
        gpu.destroy()


def test_trained_network_beats_the_mixture_grid_on_glossy(device):
    """What the trained network gains over the method it stands in for: on
    the glossy workload at equal samples (64, three seeds) the learned
    guide's squared error against a long plain reference is lower than the
    hash grid's (the grid mixes its lobes into every lobe's sampling at a
    fixed share; the network only into the diffuse lobe's share)."""
    gpu, run = _glossy(device, 240, 136)
    try:
        ref = run(2048, 1, 0)
        err = {1: 0.0, 2: 0.0}
        for seed in (2, 3, 4):
            for mode in (1, 2):
                err[mode] += float(((run(64, seed, mode) - ref) ** 2).sum())
        say(f"  squared error: grid {err[1]:.4g}, learned {err[2]:.4g} ({err[1] / err[2]:.2f}x lower)")
        assert err[2] < err[1]
    finally:
        gpu.destroy()
