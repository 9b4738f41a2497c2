"""Path guiding (the vMF mixture grid, shaders/guide.glsl) on the GPU: it must
keep a furnace white, and at equal samples it must bring a hard-to-light
room closer to a long reference than BSDF sampling alone
(tests/gpu/integrator.lucb:87-109 and :390-428)."""
import numpy as np
import pytest

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
