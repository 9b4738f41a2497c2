"""The oracle's light tracing (light_emit.comp:21-137, light_paths.glsl:15-189,
connect.glsl:30-151, and the camera paths' side of common.glsl:238-242's
three-technique power heuristic) pinned without the reference: the weights
of the three ways a path can be made sum to one, so light tracing on or off
renders the same image mean; it lights a floor from an emissive square as
the closed-form irradiance says; and dropping the light tracer's term from
the light sample's weight, a planted mistake, breaks the first."""
import math

import numpy as np
import pytest

from oracle import oracle as O


def _mean(sc, frame, width, height, samples, lt, spectral=False, **kw):
    from luce_render.loading import Loaded, RenderSettings
    st = RenderSettings(width, height, samples, spectral=spectral, light_tracing=lt, **kw)
    ld = Loaded(sc, st, frame)
    px = np.arange(width * height, dtype=np.uint32)
    return O.render(sc, ld, frame.pack(), width, height, samples, st.max_bounces, spectral, st.seed, px), ld


@pytest.mark.parametrize("spectral", [False, True])
def test_light_tracing_leaves_the_cornell_mean(spectral):
    from luce_render import render, scenes
    sc = render.compile_scene(scenes.cornell_render_set())
    frame = render.ray_frame(sc.cameras[0], 48, 48)
    base, _ = _mean(sc, frame, 48, 48, 256, 0.0, spectral, max_bounces=8)
    for lt in (0.5, 1.0):
        img, ld = _mean(sc, frame, 48, 48, 256, lt, spectral, max_bounces=8)
        assert ld.light_paths > 0
        assert not np.allclose(img, base)  # the light paths are in the image
        assert np.all(np.abs(img.mean(0) / base.mean(0) - 1.0) < 0.005), (lt, img.mean(0), base.mean(0))
    O.plant(1)
    try:
        img, _ = _mean(sc, frame, 48, 48, 256, 0.5, spectral, max_bounces=8)
    finally:
        O.plant(0)
    assert np.all(np.abs(img.mean(0) / base.mean(0) - 1.0) > 0.05)


def test_light_paths_from_afar_leave_the_dome_mean():
    """A textured dome's light paths start from afar (light_emit.comp:33-61:
    a direction by the dome's texels, a point on a disk facing the scene);
    with them the image mean stays the camera paths' alone."""
    from luce_render import render, scenes
    metal = {"base_color": [0.9, 0.8, 0.6], "base_metalness": 1.0, "specular_roughness": 0.05}
    sc = render.compile_scene(scenes.dome_set(ball=metal))  # a caustic caster: light paths from afar aim at it
    frame = render.ray_frame(sc.cameras[0], 48, 36)
    base, _ = _mean(sc, frame, 48, 36, 256, 0.0, False, max_bounces=8)
    img, ld = _mean(sc, frame, 48, 36, 256, 1.0, False, max_bounces=8)
    assert ld.light_paths > 0 and not np.allclose(img, base)
    assert np.all(np.abs(img.mean(0) / base.mean(0) - 1.0) < 0.005), (img.mean(0), base.mean(0))


def test_light_paths_light_the_floor_as_the_closed_form():
    """The floor under an emissive 1x1 square of luminance 4 at height 1
    (tests/gpu/integrator.lucb:251-289): albedo/pi x 4 x 4 corner_irradiance
    (0.5, 0.5, 1) within 2%, with every pixel also reached by light paths."""
    from luce_render import render
    from luce_render.loading import OpenPbr
    from test_materials_gpu import _floor_scene, corner_irradiance, look_at, window
    g = _floor_scene(emissive_quad=True)
    g.cameras.append(render.Camera(look_at((0, 0.5, 3), (0, 0, 0)), "perspective", 200.0, 36.0))
    sc = render.compile_scene(g)
    frame = render.ray_frame(sc.cameras[0], 32, 32)
    img, _ = _mean(sc, frame, 32, 32, 1024, 1.0, False, max_bounces=1,
                   material=OpenPbr(base_color=(0.5, 0.5, 0.5), specular_weight=0.0))
    measured = window(img.reshape(32, 32, 3), 14, 14, 4)[1]
    expected = 0.5 / math.pi * 4.0 * 4.0 * corner_irradiance(0.5, 0.5, 1.0)
    assert abs(measured / expected - 1.0) <= 0.02, (measured, expected)
