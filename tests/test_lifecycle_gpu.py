"""The interactive lifecycle (integrator.lucb:463-597): restart through a new
view, reload of lights and materials, replacement of the scene. Each must
render exactly what a fresh renderer of the same inputs renders (camera paths
only, so no float-atomic order enters: bit for bit)."""
import dataclasses

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _settings(**kw):
    from luce_render.loading import RenderSettings
    base = dict(width=48, height=40, samples=6, max_bounces=5, spectral=False, light_tracing=0.0)
    base.update(kw)
    return RenderSettings(**base)


def _fresh(device, sc, st, frame):
    from luce_render import render
    r = render.Renderer.create_viewed(device, sc, st, frame)
    try:
        r.finish()
        return r.read().copy()
    finally:
        r.destroy()


def test_restart_equals_fresh_render(device):
    from luce_render import render, scenes
    sc = render.compile_scene(scenes.cornell_render_set())
    st = _settings()
    first = render.ray_frame(sc.cameras[0], 48, 40)
    moved = render.look_frame((0.3, 0.2, 2.8), (-0.1, -0.05, -1.0), (0.0, 1.0, 0.0), 0.45, 64, 32)
    r = render.Renderer.create_viewed(device, sc, st, first)
    try:
        r.advance()  # a pass in flight, forgotten by the restart
        r.restart(moved)
        assert r.samples_done() == 0
        r.finish()
        got = r.read().copy()
        assert got.shape == (32, 64, 4)
    finally:
        r.destroy()
    ref = _fresh(device, sc, dataclasses.replace(st, width=64, height=32), moved)
    assert np.array_equal(got, ref)


def test_restart_of_the_same_size_keeps_the_buffers(device):
    """A restart whose view fits the buffers (integrator.lucb:468-480: made
    again only when the view outgrows them): at the same size the image and
    display stay at their addresses and show the last frame until the first
    new pass; a smaller view keeps them too; each render still equals a fresh
    one bit for bit (camera paths alone), and moving the view back gives the
    first image again."""
    from luce_render import render, scenes
    sc = render.compile_scene(scenes.cornell_render_set())
    st = _settings()
    first = render.ray_frame(sc.cameras[0], 48, 40)
    moved = render.look_frame((0.3, 0.2, 2.8), (-0.1, -0.05, -1.0), (0.0, 1.0, 0.0), 0.45, 48, 40)
    r = render.Renderer.create_viewed(device, sc, st, first)
    try:
        r.finish()
        before = r.read().copy()
        im = r.image()
        r.restart(moved)
        assert r.samples_done() == 0
        again = r.image()
        assert (again.rgba, again.display) == (im.rgba, im.display)
        assert np.array_equal(r.read(), before)  # the last frame until a new pass
        r.finish()
        got = r.read().copy()
        # a smaller view fits the buffers too
        small = render.look_frame((0.3, 0.2, 2.8), (-0.1, -0.05, -1.0), (0.0, 1.0, 0.0), 0.45, 40, 24)
        r.restart(small)
        assert r.image().rgba == im.rgba
        r.finish()
        got_small = r.read().copy()
        assert got_small.shape == (24, 40, 4)
        r.restart(first)
        assert r.image().rgba == im.rgba
        r.finish()
        back = r.read().copy()
    finally:
        r.destroy()
    assert np.array_equal(got, _fresh(device, sc, st, moved))
    assert np.array_equal(got_small, _fresh(device, sc, dataclasses.replace(st, width=40, height=24), small))
    assert np.array_equal(back, before)


def test_a_growing_view_gets_a_quarter_more_room(device):
    """integrator.lucb:474-478, 986-989: a restart into a view with more
    pixels than the buffers makes them again with a quarter more room each way
    (rounded up to 64: 64 x 48 gets 128 x 64), so the next larger view that
    fits (100 x 60) is taken in place; each render equals a fresh one bit for
    bit (camera paths alone). With light tracing the roomy buffers hold the
    larger view's light paths too: the image equals a fresh render's but for
    the order of the light paths' float additions."""
    from luce_render import render, scenes
    sc = render.compile_scene(scenes.cornell_render_set())
    for light_tracing in (0.0, 0.5):
        st = _settings(light_tracing=light_tracing)
        first = render.ray_frame(sc.cameras[0], 48, 40)
        grown = render.look_frame((0.3, 0.2, 2.8), (-0.1, -0.05, -1.0), (0.0, 1.0, 0.0), 0.45, 64, 48)
        more = render.look_frame((0.3, 0.2, 2.8), (-0.1, -0.05, -1.0), (0.0, 1.0, 0.0), 0.45, 100, 60)
        r = render.Renderer.create_viewed(device, sc, st, first)
        try:
            r.finish()
            im = r.image()
            r.restart(grown)
            roomy = r.image()
            assert roomy.rgba != im.rgba
            r.finish()
            got_grown = r.read().copy()
            r.restart(more)
            assert r.image().rgba == roomy.rgba
            r.finish()
            got_more = r.read().copy()
        finally:
            r.destroy()
        assert got_grown.shape == (48, 64, 4) and got_more.shape == (60, 100, 4)
        for got, frame in ((got_grown, grown), (got_more, more)):
            ref = _fresh(device, sc, dataclasses.replace(st, width=frame.width, height=frame.height), frame)
            say(f"light tracing {light_tracing}, {frame.width}x{frame.height}: mean {got[:, :, :3].mean()} "
                f"fresh {ref[:, :, :3].mean()}")
            if light_tracing == 0.0:
                assert np.array_equal(got, ref)
            else:
                assert np.allclose(got, ref, rtol=1e-4, atol=1e-5)


def test_reload_equals_fresh_render(device):
    from luce_render import render, scenes
    base = scenes.cornell_render_set()
    sc = render.compile_scene(base)
    brighter = scenes.cornell_render_set(light_radiance=40.0)
    brighter.materials["white"] = {"base_color": (0.5, 0.6, 0.7), "specular_weight": 0.0}
    sc2 = render.compile_scene(brighter)
    st = _settings()
    frame = render.ray_frame(sc.cameras[0], 48, 40)
    st2 = _settings(samples=4, max_bounces=3, seed=9)
    r = render.Renderer.create_viewed(device, sc, st, frame)
    try:
        r.finish()
        r.reload(sc2, st2)
        r.finish()
        got = r.read().copy()
    finally:
        r.destroy()
    ref = _fresh(device, sc2, st2, frame)
    assert np.array_equal(got, ref)
    with pytest.raises(render.LuceError) as e:
        r2 = render.Renderer.create_viewed(device, sc, st, frame)
        try:
            r2.reload(render.compile_scene(scenes.furnace_set()), st)
        finally:
            r2.destroy()
    assert e.value.code == 400


def test_replace_scene_equals_fresh_render(device):
    from luce_render import render, scenes
    sc = render.compile_scene(scenes.cornell_render_set())
    other = render.compile_scene(scenes.furnace_set())
    st = _settings()
    frame = render.ray_frame(sc.cameras[0], 48, 40)
    st2 = _settings(background=(1.0, 1.0, 1.0))
    frame2 = render.ray_frame(other.cameras[0], 40, 30)
    r = render.Renderer.create_viewed(device, sc, st, frame)
    try:
        r.finish()
        r.replace_scene(other, st2, frame2)
        r.finish()
        got = r.read().copy()
    finally:
        r.destroy()
    ref = _fresh(device, other, dataclasses.replace(st2, width=40, height=30), frame2)
    assert np.array_equal(got, ref)


def test_restart_keeps_the_display_settings(device):
    """A restart keeps what set_display wrote (integrator.lucb:463-466 uploads
    the view rows alone; :753-756 writes the display row into the loaded
    constants): after a restart in place, and after one whose view outgrows
    the buffers, the display equals that of a fresh render with the same
    exposure set before its first pass."""
    from luce_render import render, scenes
    sc = render.compile_scene(scenes.cornell_render_set())
    st = _settings(samples=3)
    first = render.ray_frame(sc.cameras[0], 48, 40)
    moved = render.look_frame((0.3, 0.2, 2.8), (-0.1, -0.05, -1.0), (0.0, 1.0, 0.0), 0.45, 48, 40)
    grown = render.look_frame((0.3, 0.2, 2.8), (-0.1, -0.05, -1.0), (0.0, 1.0, 0.0), 0.45, 96, 64)

    def fresh_display(frame):
        r = render.Renderer.create_viewed(device, sc, dataclasses.replace(st, width=frame.width,
                                                                           height=frame.height), frame)
        try:
            r.set_display(1.0, 1)
            r.finish()
            return r.read_display().copy()
        finally:
            r.destroy()

    r = render.Renderer.create_viewed(device, sc, st, first)
    try:
        r.set_display(1.0, 1)
        r.finish()
        r.restart(moved)
        r.finish()
        in_place = r.read_display().copy()
        r.restart(grown)
        r.finish()
        regrown = r.read_display().copy()
    finally:
        r.destroy()
    assert np.array_equal(in_place, fresh_display(moved))
    assert np.array_equal(regrown, fresh_display(grown))
    plain = _fresh(device, sc, st, moved)
    assert plain.shape == in_place.shape


def test_a_view_of_more_slots_than_one_launch_grid_spans(device):
    """A 3840x2160 render in one band (8.3M path slots, more than the
    4.19M threads of 65,535 workgroups of 64) over a tree whose stack bound
    exceeds the LDS stack, so rays spill: every slot has its own spill column,
    and the one-band render equals the same render in four bands bit for bit
    (camera paths only)."""
    from luce_render import render, scenes
    sc = render.compile_scene(scenes.cornell_render_set(random_count=10000))
    gpu = render.GpuScene.upload(device, sc)
    try:
        bound = gpu.info()["stack_bound"]
        say(f"stack bounds {bound}")
        assert max(bound) > 16
        frame = render.ray_frame(sc.cameras[0], 3840, 2160)
        st = _settings(width=3840, height=2160, samples=1, max_bounces=2, trace=(2, 64, 16))
        images = []
        for pool in (3840 * 2160, 1 << 21):
            say(f"3840x2160 pool {pool}")
            r = render.Renderer.create_viewed(device, sc, dataclasses.replace(st, pool=pool), frame, gpu_scene=gpu)
            try:
                r.finish()
                images.append(r.read().copy())
            finally:
                r.destroy()
    finally:
        gpu.destroy()
    assert np.isfinite(images[0]).all() and images[0][..., :3].max() > 0
    assert np.array_equal(images[0], images[1])
