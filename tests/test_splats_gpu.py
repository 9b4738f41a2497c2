"""Gaussian splats in the integrator on the GPU (shaders/splats.glsl; the
splat stages of the wavefront), against the f64 oracle's composite of the
splats' description (tests/gpu/splats.lucb): primary visibility of
overlapping, rotated, anisotropic splats with SH over an emissive quad; the
shadow a splat casts (1 - alpha of the light); a ball of 3,000 splats (several
k-buffer rounds a ray) against a background; light tracing through splats."""
import math

import numpy as np
import pytest

from oracle import splats as OS

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
    say(f"render {st.width}x{st.height} spp {st.samples} spectral {st.spectral} lt {st.light_tracing} "
        f"splats {len(sc.splats.records)}")
    return sc, st, render.render(device, sc, st)


def _unit(v):
    return v / np.linalg.norm(v)


def _compare(sc, st, img, expected_at, bound, share=0.5):
    """The reference's comparison: every pixel against the CPU's value at its
    centre, skipping pixels the image changes across (a corner 0.45 px away
    differing by more than 0.02); rms under bound / 2, worst under 2.5 bound,
    at least `share` of the pixels compared (half in the reference's)."""
    from luce_render import render
    frame = render.ray_frame(sc.cameras[0], st.width, st.height)
    worst, squares, compared = 0.0, 0.0, 0
    for y in range(st.height):
        for x in range(st.width):
            want = expected_at(frame, x + 0.5, y + 0.5)
            spread = max(np.linalg.norm(expected_at(frame, x + 0.5 + dx, y + 0.5 + dy) - want)
                         for dx, dy in ((0.45, 0.45), (-0.45, 0.45), (0.45, -0.45), (-0.45, -0.45)))
            if spread > 0.02:
                continue
            miss = float(np.linalg.norm(img[y, x, :3].astype(np.float64) - want))
            worst = max(worst, miss)
            squares += miss * miss
            compared += 1
    rms = math.sqrt(squares / max(compared, 1))
    say(f"rms {rms:.5f}, worst {worst:.5f} over {compared} pixels")
    assert rms <= bound * 0.5 and worst <= bound * 2.5 and compared >= share * st.width * st.height


@pytest.mark.parametrize("spectral", [False, True])
def test_primary_visibility_matches_the_cpu_composite(device, spectral):
    """tests/gpu/splats.lucb:143-224: six splats over an emissive quad, one
    hidden behind it; bounds 0.012 (RGB) and 0.03 (spectral: colours go to
    spectra and back)."""
    from luce_render import scenes
    splats = scenes.splat_primary_splats()
    sc, st, img = _render(device, scenes.splat_primary_set(splats), width=96, height=64, samples=64,
                          spectral=spectral, max_bounces=1)
    assert len(sc.splats.records) == 6
    emitted = np.array(scenes.SPLAT_QUAD_EMISSION)
    z, h = scenes.SPLAT_QUAD_Z, scenes.SPLAT_QUAD_HALF

    def expected(frame, x, y):
        o, toward = frame.ray(x, y)
        d = _unit(toward)
        t_quad = (z - o[2]) / d[2]
        p = o + d * t_quad
        on_quad = t_quad > 0 and abs(p[0]) <= h and abs(p[1]) <= h
        c, _ = OS.composite(splats, o, d, t_quad if on_quad else 1e30, emitted if on_quad else (0.0, 0.0, 0.0))
        return c

    _compare(sc, st, img, expected, 0.03 if spectral else 0.012)


def _floor_point(frame, x, y):
    o, d = frame.ray(x, y)
    if d[1] >= 0:
        return None
    p = o + d * (-o[1] / d[1])
    if abs(p[0]) > 3.9 or abs(p[2]) > 3.9:
        return None
    return np.array([p[0], 0.0, p[2]])


def test_a_splat_shadow_keeps_one_minus_alpha(device):
    """tests/gpu/splats.lucb:226-297: a black splat of opacity 1 and sigma
    0.25 one unit above a white floor under the sun: at distance r from under
    its centre the floor keeps 1 - min(0.99, exp(-r^2 / 2 sigma^2)) of the
    light (times what the camera's own ray passes), camera paths only and
    with light tracing (the light paths' connections carry the same
    transmittance; the two images differ, so light paths ran)."""
    images = [_shadow_check(device, lt) for lt in (0.0, 0.5)]
    assert not np.array_equal(images[0], images[1])


def _shadow_check(device, light_tracing):
    from luce_render import render, scenes
    gset, splats = scenes.splat_shadow_set()
    sc, st, img = _render(device, gset, width=96, height=64, samples=64 if light_tracing == 0 else 256,
                          max_bounces=1, light_tracing=light_tracing)
    frame = render.ray_frame(sc.cameras[0], st.width, st.height)
    opened = []
    for y in range(st.height):
        for x in range(st.width):
            p = _floor_point(frame, x + 0.5, y + 0.5)
            if p is not None and 1.5 < np.linalg.norm(p) < 3.0:
                opened.append(img[y, x, 1])
    assert len(opened) > 50
    open_level = float(np.mean(opened))
    checked = 0
    for y in range(st.height):
        for x in range(st.width):
            p = _floor_point(frame, x + 0.5, y + 0.5)
            if p is None:
                continue
            r = float(np.linalg.norm(p))
            if r > 0.9:
                continue
            o, toward = frame.ray(x + 0.5, y + 0.5)
            seen, _ = OS.alpha(splats, o, _unit(toward))
            kept = (1.0 - min(0.99, math.exp(-0.5 * r * r / 0.0625))) * (1.0 - seen[0])
            got = float(img[y, x, 1]) / open_level
            slope = abs(0.99 * math.exp(-0.5 * r * r / 0.0625) * r / 0.0625) * 0.05
            assert abs(got - kept) <= 0.02 + slope + (0.02 if light_tracing else 0.0), (r, got, kept)
            checked += 1
    say(f"open floor {open_level:.4f}, checked {checked} pixels")
    assert checked > 20
    return img


def test_ball_of_splats_matches_the_cpu_composite(device):
    """3,000 random splats (rays meet up to hundreds: several rounds of the
    32-hit buffer, roulette past 99% opacity) in front of a (0.1, 0.2, 0.3)
    background, RGB: the reference's comparison at bound 0.03 over the pixels
    the image is smooth across (a fifth or more here: the ball is fine
    grained), and the image mean within 1% of the CPU's."""
    from luce_render import scenes
    gset, splats = scenes.splat_ball_set()
    bg = np.array([0.1, 0.2, 0.3])
    sc, st, img = _render(device, gset, width=64, height=48, samples=256, spectral=False, max_bounces=1,
                          background=tuple(bg))

    def expected(frame, x, y):
        o, toward = frame.ray(x, y)
        c, _ = OS.composite(splats, o, _unit(toward), 1e30, bg)
        return c

    _compare(sc, st, img, expected, 0.03, share=0.2)
    from luce_render import render
    frame = render.ray_frame(sc.cameras[0], st.width, st.height)
    want = np.mean([expected(frame, x + 0.5, y + 0.5) for y in range(st.height) for x in range(st.width)], axis=0)
    got = img[:, :, :3].astype(np.float64).reshape(-1, 3).mean(0)
    say(f"ball mean gpu {got} cpu {want}")
    assert np.all(np.abs(got - want) <= 0.01 * want)
