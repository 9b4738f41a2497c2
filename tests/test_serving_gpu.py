"""The serving loop on the GPU: a request renders the same alone and in a
mix of other requests (scenes, resolutions, sample targets) sharing every
launch; admission respects the memory budget; cancellation frees slots."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def scenes(device):
    from luce_render import render, scenes as S
    out = []
    for g in (S.cornell_render_set(), S.cornell_render_set(random_count=10000), S.furnace_set()):
        sc = render.compile_scene(g)
        out.append((render.GpuScene.upload(device, sc), sc))
    yield out
    for g, _ in out:
        g.destroy()


def _serve(device, scenes, jobs, lt, budget=8 << 30, pass_samples=0):
    """jobs: [(scene index, w, h, spp)]; returns the images in order and the
    steps taken (a step runs one pass of every active request; a pass of a
    small request carries up to 16 samples unless pass_samples says)."""
    from luce_render import serving
    from luce_render.loading import RenderSettings
    server = serving.Server(device, budget_bytes=budget)
    try:
        ids = []
        for (k, w, h, spp) in jobs:
            st = RenderSettings(w, h, spp, max_bounces=6, spectral=False, light_tracing=lt,
                                background=(0.2, 0.2, 0.2) if k == 2 else (0.0, 0.0, 0.0), pass_samples=pass_samples)
            ids.append(server.submit(scenes[k][0], scenes[k][1], st))
        steps = 0
        while server.busy() or steps == 0:
            server.step()
            steps += 1
            assert steps < 1000
        out = [server.read(i, w, h).copy() for i, (k, w, h, spp) in zip(ids, jobs)]
        for i in ids:
            server.release(i)
        return out, steps
    finally:
        server.destroy()


@pytest.mark.parametrize("pass_samples", [1, 0])
@pytest.mark.parametrize("lt", [0.0, 0.5])
def test_alone_equals_mixed(device, scenes, lt, pass_samples):
    """A request renders the same alone and beside four others sharing every
    launch, with a sample a pass (20 steps for the longest request, 20
    samples) and with small requests' passes carrying up to 16 samples (2
    steps), alone and mixed alike."""
    target = (0, 96, 64, 12)
    mix = [(1, 128, 128, 5), target, (2, 64, 48, 20), (0, 32, 32, 3), (1, 200, 100, 12)]
    say(f"serving mix lt {lt} pass samples {pass_samples}")
    (alone,), _ = _serve(device, scenes, [target], lt, pass_samples=pass_samples)
    mixed, steps = _serve(device, scenes, mix, lt, pass_samples=pass_samples)
    got = mixed[1]
    assert steps == (20 if pass_samples == 1 else 2)
    if lt == 0.0:
        assert np.array_equal(alone, got)  # camera paths: the same bits
    else:
        # light paths splat with float atomics, whose order differs
        rel = np.abs(alone - got).max() / (np.abs(alone).max() + 1e-12)
        say("max relative difference", rel)
        assert rel < 1e-4


def test_budget_admits_in_order(device, scenes):
    """A budget that holds one 96x64 request at a time still finishes them all,
    with the same images."""
    from luce_render import serving
    from luce_render.loading import RenderSettings
    jobs = [(0, 96, 64, 4), (1, 96, 64, 3), (0, 96, 64, 2)]
    free, _ = _serve(device, scenes, jobs, 0.0, pass_samples=1)
    # find one request's footprint, then a budget of one and a half of it
    server = serving.Server(device, budget_bytes=8 << 30)
    rid = server.submit(scenes[0][0], scenes[0][1], RenderSettings(96, 64, 1, max_bounces=6, spectral=False,
                                                                     light_tracing=0.0, pass_samples=1))
    need = server.status(rid)["bytes"]
    server.release(rid)
    server.destroy()
    tight, steps = _serve(device, scenes, jobs, 0.0, budget=int(need * 1.5), pass_samples=1)
    assert steps == 4 + 3 + 2
    for a, b in zip(free, tight):
        assert np.array_equal(a, b)


def test_oversized_request_and_cancel(device, scenes):
    from luce_render import render, serving
    from luce_render.loading import RenderSettings
    server = serving.Server(device, budget_bytes=1 << 20)
    try:
        with pytest.raises(render.LuceError) as e:
            server.submit(scenes[0][0], scenes[0][1], RenderSettings(512, 512, 4, spectral=False))
        assert e.value.code == 401
    finally:
        server.destroy()
    server = serving.Server(device, budget_bytes=8 << 30)
    try:
        a = server.submit(scenes[0][0], scenes[0][1], RenderSettings(64, 64, 100, spectral=False, light_tracing=0.0))
        b = server.submit(scenes[1][0], scenes[1][1], RenderSettings(64, 64, 2, spectral=False, light_tracing=0.0))
        server.step()
        server.release(a)  # cancel while rendering
        while server.busy():
            server.step()
        assert server.status(b)["state"] == 2
        with pytest.raises(render.LuceError):
            server.status(a)
    finally:
        server.destroy()


def test_poisson_serving_reports_latency(device, scenes):
    """A short Poisson stream (40 requests a second) through the loop:
    every request finishes and the report is consistent."""
    from luce_render import serving
    sched = serving.poisson_schedule(count=24, rate=40.0, scenes=3, seed=5, spp_range=(4, 16),
                                     resolutions=[(64, 64), (128, 96)])
    rep = serving.run(device, scenes, sched, max_bounces=4, server_kw={"budget_bytes": 8 << 30})
    say(rep.p50_s, rep.p99_s, rep.rays_per_s, rep.peak_active)
    assert rep.requests == 24 and rep.p99_s >= rep.p50_s > 0 and rep.rays > 0


@pytest.fixture(scope="module")
def serving_scenes(device):
    """The serving workload's four scenes (scenes.serving_sets: the Cornell
    box with 10,000 random triangles, the plain Cornell render, and the
    Cornell render with 500,000 and 5,000,000 random triangles) as the
    server takes them (serving.prepare_scenes: the two clouds built over
    pieces of their triangles, the others from their triangles, all on the
    device)."""
    from luce_render import scenes as S, serving
    out = serving.prepare_scenes(device, S.serving_sets())
    for _, sc in out:
        say(f"serving scene: {sc.triangle_count()} leaf slots ({sc.split_pieces} pieces)")
    yield out
    for g, _ in out:
        g.destroy()


@pytest.mark.parametrize("lt", [0.0, 0.5])
def test_serving_scenes_alone_equal_mixed(device, serving_scenes, lt):
    """The serving workload's scenes at its resolutions, a 1920x1080 request
    on the 5M-triangle soup among them: each request renders the same alone
    and beside the others in shared launches (camera paths bit for bit;
    with light paths, whose splats add by float atomics, within 1e-4 of the
    image's peak)."""
    jobs = [(3, 1920, 1080, 2), (2, 512, 512, 4), (0, 256, 256, 6), (1, 1024, 768, 3)]
    mixed, steps = _serve(device, serving_scenes, jobs, lt, budget=64 << 30)
    assert steps == 1  # every request's samples fit one pass (the 1080p one's two)
    for k, job in enumerate(jobs):
        say(f"serving scenes lt {lt}: request {job} alone")
        (alone,), _ = _serve(device, serving_scenes, [job], lt, budget=64 << 30)
        got = mixed[k]
        assert np.isfinite(alone).all() and alone[..., :3].max() > 0.0
        if lt == 0.0:
            assert np.array_equal(alone, got), job
        else:
            rel = np.abs(alone - got).max() / (np.abs(alone).max() + 1e-12)
            say(f"  max relative difference {rel:.3g}")
            assert rel < 1e-4, job
