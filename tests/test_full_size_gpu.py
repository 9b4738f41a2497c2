"""Size-independent properties at the benchmark's full size (the instanced
30.4M-triangle scene, 1920x1080): every traversal width and ray schedule
and LDS stack (16, the headline configuration's, and 24) returns the same
polygon ids and distance bits on all 2,073,600 primary and bounce rays; any-hit occlusion equals "a closest hit exists before t_max";
the closest hits of a sampled subset equal the oracle's f32 mirror."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def big(device):
    from luce_render import render, scenes
    sc = render.compile_scene(scenes.instanced_set(instances=1000))
    g = render.GpuScene.upload(device, sc)
    yield sc, g
    g.destroy()


def test_widths_and_schedules_agree_at_full_size(device, big):
    import bench
    from luce_render import render
    sc, g = big
    W, H = 1920, 1080
    n = W * H
    frame = render.ray_frame(sc.cameras[0], W, H)
    hp = render.Buffer(device, n * 8)
    hb = render.Buffer(device, n * 8)
    ob = render.Buffer(device, n * 4)
    g.set_trace_config(4, 64, 24, 0)
    render.probe_device(g, frame, hp)
    render.synchronize(device)
    primary = hp.read(np.uint32).reshape(n, 2)
    rays = bench.bounce_rays(sc, frame, primary, np.random.default_rng(7))
    rays[1::3, 3] = 2.5  # some finite t_max
    rb = render.Buffer.create(device, rays)
    ref = None
    for width, stack in [(w, s) for w in (2, 4, 8, 20, 24) for s in (16, 24)]:
        for refill in (0, 32):
            g.set_trace_config(width, 64, stack, refill)
            say(f"full size: width {width} stack {stack} refill {refill}")
            render.probe_device(g, frame, hp)
            render.trace_closest_device(g, rb, n, hb)
            render.trace_any_device(g, rb, n, ob)
            render.synchronize(device)
            got = (hp.read(np.uint32), hb.read(np.uint32))
            occluded = ob.read(np.uint32)
            hits = got[1].reshape(n, 2)
            assert np.array_equal(occluded, (hits[:, 0] != 0xFFFFFFFF).astype(np.uint32))
            if ref is None:
                ref = got
            else:
                assert np.array_equal(got[0], ref[0]), (width, stack, refill)
                assert np.array_equal(got[1], ref[1]), (width, stack, refill)
    # a sample of the bounce rays against the oracle's mirror of bvh.glsl
    rng = np.random.default_rng(3)
    pick = rng.choice(n, 4096, replace=False)
    hits = ref[1].reshape(n, 2)
    for tm in (3.0e38, 2.5):
        sel = pick[rays[pick, 3] == np.float32(tm)]
        f, t, _ = O.trace_rays(sc.nodes, sc.positions, sc.indices, sc.faces, rays[sel][:, [0, 1, 2, 4, 5, 6]], tm)
        assert np.array_equal(hits[sel, 0], f)
        assert np.array_equal(hits[sel, 1], t.view(np.uint32))
    for b in (hp, hb, ob, rb):
        b.destroy()
    g.set_trace_config(20, 64, 16, 32)
