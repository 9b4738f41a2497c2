"""GPU parity of the tracing kernels against the oracle: polygon ids and
distance bits identical to the f32 mirror of bvh.glsl, and agreement with the
f64 brute force under the reference's margin rule (tests/gpu/main.lucb:61-78)."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu

WIDTHS = (2, 4, 8, 20, 24)  # 20, 24: the quantized 4- and 8-wide layouts
REFILLS = (0, 1, 16, 64)  # 0: one ray per thread; k: persistent waves refilled at k idle lanes


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def probe_scene(device, cornell_probe):
    from luce_render import render
    g = render.GpuScene.upload(device, cornell_probe)
    yield g
    g.destroy()


def _mirror(sc, frame):
    return O.probe(sc.nodes, sc.positions, sc.indices, sc.faces, frame.pack(), frame.width, frame.height)


@pytest.mark.parametrize("width", WIDTHS)
def test_probe_bit_exact_against_mirror(device, cornell_probe, probe_scene, width):
    from luce_render import render
    sc = cornell_probe
    frame = render.ray_frame(sc.cameras[0], 256, 256)
    probe_scene.set_trace_config(width, 128, 32)
    say(f"probe 256x256 width {width}")
    faces, t = render.probe_arrays(device, probe_scene, frame)
    mf, mt, _, _ = _mirror(sc, frame)
    assert np.array_equal(faces, mf), int((faces != mf).sum())
    assert np.array_equal(t.view(np.uint32), mt.view(np.uint32))


def test_probe_agrees_with_brute_force(device, cornell_probe, probe_scene):
    from luce_render import render
    sc = cornell_probe
    frame = render.ray_frame(sc.cameras[0], 256, 256)
    probe_scene.set_trace_config(4, 128, 32)
    faces, t = render.probe_arrays(device, probe_scene, frame)
    f64 = O.ray_frame(sc.cameras[0].matrix, False, 36.0, 30.0, 256, 256)
    bf, bd, bm = O.brute_force(sc.positions, sc.indices, sc.faces, frame=f64, width=256, height=256, threads=8)
    same, differ = O.agrees(faces, t, bf, bd, bm)
    assert (bf >= 0).sum() > 256 * 256 // 2
    assert differ.sum() == 0
    # Hit polygons equal the brute force's on every pixel and distances lie
    # within 1e-5 relative of its f64 distances.
    hit = bf >= 0
    assert np.array_equal(np.where(faces == 0xFFFFFFFF, -1, faces.astype(np.int64)), bf)
    rel = np.abs(t[hit].astype(np.float64) - bd[hit]) / bd[hit]
    say(f"probe vs brute force: max relative distance difference {rel.max():.3e}")
    assert rel.max() <= 1e-5


def test_probe_orthographic(device, cornell_probe, probe_scene):
    from luce_render import render
    sc = cornell_probe
    frame = render.look_frame((0.031, 0.017, 4.0), (0.05, -0.1, -1.0), (0.0, 1.0, 0.0), 1.3, 96, 64, orthographic=True)
    f64 = O.look_frame((0.031, 0.017, 4.0), (0.05, -0.1, -1.0), (0.0, 1.0, 0.0), 1.3, 96, 64, True)
    mf, mt, _, _ = _mirror(sc, frame)
    bf, bd, bm = O.brute_force(sc.positions, sc.indices, sc.faces, frame=f64, width=96, height=64, threads=8)
    for width in WIDTHS:
        probe_scene.set_trace_config(width, 64, 24)
        say(f"probe orthographic width {width}")
        faces, t = render.probe_arrays(device, probe_scene, frame)
        assert np.array_equal(faces, mf) and np.array_equal(t.view(np.uint32), mt.view(np.uint32))
        _, differ = O.agrees(faces, t, bf, bd, bm)
        assert differ.sum() == 0


def test_every_variant_agrees(device, cornell_probe, probe_scene):
    from luce_render import render
    sc = cornell_probe
    frame = render.ray_frame(sc.cameras[0], 100, 75)
    mf, mt, _, _ = _mirror(sc, frame)
    for width in WIDTHS:
        for group in (64, 128, 256):
            for stack in (16, 24, 32, 48):
                probe_scene.set_trace_config(width, group, stack)
                say(f"probe 100x75 variant w{width} g{group} s{stack}")
                faces, t = render.probe_arrays(device, probe_scene, frame)
                assert np.array_equal(faces, mf) and np.array_equal(t.view(np.uint32), mt.view(np.uint32)), \
                    (width, group, stack)
    probe_scene.set_trace_config(4, 128, 32)


def _random_rays(n, seed):
    rng = np.random.default_rng(seed)
    o = rng.uniform(-0.9, 0.9, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    # a few axis-parallel directions exercise the 1e-20 guard
    d[:8] = [[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1], [0, 0.6, 0.8], [-0.0, 0, -1]]
    return o.astype(np.float32), d.astype(np.float32)


def test_ray_batches_closest_and_any(device, cornell_probe, probe_scene):
    from luce_render import render
    sc = cornell_probe
    n = 20000
    o, d = _random_rays(n, 9)
    tmax = np.where(np.arange(n) % 3 == 0, 3.0e38, 0.35).astype(np.float32)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3], rays[:, 3], rays[:, 4:7] = o, tmax, d
    rb = render.Buffer.create(device, rays)
    hb = render.Buffer(device, n * 8)
    ob = render.Buffer(device, n * 4)
    rays6 = np.concatenate([o, d], 1)
    # the mirror's closest hit under each ray's own t_max
    mf = np.zeros(n, np.uint32)
    mt = np.zeros(n, np.float32)
    for tm in (3.0e38, 0.35):
        sel = tmax == np.float32(tm)
        f_, t_, _ = O.trace_rays(sc.nodes, sc.positions, sc.indices, sc.faces, rays6[sel], tm)
        mf[sel], mt[sel] = f_, t_
    for width in WIDTHS:
        for refill in REFILLS:
            probe_scene.set_trace_config(width, 128, 32, refill)
            say(f"ray batch closest/any width {width} refill {refill}")
            hb.fill(0xAB)
            ob.fill(0xAB)
            render.trace_closest_device(probe_scene, rb, n, hb)
            render.trace_any_device(probe_scene, rb, n, ob)
            render.synchronize(device)
            h = hb.read(np.uint32).reshape(n, 2)
            assert np.array_equal(h[:, 0], mf)
            assert np.array_equal(h[:, 1], mt.view(np.uint32))
            occluded = ob.read(np.uint32)
            # any-hit visits every leaf whose box passes at fixed t_max, so it
            # is occluded exactly when some tested triangle lies before t_max
            assert np.array_equal(occluded, (mf != 0xFFFFFFFF).astype(np.uint32))
    for b in (rb, hb, ob):
        b.destroy()
    probe_scene.set_trace_config(4, 128, 32)


def test_persistent_launches_back_to_back(device, cornell_probe, probe_scene):
    """The persistent launches share one ray counter that the last wave of
    each launch clears: launches in a row, of falling and rising sizes (down
    to a single ray, fewer than a wave), each cover every ray exactly once."""
    from luce_render import render
    sc = cornell_probe
    o, d = _random_rays(5000, 21)
    rays = np.zeros((5000, 8), np.float32)
    rays[:, 0:3], rays[:, 3], rays[:, 4:7] = o, 3.0e38, d
    mf, mt, _ = O.trace_rays(sc.nodes, sc.positions, sc.indices, sc.faces, np.concatenate([o, d], 1))
    rb = render.Buffer.create(device, rays)
    hb = render.Buffer(device, 5000 * 8)
    probe_scene.set_trace_config(4, 64, 24, 16)
    for n in (5000, 1, 37, 5000, 64, 4999):
        say(f"persistent closest batch of {n}")
        hb.fill(0xAB)
        render.trace_closest_device(probe_scene, rb, n, hb)
        render.synchronize(device)
        h = hb.read(np.uint32).reshape(5000, 2)
        assert np.array_equal(h[:n, 0], mf[:n]) and np.array_equal(h[:n, 1], mt[:n].view(np.uint32)), n
        assert (h[n:] == 0xABABABAB).all(), n
    rb.destroy()
    hb.destroy()
    probe_scene.set_trace_config(4, 128, 32)


@pytest.mark.parametrize("size", [(256, 256), (100, 37)])
def test_frame_and_batch_in_one_launch(device, cornell_probe, probe_scene, size):
    """lr_trace_frame_and_closest_device: a frame's primary rays and a ray
    batch from one counter (refilled persistent waves) give the probe's and
    the batch call's hits bit for bit, for every layout and refill (0: the
    two launches), a frame whose tiles overhang its edges, and batches down
    to none; nothing outside the outputs is written."""
    from luce_render import render
    sc = cornell_probe
    w, h = size
    frame = render.ray_frame(sc.cameras[0], w, h)
    mf, mt, _, _ = _mirror(sc, frame)
    n = 7000
    o, d = _random_rays(n, 33)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3], rays[:, 3], rays[:, 4:7] = o, 3.0e38, d
    bf, bt, _ = O.trace_rays(sc.nodes, sc.positions, sc.indices, sc.faces, np.concatenate([o, d], 1))
    rb = render.Buffer.create(device, rays)
    hb = render.Buffer(device, (n + 16) * 8)
    fb = render.Buffer(device, (w * h + 16) * 8)
    for width in WIDTHS:
        for refill in REFILLS:
            for count in (n, 1, 0):
                probe_scene.set_trace_config(width, 64, 16, refill)
                say(f"frame {w}x{h} and batch of {count}: width {width} refill {refill}")
                hb.fill(0xAB)
                fb.fill(0xAB)
                render.trace_frame_and_closest_device(probe_scene, frame, fb, rb, count, hb)
                render.synchronize(device)
                f = fb.read(np.uint32).reshape(-1, 2)
                assert np.array_equal(f[: w * h, 0], mf), int((f[: w * h, 0] != mf).sum())
                assert np.array_equal(f[: w * h, 1], mt.view(np.uint32))
                assert (f[w * h:] == 0xABABABAB).all()
                r = hb.read(np.uint32).reshape(-1, 2)
                assert np.array_equal(r[:count, 0], bf[:count]) and np.array_equal(r[:count, 1], bt[:count].view(np.uint32))
                assert (r[count:] == 0xABABABAB).all()
    for b in (rb, hb, fb):
        b.destroy()
    probe_scene.set_trace_config(4, 128, 32)


def test_empty_and_tiny_scenes(device):
    from luce_render import render
    empty = render.compile_scene(render.GeometrySet())
    g = render.GpuScene.upload(device, empty)
    frame = render.look_frame((0, 0, 3), (0, 0, -1), (0, 1, 0), 0.5, 16, 8)
    for refill in (0, 16):
        g.set_trace_config(4, 128, 32, refill)
        faces, t = render.probe_arrays(device, g, frame)
        assert (faces == 0xFFFFFFFF).all() and (t == np.float32(3.0e38)).all()
    g.destroy()
    one = render.compile_scene(render.GeometrySet(meshes=[render.Mesh([[-1, -1, 0], [1, -1, 0], [0, 1, 0]], [[0, 1, 2]])]))
    g = render.GpuScene.upload(device, one)
    mf, mt, _, _ = _mirror(one, frame)
    for refill in (0, 16):
        g.set_trace_config(4, 128, 32, refill)
        faces, t = render.probe_arrays(device, g, frame)
        assert np.array_equal(faces, mf) and np.array_equal(t.view(np.uint32), mt.view(np.uint32))
        assert (faces == 0).sum() > 0
    g.destroy()


def test_probe_count_must_match(device, probe_scene):
    from luce_render import render
    frame = render.look_frame((0, 0, 3), (0, 0, -1), (0, 1, 0), 0.5, 16, 8)
    with pytest.raises(render.LuceError) as e:
        render.probe(device, probe_scene, frame, out=[None] * 10)
    assert e.value.code == 400


def test_each_variant_sizes_its_own_persistent_grid(device, cornell_probe):
    """A refilled launch takes the resident workgroups of its own kernel:
    variants of different group sizes (asked in either order) get different
    grids, each at most what the rays need, and asking again gives the
    same answer."""
    from luce_render import render
    gpu = render.GpuScene.upload(device, cornell_probe)
    try:
        grids = {}
        for g in (64, 256, 128, 64):
            gpu.set_trace_config(20, g, 16, 32)
            grids.setdefault(g, []).append((gpu.resident_grid(1 << 26), gpu.resident_grid(1 << 26, any_hit=True)))
            assert gpu.resident_grid(100) == 2 if g == 64 else gpu.resident_grid(100) == 1
        say(f"resident grids {grids}")
        assert grids[64][0] == grids[64][1]
        for any_hit in (0, 1):
            sizes = {g: grids[g][0][any_hit] for g in (64, 128, 256)}
            assert len(set(sizes.values())) == 3
            threads = {g: sizes[g] * g for g in sizes}
            # no more waves than a compute unit holds (32 of 64 lanes)
            assert max(threads.values()) <= 2048 * 256
    finally:
        gpu.destroy()
