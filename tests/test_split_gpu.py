"""A build over pieces of triangles on the GPU (compile_scene(split_depth=),
no reference counterpart): the probe through the split tree finds the
reference tree's distance bits on every pixel and its polygon except at
exact ties, at several traversal layouts; the device-built split tree equals
the host-built one; and a render of a split scene equals the render of the
whole-triangle scene."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def test_split_probe_matches_the_whole_tree(device):
    from luce_render import render, scenes
    gset = scenes.cornell_render_set(random_count=20000)
    whole = render.compile_scene(gset)
    split = render.compile_scene(gset, split_depth=4)
    on_device = render.compile_scene(gset, split_depth=4, device=device)
    assert np.array_equal(on_device.order, split.order)
    assert np.array_equal(on_device.nodes["first"], split.nodes["first"])
    assert np.array_equal(on_device.nodes["count"], split.nodes["count"])
    assert np.allclose(on_device.nodes["lo"], split.nodes["lo"], atol=0) and np.allclose(on_device.nodes["hi"], split.nodes["hi"], atol=0)
    say(f"split build: {whole.triangle_count()} triangles -> {split.triangle_count()} pieces")
    W = H = 128
    frame = render.ray_frame(whole.cameras[0], W, H)
    gw = render.GpuScene.upload(device, whole)
    gs = render.GpuScene.upload(device, split)
    try:
        fw, tw = render.probe_arrays(device, gw, frame)
        f64 = O.ray_frame(whole.cameras[0].matrix, False, 36.0, 30.0, W, H)
        bf, bd, bm = O.brute_force(whole.positions, whole.indices, whole.faces, frame=f64, width=W, height=H,
                                   threads=8)
        assert (bf >= 0).sum() > W * H // 2
        for width in (4, 20, 8):
            gs.set_trace_config(width, 64, 16, 32)
            say(f"split probe layout {width}")
            fs, ts = render.probe_arrays(device, gs, frame)
            assert np.array_equal(ts.view(np.uint32), tw.view(np.uint32)), width
            differ = np.nonzero(fs != fw)[0]
            assert len(differ) <= 16 and np.all(bm[differ] <= 1e-6 * bd[differ]), (width, differ)
            # against the f64 brute force: same polygon but at ties, 1e-5 relative distance
            hit = bf >= 0
            got = np.where(fs == 0xFFFFFFFF, -1, fs.astype(np.int64))
            assert np.all((got == bf) | (bm <= 1e-6 * bd))
            rel = np.abs(ts[hit].astype(np.float64) - bd[hit]) / bd[hit]
            assert rel.max() <= 1e-5
    finally:
        gw.destroy()
        gs.destroy()


def test_split_scene_renders_the_same_image(device):
    from luce_render import render, scenes
    from luce_render.loading import RenderSettings
    gset = scenes.cornell_render_set(random_count=20000)
    st = RenderSettings(64, 64, 8, max_bounces=6, spectral=False, light_tracing=0.0)
    say("render whole and split soup")
    a = render.render(device, render.compile_scene(gset), st)
    b = render.render(device, render.compile_scene(gset, split_depth=3), st)
    same = (a == b).all(axis=2).mean()
    say(f"  pixels bit-identical: {same:.4f}; means {a[..., :3].mean():.6f} {b[..., :3].mean():.6f}")
    assert same >= 0.995
    assert abs(a[..., :3].mean() - b[..., :3].mean()) <= 1e-4 * a[..., :3].mean()
