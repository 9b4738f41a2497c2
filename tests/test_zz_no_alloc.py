"""The allocation-free closest-hit entry point (lr_trace_closest_device_noalloc)
against lr_trace_closest_device, through the C ABI: the same polygon and
distance bits on two calls in a row, the second with fewer rays, for several
traversal variants (the spill area and counters live in the library's scratch
array, which keeps the last call's contents); rays past n are left alone; too
many rays are refused before anything runs."""
import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _rays(n, seed):
    rng = np.random.default_rng(seed)
    o = rng.uniform(-0.9, 0.9, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3], rays[:, 4:7] = o, d
    rays[:, 3] = np.where(np.arange(n) % 3 == 0, 3.0e38, 0.35)
    return rays


def test_noalloc_equals_the_allocating_entry_point(device):
    from luce_render import render, scenes
    sc = render.compile_scene(scenes.cornell_probe_set())
    gpu = render.GpuScene.upload(device, sc)
    big, small = 20000, 3001
    rays = _rays(big, 31)
    rb = render.Buffer.create(device, rays)
    ha = render.Buffer(device, big * 8)
    hb = render.Buffer(device, big * 8)
    try:
        bounds = gpu.info()["stack_bound"]
        say(f"no-alloc: stack bounds {bounds}")
        assert max(bounds) > 16  # rays spill into the scratch array
        mirror = {}
        for tm in (3.0e38, 0.35):
            sel = rays[:, 3] == np.float32(tm)
            f_, t_, _ = O.trace_rays(sc.nodes, sc.positions, sc.indices, sc.faces,
                                     np.concatenate([rays[sel, 0:3], rays[sel, 4:7]], 1), tm)
            mirror[tm] = (sel, f_, t_)
        mf = np.zeros(big, np.uint32)
        mt = np.zeros(big, np.float32)
        for sel, f_, t_ in mirror.values():
            mf[sel], mt[sel] = f_, t_
        for cfg in ((20, 64, 16, 32), (2, 256, 16, 0), (24, 128, 24, 16), (4, 64, 32, 48)):
            gpu.set_trace_config(*cfg)
            for n in (big, small):
                say(f"no-alloc: config {cfg} n {n}")
                ha.fill(0xAB)
                hb.fill(0xAB)
                render.trace_closest_device(gpu, rb, n, ha)
                render.trace_closest_device_noalloc(gpu, rb, n, hb)
                render.synchronize(device)
                a = ha.read(np.uint32).reshape(big, 2)
                b = hb.read(np.uint32).reshape(big, 2)
                assert np.array_equal(a[:n], b[:n]), (cfg, n)
                assert np.array_equal(b[:n, 0], mf[:n]) and np.array_equal(b[:n, 1], mt[:n].view(np.uint32))
                assert (b[n:] == 0xABABABAB).all()
        with pytest.raises(render.LuceError) as e:
            render._lib.check(render._lib.lib().lr_trace_closest_device_noalloc(
                gpu.handle, rb.ptr, 1 << 31, hb.ptr, render.device_stream(device)))
        assert e.value.code == 401
    finally:
        for b in (rb, ha, hb):
            b.destroy()
        gpu.destroy()
