"""lr_bvh_build_device: the GPU build of build_bvh (bvh.lucb:66-217) gives the
reference's tree node for node (first/count equal, boxes equal as values) and
its leaf order: against the oracle's sequential restatement on triangle soups,
degenerate sets (one point, coincident triangles that force halving, a line)
and the Cornell probe scene; against the host builder on the instanced scene."""
import time

import numpy as np
import pytest

from oracle import oracle as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _soup(n, seed, spread=1.0):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-spread, spread, (n, 1, 3))
    p = (c + rng.normal(0, 0.05, (n, 3, 3))).reshape(-1, 3).astype(np.float32)
    return p, np.arange(3 * n, dtype=np.uint32).reshape(n, 3)


def _scene(p, idx):
    from luce_render import render
    n = len(idx)
    sc = render.CompiledScene()
    sc.positions, sc.indices = np.ascontiguousarray(p.reshape(-1, 3)), np.ascontiguousarray(idx)
    sc.faces = np.arange(n, dtype=np.uint32)
    sc.shading = np.zeros((n, 4), np.uint32)
    sc.uvs = np.zeros((n, 6), np.float32)
    return sc


def _same_tree(got, want):
    assert got.shape == want.shape
    assert np.array_equal(got["first"], want["first"])
    assert np.array_equal(got["count"], want["count"])
    assert np.array_equal(got["lo"], want["lo"]) and np.array_equal(got["hi"], want["hi"])


def _device_tree(device, p, idx):
    from luce_render import render
    sc = _scene(p, idx)
    render.build_bvh(sc, device=device)
    return sc.nodes, sc.order


@pytest.mark.parametrize("n", [0, 1, 3, 4, 5, 16, 17, 100, 1000, 20000, 300000])
def test_device_bvh_equals_oracle_on_soups(device, n):
    p, idx = _soup(n, n + 11)
    say(f"device build: soup of {n}")
    nodes, order = _device_tree(device, p, idx)
    want, want_order, _ = O.bvh_build(p, idx)
    _same_tree(nodes, want)
    assert np.array_equal(order, want_order[:n])


def _degenerate(kind, n):
    rng = np.random.default_rng(5)
    if kind == "point":  # every corner the same point: no split has area, halving throughout
        p = np.zeros((3 * n, 3), np.float32) + np.float32(0.25)
    elif kind == "coincident":  # groups of identical triangles: bins cannot separate a group
        base = rng.uniform(-1, 1, (n // 40 + 1, 3, 3)).astype(np.float32)
        p = base[np.arange(n) % len(base)].reshape(-1, 3)
    elif kind == "line":  # centroids on one axis, zero extent on the others
        x = rng.uniform(-1, 1, n).astype(np.float32)
        p = np.stack([x, x + 0.01, x + 0.02], 1)[:, :, None] * np.array([1, 0, 0], np.float32)
        p = p.reshape(-1, 3).astype(np.float32)
    else:  # signed zeros: planes through the origin with both signs of zero
        c = rng.integers(-3, 4, (n, 3, 3)).astype(np.float32) * np.float32(0.5)
        c[rng.random((n, 3, 3)) < 0.3] = -0.0
        p = c.reshape(-1, 3)
    return p, np.arange(3 * n, dtype=np.uint32).reshape(n, 3)


@pytest.mark.parametrize("kind", ["point", "coincident", "line", "zeros"])
def test_device_bvh_equals_oracle_on_degenerate_sets(device, kind):
    p, idx = _degenerate(kind, 5000)
    say(f"device build: degenerate {kind}")
    nodes, order = _device_tree(device, p, idx)
    want, want_order, _ = O.bvh_build(p, idx)
    _same_tree(nodes, want)
    assert np.array_equal(order, want_order)


def test_device_bvh_equals_oracle_on_cornell_probe(device, cornell_probe):
    sc = cornell_probe
    inv = np.empty_like(sc.order)
    inv[sc.order] = np.arange(len(sc.order), dtype=np.uint32)
    original = sc.indices[inv]
    nodes, order = _device_tree(device, sc.positions, original)
    _same_tree(nodes, sc.nodes)
    assert np.array_equal(order, sc.order)


def test_device_bvh_equals_host_build_on_instanced_scene(device):
    from luce_render import render, scenes
    sc = render.compile_scene(scenes.instanced_set(instances=60), geometry=False)
    p, idx = sc.positions, sc.indices
    say(f"device build: instanced, {len(idx)} triangles")
    t = time.perf_counter()
    host = _scene(p, idx)
    render.build_bvh(host)
    t_host = time.perf_counter() - t
    t = time.perf_counter()
    nodes, order = _device_tree(device, p, idx)
    t_dev = time.perf_counter() - t
    say(f"host {t_host:.3f} s, device {t_dev:.3f} s (host arrays in and out)")
    _same_tree(nodes, host.nodes)
    assert np.array_equal(order, host.order)
    assert O.check_bvh(nodes, p, idx[order]) == 0


def _host_layout(nodes, width):
    from luce_render import _lib
    import ctypes
    size, bound = ctypes.c_size_t(0), ctypes.c_uint32(0)
    nodes = np.ascontiguousarray(nodes)
    _lib.check(_lib.lib().lr_bvh_collapse(_lib.ptr(nodes), len(nodes), width, None, 0, ctypes.byref(size),
                                          ctypes.byref(bound)))
    out = np.zeros(size.value, np.uint8)
    _lib.check(_lib.lib().lr_bvh_collapse(_lib.ptr(nodes), len(nodes), width, _lib.ptr(out), size.value,
                                          ctypes.byref(size), ctypes.byref(bound)))
    return out, bound.value


@pytest.mark.parametrize("which", ["empty", "one", "soup", "degenerate", "cornell", "instanced"])
def test_device_collapse_equals_host_collapse(device, cornell_probe, which):
    """Upload collapses the tree on the device into the host collapse's bytes
    (bvh_wide.cpp) at widths 2, 4, 8, with the same stack bounds."""
    from luce_render import render, scenes
    if which == "cornell":
        sc = cornell_probe
    else:
        if which == "instanced":
            flat = render.compile_scene(scenes.instanced_set(instances=20), geometry=False)
            p, idx = flat.positions, flat.indices
        elif which == "degenerate":
            p, idx = _degenerate("coincident", 3000)
        else:
            p, idx = _soup({"empty": 0, "one": 1, "soup": 50000}[which], 9)
        sc = _scene(p, idx)
        render.build_bvh(sc)
    say(f"device collapse: {which}, {sc.triangle_count()} triangles")
    g = render.GpuScene.upload(device, sc)
    info = g.info()
    for wi, width in enumerate((2, 4, 8)):
        want, bound = _host_layout(sc.nodes, width)
        got = g.layout(width)
        assert np.array_equal(got, want), width
        assert info["stack_bound"][wi] == bound
    g.destroy()


def test_scene_build_on_device_equals_compile_then_upload(device, cornell_probe):
    """lr_scene_build (device BVH build, reorder, collapse) gives the scene
    that compile_scene + upload give: same tree and order, same layouts, the
    same probe image bit for bit."""
    from luce_render import render, scenes
    gset = scenes.cornell_probe_set(random_count=10000)
    ref = render.compile_scene(gset)
    flat = render.compile_scene(gset, geometry=False)
    g = render.GpuScene.build(device, flat)
    _same_tree(flat.nodes, ref.nodes)
    assert np.array_equal(flat.order, ref.order)
    for a in ("indices", "faces", "shading", "uvs"):
        assert np.array_equal(getattr(flat, a), getattr(ref, a)), a
    gr = render.GpuScene.upload(device, ref)
    for width in (2, 4, 8):
        assert np.array_equal(g.layout(width), gr.layout(width)), width
    frame = render.ray_frame(ref.cameras[0], 256, 256)
    f1, t1 = render.probe_arrays(device, g, frame)
    f2, t2 = render.probe_arrays(device, gr, frame)
    assert np.array_equal(f1, f2) and np.array_equal(t1.view(np.uint32), t2.view(np.uint32))
    mf, mt, _, _ = O.probe(ref.nodes, ref.positions, ref.indices, ref.faces, frame.pack(), 256, 256)
    assert np.array_equal(f1, mf) and np.array_equal(t1.view(np.uint32), mt.view(np.uint32))
    g.destroy()
    gr.destroy()


def _refit_tree(nodes, positions, indices):
    """The refit tree on the CPU: a leaf's box grows over its triangles'
    corners, an interior node's is the union of its children's (children
    follow their parent in the reference's numbering: one backward sweep)."""
    out = nodes.copy()
    tri = positions[indices]
    tlo, thi = tri.min(1), tri.max(1)
    for i in range(len(out) - 1, -1, -1):
        c, f = int(out["count"][i]), int(out["first"][i])
        if c:
            out["lo"][i], out["hi"][i] = tlo[f:f + c].min(0), thi[f:f + c].max(0)
        else:
            out["lo"][i] = np.minimum(out["lo"][f], out["lo"][f + 1])
            out["hi"][i] = np.maximum(out["hi"][f], out["hi"][f + 1])
    return out


def test_refit_moves_boxes_and_keeps_the_tree(device):
    """lr_scene_refit: unchanged points leave every layout byte for byte; moved
    points give the width-2 layout of the CPU-refit tree byte for byte, probe
    images at every width equal to the oracle's mirror over the refit tree,
    and moving back restores the original layouts."""
    from luce_render import render, scenes
    sc = render.compile_scene(scenes.cornell_probe_set(random_count=10000))
    g = render.GpuScene.upload(device, sc)
    before = [g.layout(w) for w in (2, 4, 8)]
    g.refit(sc.positions)
    for w, b in zip((2, 4, 8), before):
        assert np.array_equal(g.layout(w), b), w
    rng = np.random.default_rng(21)
    moved = (sc.positions + rng.normal(0, 0.01, sc.positions.shape)).astype(np.float32)
    moved[:, 1] += np.float32(0.05)
    g.refit(moved)
    tree = _refit_tree(sc.nodes, moved, sc.indices)
    want, _ = _host_layout(tree, 2)
    assert np.array_equal(g.layout(2), want)
    frame = render.ray_frame(sc.cameras[0], 128, 128)
    mf, mt, _, _ = O.probe(tree, moved, sc.indices, sc.faces, frame.pack(), 128, 128)
    assert (mf != 0xFFFFFFFF).sum() > 128 * 128 // 2
    for w in (2, 4, 8):
        g.set_trace_config(w, 64, 24, 0)
        f, t = render.probe_arrays(device, g, frame)
        assert np.array_equal(f, mf), w
        assert np.array_equal(t.view(np.uint32), mt.view(np.uint32)), w
    g.refit(sc.positions)
    for w, b in zip((2, 4, 8), before):
        assert np.array_equal(g.layout(w), b), w
    with pytest.raises(render.LuceError) as e:
        g.refit(sc.positions[:-1])
    assert e.value.code == 400
    g.destroy()


def test_refit_of_a_device_built_scene(device):
    """A scene built on the device refits like an uploaded one."""
    from luce_render import render, scenes
    gset = scenes.cornell_probe_set(random_count=2000)
    ref = render.compile_scene(gset)
    flat = render.compile_scene(gset, geometry=False)
    g = render.GpuScene.build(device, flat, keep_tree=False)
    moved = (ref.positions * np.float32(0.9)).astype(np.float32)
    g.refit(moved)
    tree = _refit_tree(ref.nodes, moved, ref.indices)
    want, _ = _host_layout(tree, 2)
    assert np.array_equal(g.layout(2), want)
    g.destroy()


def _decode_quantized(raw, width):
    """Child boxes (nodes, width, 2, 3) of a quantized layout, decoded in f32
    as the traversal does, with the child words."""
    node = 64 if width == 4 else 128
    q = raw.reshape(-1, node)
    words = q.view(np.uint32)
    origin = words[:, 0:3].view(np.float32)
    e = words[:, 3]
    step = np.stack([((e >> (8 * a)) & 0xFF) << 23 for a in range(3)], 1).astype(np.uint32).view(np.float32)
    planes = q[:, 16:16 + 6 * width].reshape(-1, 6, width).astype(np.float32)
    lo = origin[:, None, :] + np.stack([planes[:, 0], planes[:, 2], planes[:, 4]], 2) * step[:, None, :]
    hi = origin[:, None, :] + np.stack([planes[:, 1], planes[:, 3], planes[:, 5]], 2) * step[:, None, :]
    child = words[:, 4 + 6 * width // 4:4 + 6 * width // 4 + width]
    return lo.astype(np.float32), hi.astype(np.float32), child


def _decode_float(raw, width):
    node = 128 if width == 4 else 256
    f = raw.reshape(-1, node).view(np.float32)
    u = raw.reshape(-1, node).view(np.uint32)
    lo = np.stack([f[:, 0:width], f[:, 2 * width:3 * width], f[:, 4 * width:5 * width]], 2)
    hi = np.stack([f[:, width:2 * width], f[:, 3 * width:4 * width], f[:, 5 * width:6 * width]], 2)
    return lo, hi, u[:, 6 * width:7 * width]


def _check_quantized(g, tight=True):
    for width, code in ((4, 20), (8, 24)):
        flo, fhi, fch = _decode_float(g.layout(width), width)
        qlo, qhi, qch = _decode_quantized(g.layout(code), width)
        assert np.array_equal(fch, qch)
        valid = fch != 0xFFFFFFFF
        assert np.all(qlo[valid] <= flo[valid]) and np.all(qhi[valid] >= fhi[valid])
        if tight:  # the grid is at most 255 steps of the node's extent: boxes grow by a step at most per side
            ext = np.maximum(fhi.max(1) - flo.min(1), 1e-30)[:, None, :]
            grow = np.maximum(flo - qlo, qhi - fhi)
            assert np.all((grow / ext)[valid] <= 2.0 / 255.0 + 1e-6)


def test_quantized_layouts_hold_the_float_boxes(device, cornell_probe):
    """Layouts 20 and 24 (widths 4 and 8 on an 8-bit grid): the same children,
    every decoded child box holding the float layout's and at most a grid step
    larger on each side; after a refit too."""
    from luce_render import render, scenes
    g = render.GpuScene.upload(device, cornell_probe)
    _check_quantized(g)
    moved = (cornell_probe.positions * np.float32(1.3) + np.float32(0.2)).astype(np.float32)
    g.refit(moved)
    _check_quantized(g)
    g.destroy()
    flat = render.compile_scene(scenes.instanced_set(instances=20), geometry=False)
    g = render.GpuScene.build(device, flat, keep_tree=False)
    _check_quantized(g)
    g.destroy()
