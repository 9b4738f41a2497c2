"""Host-side checks that need no GPU: the library exports every declared
entry point, the native BVH builder reproduces the oracle's tree, scene
compilation follows the reference's rules, and errors carry its codes."""
import hashlib
import json
import re
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle as O

ROOT = Path(__file__).resolve().parent.parent


def _declared():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        names |= set(re.findall(r"^\s*(?:[\w\s\*]+?)\b(lr_\w+)\s*\(", h.read_text(), re.M))
    return names


def test_library_exports_every_declared_symbol():
    from luce_render import _lib
    lib = _lib.lib()
    declared = _declared()
    assert len(declared) >= 20
    for name in declared:
        assert hasattr(lib, name), name
    assert declared == set(_lib.exported_names()) | (declared - set(_lib.exported_names()))
    assert set(_lib.exported_names()) <= declared


def test_device_open_without_gpu_reports_unavailable():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from luce_render import render
    with pytest.raises(render.GpuUnavailable) as e:
        render.Device.open()
    assert e.value.code == 503


def _soup(n, seed, spread=1.0):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-spread, spread, (n, 1, 3))
    p = (c + rng.normal(0, 0.05, (n, 3, 3))).reshape(-1, 3).astype(np.float32)
    return p, np.arange(3 * n, dtype=np.uint32).reshape(n, 3)


@pytest.mark.parametrize("n", [0, 1, 4, 5, 17, 1000, 20000])
@pytest.mark.parametrize("threads", [1, 3, 8])
def test_native_bvh_equals_oracle(n, threads):
    """lr_bvh_build (parallel) gives the reference's tree and leaf order."""
    from luce_render import render
    p, idx = _soup(n, n + 11)
    nodes, order, _ = O.bvh_build(p, idx)
    sc = render.CompiledScene()
    sc.positions, sc.indices = p.reshape(-1, 3), idx
    sc.faces = np.arange(n, dtype=np.uint32)
    sc.shading = np.zeros((n, 4), np.uint32)
    sc.uvs = np.zeros((n, 6), np.float32)
    render.build_bvh(sc, threads=threads)
    assert np.array_equal(sc.nodes.view(np.uint8), nodes.view(np.uint8))
    assert np.array_equal(sc.order, order[:n])


def test_native_bvh_equals_oracle_on_cornell_probe(cornell_probe):
    sc = cornell_probe
    # Rebuild the oracle tree from the pre-BVH triangle order.
    inv = np.empty_like(sc.order)
    inv[sc.order] = np.arange(len(sc.order), dtype=np.uint32)
    original = sc.indices[inv]
    nodes, order, _ = O.bvh_build(sc.positions, original)
    assert np.array_equal(sc.nodes.view(np.uint8), nodes.view(np.uint8))
    assert np.array_equal(sc.order, order)
    assert O.check_bvh(sc.nodes, sc.positions, sc.indices) == 0


def test_cube_compiles_to_12_triangles_of_6_faces():
    """tests.lucb:30-41."""
    from luce_render import render, scenes
    cube = scenes.box_mesh((2.0, 2.0, 2.0))
    # the box mesh keeps 4 corners per face; weld them as Mesh.cube does
    pts, inverse = np.unique(cube.points, axis=0, return_inverse=True)
    welded = render.Mesh(pts, inverse.reshape(-1)[cube.triangles], cube.triangle_faces, 6)
    sc = render.compile_scene(render.GeometrySet(meshes=[welded]))
    assert sc.triangle_count() == 12 and sc.point_count() == 8
    assert np.bincount(sc.faces, minlength=6).tolist() == [2] * 6
    assert O.check_bvh(sc.nodes, sc.positions, sc.indices) == 0


def _grid(size, n):
    xs = np.linspace(-size / 2, size / 2, n + 1)
    px, pz = np.meshgrid(xs, xs, indexing="ij")
    pts = np.stack([px.ravel(), np.zeros(px.size), pz.ravel()], 1)
    tris, faces = [], []
    for i in range(n):
        for j in range(n):
            a, b, c, d = i * (n + 1) + j, (i + 1) * (n + 1) + j, (i + 1) * (n + 1) + j + 1, i * (n + 1) + j + 1
            tris += [[a, b, c], [a, c, d]]
            faces += [i * n + j] * 2
    from luce_render import render
    return render.Mesh(pts, np.array(tris), np.array(faces), n * n)


def test_instances_are_placed_and_split_into_leaves_of_at_most_16():
    """tests.lucb:43-61: a 40x40 grid plus an instance of it 5 units up."""
    from luce_render import render, scenes
    grid = _grid(10.0, 40)
    plane = render.GeometrySet(meshes=[grid])
    base = render.GeometrySet(meshes=[grid], instances=[render.Instance(plane, scenes.placement(translation=(0, 5, 0)))])
    sc = render.compile_scene(base)
    assert sc.triangle_count() == 2 * 2 * 40 * 40
    assert sc.positions[:, 1].max() == 5.0
    assert (sc.nodes["count"] <= 16).all() and len(sc.nodes) > 100
    assert O.check_bvh(sc.nodes, sc.positions, sc.indices) == 0
    # polygon ids of the instance follow the base mesh's
    assert sc.faces.max() == 2 * 40 * 40 - 1


def test_octahedral_round_trip():
    """tests.lucb:134-144: encode then the kernels' decode within 1e-3."""
    from luce_render import render
    for d in ([0.3, -0.5, 0.8], [-1.0, 0.0, 0.0], [0.0, 0.0, -1.0], [0.6, 0.6, -0.5]):
        e = np.asarray(d) / np.linalg.norm(d)
        back = render.decode_octahedral(render.octahedral(e))[0]
        assert np.linalg.norm(back - e) < 1e-3


def test_compile_stops_when_asked():
    """scene.lucb:143-161: a stopped compile fails with `stopped` (402)."""
    from luce_render import render, scenes
    stop = render.Stopping()
    stop.ask()
    with pytest.raises(render.LuceError) as e:
        render.compile_scene(scenes.cornell_probe_set(random_count=10), stopping=stop)
    assert e.value.code == 402


def test_invalid_inputs_carry_the_reference_codes():
    from luce_render import render, _lib
    with pytest.raises(render.LuceError) as e:
        render.ray_frame(render.Camera(np.eye(4)), 0, 10)
    assert e.value.code == 400
    # a BVH leaf naming triangles past the array is refused before any device work
    sc = render.CompiledScene()
    sc.positions = np.zeros((3, 3), np.float32)
    sc.indices = np.array([[0, 1, 2]], np.uint32)
    sc.faces = np.zeros(1, np.uint32)
    rc = _lib.lib().lr_bvh_build(None, 3, None, 1, 1, None, None, None)
    assert rc == 400 and "needs" in _lib.last_error()
    frame = np.zeros(16, np.float32)
    rc = _lib.lib().lr_probe(None, None, _lib.ptr(frame), 4, 4, None, None, 15)
    assert rc == 400


def test_cornell_probe_scene_is_the_committed_one(cornell_probe):
    """The generated workload's arrays hash to the committed fixture."""
    golden = json.loads((ROOT / "tests" / "golden" / "cornell_probe.json").read_text())
    sc = cornell_probe
    assert sc.triangle_count() == golden["triangles"] == 10034
    assert int(sc.faces.max()) + 1 == golden["polygons"] == 10017
    assert hashlib.sha256(np.ascontiguousarray(sc.positions).tobytes()).hexdigest() == golden["positions_sha256"]
    assert hashlib.sha256(np.ascontiguousarray(sc.nodes).tobytes()).hexdigest() == golden["nodes_sha256"]


def _swap_partition(items, left):
    """partition (bvh.lucb:185-196): the reference's in-place swap loop."""
    a = list(items)
    low, high = 0, len(a)
    while low < high:
        if left(a[low]):
            low += 1
        else:
            high -= 1
            a[low], a[high] = a[high], a[low]
    return a, low


def _closed_form_partition(items, left):
    """The device builder's closed form of the same permutation: each element's
    destination from its rank among the lefts before it and two tables."""
    n = len(items)
    f = [left(x) for x in items]
    L = sum(f)
    lb = np.concatenate([[0], np.cumsum(f)]).astype(int)
    S, T = {}, {}
    for i in range(n):
        if i < L and not f[i]:
            S[i - lb[i] + 1] = i
        if i >= L and f[i]:
            T[L - lb[i]] = i
    out = [None] * n
    for i in range(n):
        if f[i]:
            d = i if i < L else S[L - lb[i]]
        elif i < L:
            k = i - lb[i] + 1
            d = (n if k == 1 else T[k - 1]) - 1
        elif i > L:
            d = i - 1
        else:
            m = L - lb[i]
            d = (n if m == 0 else T[m]) - 1
        assert out[d] is None
        out[d] = items[i]
    return out, L


def test_closed_form_partition_equals_the_swap_loop():
    rng = np.random.default_rng(4)
    for _ in range(20000):
        n = int(rng.integers(0, 14))
        flags = rng.random(n) < rng.random()
        items = list(range(n))
        assert _closed_form_partition(items, lambda x: flags[x]) == _swap_partition(items, lambda x: flags[x])
