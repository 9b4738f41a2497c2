"""Builds over pieces of triangles (scene.split_bvh, lr_bvh_split_triangles;
no reference counterpart, for clouds of heavily overlapping triangles): the
pieces' boxes hold their triangles' parts and stay inside each triangle's
box, every triangle keeps at least one piece, emitters and cut-outs stay
whole, and the oracle's mirror of the reference traversal finds the same
polygon and distance bits through the split tree as through the reference's
tree, on primary rays and on rays from inside the cloud."""
import numpy as np

from oracle import oracle as O


def _soup(count, seed=1):
    from luce_render import render, scenes
    return render.compile_scene(scenes.cornell_render_set(random_count=count)), scenes


def test_pieces_cover_their_triangles():
    from luce_render import _lib
    rng = np.random.default_rng(5)
    pos = rng.uniform(-1, 1, (300, 3)).astype(np.float32)
    idx = rng.integers(0, 300, (100, 3)).astype(np.uint32)
    count = _lib.ctypes.c_size_t(0)
    for depth, extent in ((0, 0.0), (3, 0.2), (8, 0.01)):
        _lib.check(_lib.lib().lr_bvh_split_triangles(_lib.ptr(pos), 300, _lib.ptr(idx), 100, None, extent, depth, 1,
                                                     None, None, 0, _lib.ctypes.byref(count)))
        m = count.value
        boxes = np.zeros((m, 6), np.float32)
        tris = np.zeros(m, np.uint32)
        _lib.check(_lib.lib().lr_bvh_split_triangles(_lib.ptr(pos), 300, _lib.ptr(idx), 100, None, extent, depth, 1,
                                                     _lib.ptr(boxes), _lib.ptr(tris), m, _lib.ctypes.byref(count)))
        assert np.array_equal(np.unique(tris), np.arange(100)) and np.all(np.diff(tris.astype(np.int64)) >= 0)
        tri = pos[idx[tris]]
        assert np.all(boxes[:, :3] >= tri.min(1)) and np.all(boxes[:, 3:] <= tri.max(1))
        assert np.all(boxes[:, :3] <= boxes[:, 3:])
        if depth == 0:
            assert m == 100
            continue
        assert m > 100
        # points of each triangle (barycentric samples) lie in one of its pieces' boxes
        for t in range(0, 100, 7):
            mine = boxes[tris == t]
            w = rng.dirichlet((1, 1, 1), 400)
            pts = w @ pos[idx[t]].astype(np.float64)
            inside = ((pts[:, None, :] >= mine[None, :, :3]) & (pts[:, None, :] <= mine[None, :, 3:])).all(2).any(1)
            assert inside.all(), t
        # the pieces are smaller than the whole triangles' boxes
        assert (np.prod(boxes[:, 3:] - boxes[:, :3], 1).sum()
                < np.prod(pos[idx].max(1) - pos[idx].min(1), 1).sum())


def test_split_tree_finds_the_same_hits():
    from luce_render import render, scene as S
    sc, _ = _soup(20000)
    ref_nodes, ref_idx, ref_faces = sc.nodes, sc.indices, sc.faces
    W = H = 48
    frame = render.ray_frame(sc.cameras[0], W, H)
    f0, t0, _, c0 = O.probe(ref_nodes, sc.positions, ref_idx, ref_faces, frame.pack(), W, H)
    rng = np.random.default_rng(7)
    n = 3000
    o = rng.uniform(-0.4, 0.4, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.concatenate([o, d], 1).astype(np.float32)
    fb, tb, cb = O.trace_rays(ref_nodes, sc.positions, ref_idx, ref_faces, rays)
    sp, _ = _soup(20000)
    inv = np.argsort(sp.order)  # back to triangle order, then split
    for name in ("indices", "faces", "shading", "uvs"):
        setattr(sp, name, np.ascontiguousarray(getattr(sp, name)[inv]))
    S.split_bvh(sp, 4)
    assert sp.split_pieces == len(sp.indices) > 2 * len(ref_idx)
    assert np.array_equal(np.unique(sp.faces), np.unique(ref_faces))
    f1, t1, _, c1 = O.probe(sp.nodes, sp.positions, sp.indices, sp.faces, frame.pack(), W, H)
    fb1, tb1, cb1 = O.trace_rays(sp.nodes, sp.positions, sp.indices, sp.faces, rays)
    # The same distance bits on every ray; the same polygon except where two
    # polygons meet the ray at exactly the same distance (a shared edge),
    # where the reference keeps the first its walk finds and the two trees
    # walk in different orders: the f64 brute force over the whole triangles
    # confirms each such tie.
    for (fa, ta), (fs, ts), r in (((f0, t0), (f1, t1), None), ((fb, tb), (fb1, tb1), rays)):
        assert np.array_equal(ts.view(np.uint32), ta.view(np.uint32))
        differ = np.nonzero(fs != fa)[0]
        assert len(differ) <= max(2, len(fa) // 1000)
        if len(differ) and r is not None:
            _, dist, margin = O.brute_force(sc.positions, ref_idx, ref_faces, rays=r[differ].astype(np.float64),
                                            threads=4)
            assert np.all(margin <= 1e-6 * dist), margin
        elif len(differ):
            f64 = O.ray_frame(sc.cameras[0].matrix, False, 36.0, 30.0, W, H)
            _, dist, margin = O.brute_force(sc.positions, ref_idx, ref_faces, frame=f64, width=W, height=H, threads=4)
            assert np.all(margin[differ] <= 1e-6 * dist[differ])
    # and with fewer triangle tests from inside the cloud
    assert cb1[2] < 0.5 * cb[2]


def test_emitters_and_cutouts_stay_whole():
    from luce_render import render, scene as S, scenes
    gset = scenes.cornell_render_set(random_count=2000)
    gset.materials["random"] = dict(gset.materials["random"], emission_luminance=2.0)
    sc = render.compile_scene(gset, split_depth=4)
    random_material = sc.material_names.index("random") + 1
    em = sc.shading[:, 3] == random_material
    assert em.sum() == 2000  # each emissive triangle once
    gset = scenes.cornell_render_set(random_count=2000)
    gset.materials["random"] = dict(gset.materials["random"], geometry_opacity=0.5)
    sc = render.compile_scene(gset, split_depth=4)
    assert (sc.shading[:, 3] == random_material).sum() == 2000
    gset = scenes.cornell_render_set(random_count=2000)
    sc = render.compile_scene(gset, split_depth=4)
    assert (sc.shading[:, 3] == random_material).sum() > 4000
    assert S.box_overlap(render.compile_scene(scenes.cornell_render_set(random_count=100000), geometry=False)) > 1.0
