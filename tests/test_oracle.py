"""Pins the oracle (oracle/oracle_trace.c) to the reference's known answers and
to closed forms, before anything is compared against it."""
import math

import numpy as np
import pytest

from oracle import oracle as O


def _cam(tx, ty, tz):
    m = np.eye(4)
    m[3, :3] = (tx, ty, tz)
    return m


def test_ray_frame_known_answers():
    """tests.lucb:63-81: 36 mm across at 18 mm is 90 degrees horizontally; the
    centre looks down -Z; half as tall as wide puts the top at 0.5."""
    f = O.ray_frame(_cam(0, 0, 5), False, 36.0, 18.0, 200, 100)
    origin, left = O.frame_ray(f, 0.0, 50.0)
    assert origin[2] == 5.0 and abs(left[0] + 1.0) < 1e-9 and abs(left[2] + 1.0) < 1e-9
    _, centre = O.frame_ray(f, 100.0, 50.0)
    assert abs(centre[0]) < 1e-9 and abs(centre[1]) < 1e-9 and centre[2] == -1.0
    _, top = O.frame_ray(f, 100.0, 0.0)
    assert abs(top[1] - 0.5) < 1e-9


def test_look_frame_orthographic_known_answers():
    """tests.lucb:83-91."""
    f = O.look_frame((0.0, 10.0, 0.0), (0.0, -1.0, 0.0), (0.0, 0.0, -1.0), 2.0, 200, 100, True)
    assert f[17] == 1.0
    corner, along = O.frame_ray(f, 0.0, 0.0)
    assert abs(corner[0] + 4.0) < 1e-9 and abs(corner[2] + 2.0) < 1e-9 and corner[1] == 10.0
    assert along[1] == -1.0 and along[0] == 0.0
    middle, _ = O.frame_ray(f, 100.0, 50.0)
    assert abs(middle[0]) < 1e-9 and abs(middle[2]) < 1e-9


def test_ray_frame_rejects_empty_images():
    with pytest.raises(ValueError):
        O.ray_frame(_cam(0, 0, 5), False, 36.0, 18.0, 0, 10)


def _triangle_scene():
    # Two triangles facing +z at z = -2 and z = -3, and one far off to the side.
    p = np.array([[-1, -1, -2], [1, -1, -2], [0, 1, -2],
                  [-1, -1, -3], [1, -1, -3], [0, 1, -3],
                  [10, 10, -1], [11, 10, -1], [10, 11, -1]], np.float32)
    idx = np.arange(9, dtype=np.uint32).reshape(3, 3)
    faces = np.array([7, 8, 9], np.uint32)
    return p, idx, faces


def test_brute_force_closed_form():
    """Möller–Trumbore in f64: the nearer triangle at t = 2 along -z with the
    next polygon 50% farther; a ray past every triangle misses."""
    p, idx, faces = _triangle_scene()
    rays = np.array([[0, 0, 0, 0, 0, -1], [0, 0, 0, 0, 0, 1], [0.5, 0, 0, 0, 0, -1]], np.float64)
    face, dist, margin = O.brute_force(p, idx, faces, rays=rays)
    assert face[0] == 7 and dist[0] == 2.0 and abs(margin[0] - 0.5) < 1e-12
    assert face[1] == -1
    assert face[2] == 7 and dist[2] == 2.0


def test_brute_force_matches_analytic_plane():
    """A ray from (x, y, 0) toward (0, 0, -1) scaled by s hits z = -2 at t = 2 / s."""
    p, idx, faces = _triangle_scene()
    for s in (0.5, 1.0, 3.0):
        face, dist, _ = O.brute_force(p, idx, faces, rays=np.array([[0.1, 0.2, 0, 0, 0, -s]]))
        assert face[0] == 7 and abs(dist[0] - 2.0 / s) < 1e-15


def _random_soup(n, seed):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-1, 1, (n, 1, 3))
    p = (c + rng.normal(0, 0.05, (n, 3, 3))).reshape(-1, 3).astype(np.float32)
    return p, np.arange(3 * n, dtype=np.uint32).reshape(n, 3)


@pytest.mark.parametrize("n", [1, 3, 4, 5, 16, 17, 100, 3000])
def test_bvh_build_invariants(n):
    """check_bvh (tests.lucb:9-28): each triangle in one leaf, leaves inside
    their boxes, children nested; the f32 SAH costs agree with an f64
    recompute of the same split."""
    p, idx = _random_soup(n, n)
    nodes, order, gap = O.bvh_build(p, idx)
    assert sorted(order.tolist()) == list(range(n))
    leaf_order = idx[order]
    assert O.check_bvh(nodes, p, leaf_order) == 0
    assert gap < 1e-3
    assert (nodes["count"] <= 16).all()


def test_bvh_build_empty_and_degenerate():
    """No triangles: one inside-out root (bvh.lucb:88-90). Identical centroids:
    bins cannot separate them, so nodes halve by count (bvh.lucb:118-120)."""
    nodes, order, _ = O.bvh_build(np.zeros((0, 3), np.float32), np.zeros((0, 3), np.uint32))
    assert len(nodes) == 1 and nodes[0]["lo"][0] > nodes[0]["hi"][0]
    p = np.tile(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32), (40, 1))
    idx = np.arange(120, dtype=np.uint32).reshape(40, 3)
    nodes, order, _ = O.bvh_build(p, idx)
    leaves = nodes[nodes["count"] > 0]
    assert leaves["count"].sum() == 40 and leaves["count"].max() <= 16
    assert O.check_bvh(nodes, p, idx[order]) == 0


def test_bvh_planted_mistake_is_caught():
    """A leaf that drops a triangle fails check_bvh: the invariant check bites."""
    p, idx = _random_soup(50, 7)
    nodes, order, _ = O.bvh_build(p, idx)
    bad = nodes.copy()
    leaf = int(np.nonzero(bad["count"] > 1)[0][0])
    bad[leaf]["count"] -= 1
    assert O.check_bvh(bad, p, idx[order]) != 0


def test_mirror_agrees_with_brute_force_on_cornell_probe(cornell_probe):
    """The f32 traversal mirror against f64 brute force on every pixel of the
    256x256 Cornell probe, under the reference's rule (tests/gpu/main.lucb:
    61-78): more than half the pixels hit and none disagree."""
    from luce_render import render
    sc = cornell_probe
    frame = render.ray_frame(sc.cameras[0], 256, 256)
    m = sc.cameras[0].matrix
    f64 = O.ray_frame(m, False, 36.0, 30.0, 256, 256)
    face, t, _, counters = O.probe(sc.nodes, sc.positions, sc.indices, sc.faces, frame.pack(), 256, 256)
    bf, bd, bm = O.brute_force(sc.positions, sc.indices, sc.faces, frame=f64, width=256, height=256, threads=8)
    same, differ = O.agrees(face, t, bf, bd, bm)
    assert (bf >= 0).sum() > 256 * 256 // 2
    assert differ.sum() == 0
    assert counters[3] == 0  # no stack entry dropped
    # The stricter bound: every pixel's polygon equal to the brute force's,
    # and every hit distance within 1e-5 relative of its f64 distance.
    hit = bf >= 0
    assert np.array_equal(np.where(face == 0xFFFFFFFF, -1, face.astype(np.int64)), bf)
    rel = np.abs(t[hit].astype(np.float64) - bd[hit]) / bd[hit]
    assert rel.max() <= 1e-5, rel.max()


def test_mirror_planted_mistake_is_caught(cornell_probe):
    """Swapping two polygon ids in the mirror's input makes the comparison fail."""
    from luce_render import render
    sc = cornell_probe
    frame = render.ray_frame(sc.cameras[0], 64, 64)
    f64 = O.ray_frame(sc.cameras[0].matrix, False, 36.0, 30.0, 64, 64)
    faces = sc.faces.copy()
    wall = faces == 2  # the back wall
    faces[wall] = 3
    face, t, _, _ = O.probe(sc.nodes, sc.positions, sc.indices, faces, frame.pack(), 64, 64)
    bf, bd, bm = O.brute_force(sc.positions, sc.indices, sc.faces, frame=f64, width=64, height=64, threads=8)
    _, differ = O.agrees(face, t, bf, bd, bm)
    assert differ.sum() > 0


def test_adaptive_verdict_closed_forms():
    """The oracle's convergence check (adaptive.comp:16-42) on films whose
    errors are known in closed form: a pixel whose half estimate equals its
    full one has error 0; one whose channels differ by d at mean brightness b
    over weight w has error 3d/w / (1e-4 + sqrt(3b/w)); a single unconverged
    pixel unconverges exactly its 3 x 3 neighbourhood, clipped at the edges."""
    h, w = 6, 7
    s = np.zeros((h, w, 4), np.float32)
    s[..., :3] = 8.0
    s[..., 3] = 16.0
    half = s.copy()
    half[..., 3] = 0.0
    err = O.adaptive_error(s, half)
    assert np.all(err == 0.0)
    half[2, 3, :3] += 0.8  # d = 0.8 on each channel
    want = (3 * 0.8 / 16.0) / (1e-4 + np.sqrt(3 * 8.0 / 16.0))
    err = O.adaptive_error(s, half)
    assert abs(float(err[2, 3]) - want) < 1e-6 and (err > 0).sum() == 1
    own, conv = O.adaptive_verdict(s, half, want * 0.99)
    assert own.sum() == h * w - 1
    expect = np.ones((h, w), bool)
    expect[1:4, 2:5] = False
    assert np.array_equal(conv, expect)
    half[0, 0, :3] += 0.8  # a corner: its clipped neighbourhood is 2 x 2
    _, conv = O.adaptive_verdict(s, half, want * 0.99)
    expect[0:2, 0:2] = False
    assert np.array_equal(conv, expect)
    _, conv = O.adaptive_verdict(s, half, want * 1.01)
    assert conv.all()
