"""Gaussian splats on the CPU (src/render/splats.lucb, shaders/splats.glsl):
the oracle's model pinned by closed forms (the SH basis orthonormal over the
sphere, alpha the maximum of the Gaussian's response along the ray), and the
packing (records, harmonics, boxes, the splats' BVH, placement) against the
oracle evaluated from the splats' description."""
import math

import numpy as np
import pytest

from luce_render import render, scenes
from luce_render import splats as SP
from luce_render._lib import LuceError
from luce_render.scene import GeometrySet, Instance, SplatCloud
from oracle import splats as OS


def test_sh_basis_is_orthonormal():
    """The 15 band-1..3 functions (3DGS's constants and signs) are
    orthonormal over the sphere: Gauss-Legendre in cos(theta) x uniform phi."""
    xs, ws = np.polynomial.legendre.leggauss(24)
    phis = (np.arange(48) + 0.5) * (2 * math.pi / 48)
    gram = np.zeros((15, 15))
    for z, w in zip(xs, ws):
        r = math.sqrt(1 - z * z)
        for phi in phis:
            y = OS.sh_basis((r * math.cos(phi), r * math.sin(phi), z), 15)
            gram += np.outer(y, y) * w * (2 * math.pi / 48)
    assert np.allclose(gram, np.eye(15), atol=1e-12)


def test_alpha_is_the_maximum_response_along_the_ray():
    rng = np.random.default_rng(4)
    sp = scenes.splat_ball(count=40, seed=5)
    sp["opacity"][:] = 0.9
    for _ in range(20):
        o = rng.normal(size=3) * 3
        target = sp["center"][rng.integers(40)] + rng.normal(size=3) * 0.05
        d = (target - o) / np.linalg.norm(target - o)
        al, t = OS.alpha(sp, o, d)
        ts = np.linspace(0, 10, 200001)
        r = OS.quat_matrix(sp["orient"]).transpose(0, 2, 1)
        for i in np.nonzero(al > 0)[0]:
            p = o + ts[:, None] * d - sp["center"][i]
            local = (p @ r[i].T) / sp["scale"][i]
            resp = 0.9 * np.exp(-0.5 * (local * local).sum(1))
            k = resp.argmax()
            assert abs(ts[k] - t[i]) < 1e-3
            assert abs(min(0.99, resp[k]) - al[i]) < 1e-6


def _record_alpha(record, o, d):
    """Alpha from a packed record as the kernels compute it, in f64."""
    center, opacity, turn, inv, _, _, _, _ = SP.decode_records(record[None])
    rot = OS.quat_matrix(turn)[0]  # the record's turn, world to the splat's axes
    lo = rot @ (np.asarray(o) - center[0]) * inv[0]
    ld = rot @ np.asarray(d) * inv[0]
    dd = ld @ ld
    c = np.cross(lo, ld)
    m2 = c @ c / dd
    if m2 > 9.0:
        return 0.0, -(lo @ ld) / dd
    a = min(0.99, opacity[0] * math.exp(-0.5 * m2))
    return (a if a >= OS.CULL else 0.0), -(lo @ ld) / dd


@pytest.mark.parametrize("srgb", [False, True])
def test_records_and_harmonics_reproduce_the_description(srgb):
    """Packed records (f32) and harmonics (halves) give the description's
    alpha, distance and colour along random rays (srgb: the DC and SH in the
    sRGB encoding, decoded after the bands)."""
    sp = scenes.splat_ball(count=300, seed=9)
    sc = render.compile_scene(GeometrySet(clouds=[scenes.splat_cloud(sp, "srgb" if srgb else "linear")]))
    recs = sc.splats.records
    order = sc.splats.order
    assert sorted(order.tolist()) == list(range(300))
    rng = np.random.default_rng(2)
    for k in range(0, 300, 7):
        i = int(order[k])
        one = {key: v[i:i + 1] for key, v in sp.items()}
        for _ in range(3):
            o = rng.normal(size=3) * 2.5
            d = one["center"][0] + rng.normal(size=3) * 0.05 - o
            d /= np.linalg.norm(d)
            want_a, want_t = OS.alpha(one, o, d)
            got_a, got_t = _record_alpha(recs[k], o, d)
            assert abs(got_a - want_a[0]) < 2e-6 * max(1.0, want_a[0]) + 1e-6
            if want_a[0] > 0:
                assert abs(got_t - want_t[0]) < 1e-5 * max(1.0, abs(want_t[0]))
            _, _, _, _, first, encoded, items, flag = SP.decode_records(recs[k][None])
            frame, items_sh = SP.harmonics_of(sc.splats.harmonics, int(first[0]), int(items[0]))
            assert items[0] == 3 and bool(flag[0]) == srgb
            assert np.allclose(frame, [0, 0, 0, 1])
            assert np.allclose(items_sh, one["sh"][0], atol=1e-3)
            want = OS.display_color(one["color"][0], one["sh"][0], d, srgb)
            e = encoded[0] + OS.sh_basis(d, 3) @ items_sh
            e = np.maximum(e, 0.0)
            got = OS.srgb_decode(e) if srgb else e
            assert np.allclose(got, want, atol=2e-3)


def test_boxes_hold_the_splats_out_to_their_cull():
    """Each box holds the ellipsoid where alpha falls to the cull (at most 3
    sigma) and touches it on every side."""
    sp = scenes.splat_ball(count=200, seed=3)
    sc = render.compile_scene(GeometrySet(clouds=[scenes.splat_cloud(sp)]))
    rng = np.random.default_rng(1)
    dirs = rng.normal(size=(4000, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    for k in range(0, 200, 5):
        i = int(sc.splats.order[k])
        op = sp["opacity"][i]
        kk = min(3.0, math.sqrt(max(2 * math.log(op / OS.CULL), 0.0))) if op > OS.CULL else 0.0
        rot = OS.quat_matrix(sp["orient"][i:i + 1])[0]
        pts = sp["center"][i] + (dirs * sp["scale"][i] * kk) @ rot.T
        lo, hi = sc.splats.boxes[k, :3].astype(np.float64), sc.splats.boxes[k, 3:].astype(np.float64)
        assert np.all(pts >= lo - 1e-5) and np.all(pts <= hi + 1e-5)
        assert np.all(pts.min(0) - lo < 0.02 * (hi - lo) + 1e-6) and np.all(hi - pts.max(0) < 0.02 * (hi - lo) + 1e-6)


def test_splat_tree_covers_every_splat_once():
    sp = scenes.splat_ball(count=3000, seed=17)
    sc = render.compile_scene(GeometrySet(clouds=[scenes.splat_cloud(sp)]))
    nodes, boxes = sc.splats.nodes, sc.splats.boxes
    seen = np.zeros(3000, int)
    stack = [0]
    while stack:
        i = stack.pop()
        n = nodes[i]
        if n["count"]:
            f, c = int(n["first"]), int(n["count"])
            seen[f:f + c] += 1
            assert np.all(boxes[f:f + c, :3] >= n["lo"]) and np.all(boxes[f:f + c, 3:] <= n["hi"])
        else:
            for ch in (int(n["first"]), int(n["first"]) + 1):
                assert np.all(nodes[ch]["lo"] >= n["lo"]) and np.all(nodes[ch]["hi"] <= n["hi"])
                stack.append(ch)
    assert np.all(seen == 1)


def test_placed_cloud_is_the_cloud_seen_through_its_placement():
    """An instance turned, scaled uniformly and moved: alpha along a ray in
    world space equals the unplaced cloud's along the ray taken back."""
    sp = scenes.splat_ball(count=60, seed=21)
    rot = OS.quat_matrix(OS.about((0.3, -0.5, 0.8), 1.2))[0]
    s, tr = 1.7, np.array([0.4, -1.0, 2.0])
    m = np.eye(4)
    m[:3, :3] = (rot * s).T  # rows: images of the axes (row vectors)
    m[3, :3] = tr
    inner = GeometrySet(clouds=[scenes.splat_cloud(sp)])
    placed = render.compile_scene(GeometrySet(instances=[Instance(inner, m)]))
    plain = render.compile_scene(inner)
    rng = np.random.default_rng(5)
    where = {int(v): k for k, v in enumerate(plain.splats.order)}
    for k in range(60):
        i = int(placed.splats.order[k])
        j = where[i]
        o = rng.normal(size=3) * 3
        c_world = rot @ (sp["center"][i] * s) + tr
        d = c_world + rng.normal(size=3) * 0.05 - o
        d /= np.linalg.norm(d)
        a_world, _ = _record_alpha(placed.splats.records[k], o, d)
        o_local = rot.T @ (o - tr) / s
        d_local = rot.T @ d
        a_local, _ = _record_alpha(plain.splats.records[j], o_local, d_local)
        assert abs(a_world - a_local) < 1e-5
        # The SH follow the placement: the frame turns world d to the cloud's axes.
        _, _, _, _, first, _, items, _ = SP.decode_records(placed.splats.records[k][None])
        frame, _ = SP.harmonics_of(placed.splats.harmonics, int(first[0]), int(items[0]))
        turned = OS.quat_matrix(frame[None])[0] @ d
        assert np.allclose(turned, d_local, atol=2e-3)


def test_non_uniform_placement_and_plain_clouds():
    sp = scenes.splat_ball(count=10, seed=1)
    m = np.diag([1.0, 2.0, 1.0, 1.0])
    with pytest.raises(LuceError):
        render.compile_scene(GeometrySet(instances=[Instance(GeometrySet(clouds=[scenes.splat_cloud(sp)]), m)]))
    sc = render.compile_scene(GeometrySet(clouds=[SplatCloud(points=np.zeros((5, 3)))]))
    assert sc.splats.plain_clouds == 1 and len(sc.splats.records) == 0
    assert sc.splats.nodes["lo"][0][0] > sc.splats.nodes["hi"][0][0]  # the empty tree


def test_composite_of_one_splat_over_a_background():
    """One splat's composite is alpha c + (1 - alpha) behind."""
    sp = scenes.splat_description([(0, 0, 0)], [(0.3, 0.3, 0.3)], [0.6], [(0.5, 0.2, 0.1)])
    o, d = np.array([0.1, 0.0, 3.0]), np.array([0.0, 0.0, -1.0])
    al, _ = OS.alpha(sp, o, d)
    c, kept = OS.composite(sp, o, d, 1e30, behind=(1.0, 1.0, 1.0))
    want = OS.REC709_TO_ACESCG @ np.array([0.5, 0.2, 0.1]) * al[0] + (1 - al[0])
    assert abs(al[0] - 0.6 * math.exp(-0.5 * (0.1 / 0.3) ** 2)) < 1e-12
    assert np.allclose(c, want) and abs(kept - (1 - al[0])) < 1e-15
