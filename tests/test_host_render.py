"""Host-side pins of the render path's inputs (no GPU): the reference's own
known answers for the spectrum fit, blackbody and light tree
(src/render/tests.lucb:93-124, 155-183), the sampler's direction table and
stratification, and the packed layouts."""
import math

import numpy as np
import pytest

from oracle import oracle as O


def test_white_fit_is_constant_and_film_returns_white():
    """tests.lucb:93-99."""
    from luce_render import spectrum as S
    white = S.fit_spectrum((1.0, 1.0, 1.0))
    assert white.c0 == 0.0 and white.c1 == 0.0 and white.c2 == 0.0 and abs(white.at(500.0) - 1.0) < 1e-12
    back = S.film_matrix() @ np.ones(3)
    assert np.all(np.abs(back - 1.0) < 1e-6)


@pytest.mark.parametrize("color", [(0.8, 0.2, 0.1), (0.05, 0.5, 0.2), (0.1, 0.2, 0.9), (0.9, 0.9, 0.2)])
def test_fit_round_trip(color):
    """tests.lucb:100-109: a colour's fitted spectrum through the film comes back within 0.01."""
    from luce_render import spectrum as S
    fit = S.fit_spectrum(color)
    lam = np.arange(S.LAMBDA_MIN, S.LAMBDA_MAX + 0.5, 1.0)
    xyz = (S.cmf(lam) * fit.at(lam)[:, None]).sum(axis=0) / S.cmf_integrals()
    rgb = S.film_matrix() @ xyz
    assert np.all(np.abs(rgb - np.array(color)) < 0.01), rgb


def test_blackbody_colours():
    """tests.lucb:111-122: 5455 K near neutral, 3000 K warm, 9000 K cool."""
    from luce_render import spectrum as S
    m = S.film_matrix()
    n = m @ (S.blackbody_xyz(5455.0) / S.blackbody_xyz(5455.0)[1])
    assert np.all(np.abs(n - 1.0) < 0.06)
    w = m @ (S.blackbody_xyz(3000.0) / S.blackbody_xyz(3000.0)[1])
    assert w[0] > 1.2 * w[1] and w[1] > w[2]
    k = m @ (S.blackbody_xyz(9000.0) / S.blackbody_xyz(9000.0)[1])
    assert k[2] > k[1] > k[0]


def test_cmf_integrals_make_constant_spectrum_white():
    """spectrum.lucb:12-14: the constant spectrum 1 is XYZ (1, 1, 1) under E."""
    from luce_render import spectrum as S
    lam = np.arange(S.LAMBDA_MIN, S.LAMBDA_MAX + 0.5, 1.0)
    assert np.allclose(S.cmf(lam).sum(axis=0) / S.cmf_integrals(), 1.0)


def test_light_tree_trails_reach_every_emitter_and_hold_energy():
    """tests.lucb:155-183: an emissive 9x9 grid (162 triangles): every
    emitter's trail from the root ends at its leaf, depth <= 8, the triangle
    map points back, and the root holds the total energy."""
    from luce_render import render
    from luce_render.loading import OpenPbr, build_light_tree
    xs = np.linspace(-2, 2, 10)
    px, pz = np.meshgrid(xs, xs, indexing="ij")
    pts = np.stack([px.ravel(), np.zeros(px.size), pz.ravel()], 1)
    tris = []
    for i in range(9):
        for j in range(9):
            a, b, c, d = i * 10 + j, (i + 1) * 10 + j, (i + 1) * 10 + j + 1, i * 10 + j + 1
            tris += [[a, b, c], [a, c, d]]
    sc = render.compile_scene(render.GeometrySet(meshes=[render.Mesh(pts, np.array(tris))]))
    glowing = OpenPbr(emission_luminance=2.0)
    nodes, links, records, count, energy = build_light_tree(sc, [], 0, [glowing])
    assert count == sc.triangle_count() == 162
    total = 0.0
    for index in range(count):
        trail, depth = int(records[index * 4 + 1]), int(records[index * 4 + 2])
        assert depth <= 8
        node = 0
        for level in range(depth):
            node = int(links[node * 2]) + ((trail >> level) & 1)
        assert links[node * 2 + 1] == 0x80000000 and links[node * 2] == index
        code = int(records[index * 4])
        assert code & 0x80000000 and records[count * 4 + (code & 0x7FFFFFFF)] == index
        total += nodes[node * 12 + 3]
    assert abs(nodes[3] / total - 1.0) < 1e-4


def _sobol_directions():
    """Joe–Kuo direction numbers for dimensions 1..3 (new-joe-kuo-6.21201:
    s, a, m = (1,0,[1]), (2,1,[1,3]), (3,1,[1,3,1])), dimension 0 van der Corput."""
    params = [(1, 0, [1]), (2, 1, [1, 3]), (3, 1, [1, 3, 1])]
    rows = [[1 << (31 - b) for b in range(32)]]
    for s, a, m in params:
        v = [0] * 33
        for k in range(1, 33):
            if k <= s:
                v[k] = m[k - 1] << (32 - k)
            else:
                x = v[k - s] ^ (v[k - s] >> s)
                for i in range(1, s):
                    if (a >> (s - 1 - i)) & 1:
                        x ^= v[k - i]
                v[k] = x
        rows.append([x & 0xFFFFFFFF for x in v[1:]])
    return rows


def _table_in(path, name):
    """The 128 words of the table `name` in a C/HIP source of this build."""
    import re
    text = path.read_text()
    body = text[text.index(name):]
    body = body[body.index("{") + 1:body.index("};")]
    return [int(w, 16) for w in re.findall(r"0x([0-9a-fA-F]{8})u", body)]


def test_sobol_table_is_joe_kuo():
    """The kernels' and the oracle's direction tables equal, word for word,
    the reference's committed table (shaders/sobol_table.glsl, kept as the
    fixture tests/golden/sobol_directions.json), and that table is Joe and
    Kuo's, regenerated from their parameters."""
    import json
    from pathlib import Path
    root = Path(__file__).resolve().parents[1]
    fixture = json.loads((root / "tests" / "golden" / "sobol_directions.json").read_text())
    ref = [int(w, 16) for w in fixture["directions"]]
    assert len(ref) == 128
    rows = _sobol_directions()
    assert [w for d in range(4) for w in rows[d][:32]] == ref
    kernels = _table_in(root / "luce-render_amd" / "csrc" / "render_common.h", "kSobol[128]")
    oracle = _table_in(root / "oracle" / "oracle_render.c", "SOBOL[128]")
    assert kernels == ref and oracle == ref


def test_sampler_stratifies_every_dimension():
    """Owen scrambling keeps Sobol's elementary intervals: the first 2^k
    samples of a pixel put exactly one point in each 1/2^k interval of every
    coordinate (a property a wrong scramble, table or bit order breaks)."""
    for group in (0, 3, 7):
        for pixel in (0, 12345):
            vals = np.array([O.sample4(1, pixel, i, group)[1] for i in range(64)])
            for d in range(4):
                bins = np.floor(vals[:, d] * 64).astype(int)
                assert sorted(bins.tolist()) == list(range(64)), (group, pixel, d)


def test_sampler_differs_by_pixel_and_group():
    a = O.sample4(1, 5, 0, 0)[0]
    assert not np.array_equal(a, O.sample4(1, 6, 0, 0)[0])
    assert not np.array_equal(a, O.sample4(1, 5, 0, 1)[0])
    assert not np.array_equal(a, O.sample4(2, 5, 0, 0)[0])


def test_oracle_tables_bound_albedo():
    """The f64 albedo tables (tables.comp's estimator): E_ggx in (0, 1],
    nearly 1 when smooth, falling with roughness at grazing angles; the
    dielectric table below the GGX one (Fresnel <= 1)."""
    t = O.tables()
    ggx = t[:32 * 32].reshape(32, 32)  # [roughness][mu]
    assert np.all(ggx > 0) and np.all(ggx <= 1.0 + 1e-9)
    assert ggx[0].min() > 0.99
    assert ggx[31, 0] < ggx[0, 0]
    die = t[32 * 32 + 32:].reshape(16, 32, 32)
    assert np.all(die <= ggx[None] + 1e-12)


def test_packed_layouts():
    """Light records are 7 float4 (lights.lucb), materials 16 float4
    (materials.lucb), the constants 32 rows."""
    from luce_render import render, scenes
    from luce_render.loading import Loaded, RenderSettings
    sc = render.compile_scene(scenes.cornell_render_set())
    st = RenderSettings.of(64, 64, 1)
    L = Loaded(sc, st, render.ray_frame(sc.cameras[0], 64, 64))
    assert L.light_count == 1 and L.emitter_count == 1
    rec = L.lights[:28]
    assert rec[3] == 0.0  # rect
    assert abs(rec[15] - 0.0625) < 1e-6  # area of 0.25 x 0.25
    assert np.allclose(rec[12:15], [0, -1, 0])  # shines down
    assert abs(rec[20] - 17.0) < 1e-5  # radiance
    assert len(L.materials) == 16 * 4 * (1 + len(sc.materials))
    assert L.constants.size == 128 and L.constants[1] == 12
    assert L.light_paths == 64 * 64 // 2


def _tiled_pixel(i, w, h):
    """The camera kernel's walk of the image in 8x8 tiles (render_kernels.hip
    tiled_pixel), restated."""
    tile_row, local = divmod(i, 8 * w)
    rows = min(8, h - 8 * tile_row)
    tile, within = divmod(local, 8 * rows)
    cols = min(8, w - 8 * tile)
    return (8 * tile_row + within // cols) * w + 8 * tile + within % cols


def test_tiled_camera_walk_visits_every_pixel_once():
    for w, h in [(1, 1), (8, 8), (7, 5), (9, 17), (64, 3), (33, 41), (96, 96), (250, 13)]:
        seen = sorted(_tiled_pixel(i, w, h) for i in range(w * h))
        assert seen == list(range(w * h)), (w, h)
    # a wave of 64 consecutive paths stays in one 8x8 tile where tiles are whole
    assert {(_tiled_pixel(i, 64, 64) % 64) // 8 for i in range(64)} == {0}


def _dome_oracle(texture, width, height, samples, seed, pixels, spectral=False):
    from luce_render import render, scenes
    from luce_render.loading import Loaded, RenderSettings
    sc = render.compile_scene(scenes.dome_set(texture=texture, ground=False))
    st = RenderSettings.of(width, height, samples)
    st.spectral, st.light_tracing, st.max_bounces = spectral, 0.0, 8
    frame = render.ray_frame(sc.cameras[0], width, height)
    loaded = Loaded(sc, st, frame)
    return frame, O.render(sc, loaded, frame.pack(), width, height, samples, 8, spectral, seed,
                           np.asarray(pixels, np.uint32))


def test_dome_background_is_the_texel_radiance():
    """A textured dome (environment.lucb; lights.glsl:89-117): a camera ray that
    meets nothing sees the radiance of the texel its direction falls in, the
    image's center down the light's -Z and its top up +Y. Checked on pixels
    whose four corner rays fall in one texel of an 8 x 4 image of random
    colours and pass the sphere: the oracle's value is that texel's ACEScg."""
    from luce_render.loading import environment_of
    rng = np.random.default_rng(3)
    tex = rng.uniform(0.1, 2.0, (4, 8, 3))
    env = environment_of(tex)
    W, H = 48, 36
    frame, img = _dome_oracle(tex, W, H, 4, 1, np.arange(W * H))
    f = frame.pack().astype(np.float64)
    corner, du, dv = f[4:7], f[8:11], f[12:15]

    def texel(d):
        d = d / np.linalg.norm(d)
        u = 0.5 + np.arctan2(d[0], -d[2]) * (0.5 / np.pi)
        v = np.arccos(np.clip(d[1], -1.0, 1.0)) / np.pi
        return min(int(v * env.height), env.height - 1) * env.width + min(int(u * env.width), env.width - 1)

    def passes(d):  # the unit sphere at the origin, seen from (0, 0, 4)
        o = f[0:3]
        d = d / np.linalg.norm(d)
        return np.linalg.norm(o - d * (o @ d)) > 1.05 or o @ d > 0

    checked = 0
    rgb = env.texels.reshape(-1, 4)[:, :3].astype(np.float64)
    for y in range(H):
        for x in range(W):
            ds = [corner + (x + a) * du + (y + b) * dv for a in (0, 1) for b in (0, 1)]
            ids = {texel(d) for d in ds}
            if len(ids) != 1 or not all(passes(d) for d in ds):
                continue
            want = rgb[ids.pop()]
            got = img[y * W + x]
            assert np.all(np.abs(got - want) <= 1e-6 * want), (x, y, got, want)
            checked += 1
    assert checked >= 100


def test_dome_lighting_matches_quadrature():
    """The dome's importance sampling and pdf (lights.glsl:119-175) carry no
    bias: a Lambert sphere (albedo 0.8) under the synthetic sky, at the point
    facing the camera (normal +Z), reflects 0.8 / pi times the integral of the
    environment's radiance times the cosine, computed here by quadrature over
    every texel (32 x 32 points each). 4 x 2048 samples of the oracle's
    estimator (light samples by the texel CDFs, BSDF samples, MIS) agree
    within 1.5% (the estimate's spread is about 0.5%)."""
    from luce_render import scenes
    from luce_render.loading import environment_of
    tex = scenes.sky_texture()
    env = environment_of(tex)
    W, H = env.width, env.height
    rgb = env.texels.reshape(H, W, 4)[:, :, :3].astype(np.float64)
    k = 32
    sub = (np.arange(k) + 0.5) / k
    total = np.zeros(3)
    for y in range(H):
        th = np.pi * (y + sub) / H
        for x in range(W):
            T, P = np.meshgrid(th, 2 * np.pi * ((x + sub) / W - 0.5), indexing="ij")
            cosn = np.maximum(0.0, -np.sin(T) * np.cos(P))
            total += rgb[y, x] * (cosn * np.sin(T)).mean() * (np.pi / H) * (2 * np.pi / W)
    expected = 0.8 / np.pi * total
    got = np.mean([_dome_oracle(tex, 33, 25, 2048, seed, [12 * 33 + 16])[1][0] for seed in (1, 2, 3, 4)], axis=0)
    assert np.all(np.abs(got / expected - 1.0) < 0.015), (got, expected)


def test_environment_cdfs():
    """environment_of (environment.lucb:111-156): each row's cumulative
    weights rise to 1 (evenly across an all-black row), the rows' too; the
    weights are luminance x sin theta of the ACEScg texels; the mean is the
    solid-angle average of luminance."""
    from luce_render import spectrum as S
    from luce_render.loading import environment_of
    rng = np.random.default_rng(2)
    tex = rng.uniform(0.0, 3.0, (6, 10, 3))
    tex[2] = 0.0
    env = environment_of(tex)
    W, H = 10, 6
    t = env.texels.reshape(H, W, 4).astype(np.float64)
    sin_theta = np.sin(np.pi * (np.arange(H) + 0.5) / H)
    lum = S.luminance(np.moveaxis(t[:, :, :3], 2, 0))
    assert np.allclose(t[:, :, 3], lum * sin_theta[:, None], rtol=1e-6)
    rows = env.cdf[:W * H].reshape(H, W)
    assert np.all(np.diff(rows, axis=1) >= 0) and np.allclose(rows[:, -1], 1.0)
    assert np.allclose(rows[2], (np.arange(W) + 1.0) / W)
    assert np.all(np.diff(env.cdf[W * H:]) >= 0) and abs(env.cdf[-1] - 1.0) < 1e-6
    assert abs(env.total - (lum * sin_theta[:, None]).sum()) < 1e-6 * env.total
    assert abs(env.average - (lum * sin_theta[:, None]).sum() / (W * sin_theta.sum())) < 1e-6 * env.average
