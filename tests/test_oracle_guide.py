"""The oracle's restatement of the hash-grid guide (shaders/guide.glsl:39-139,
guide_fit.comp:16-35) pinned by what the mathematics fixes: a vMF lobe's
density integrates to one, its sampler's mean resultant length is
coth(kappa) - 1/kappa, the fit recovers a lobe's mean and Banerjee's kappa from
samples of a known lobe and decays its statistics by 0.9, and the cell key
follows the grid's power-of-two levels and the normal's octant."""
import math

import numpy as np

from oracle import oracle as O


def _t_grid(n=400):
    t, w = np.polynomial.legendre.leggauss(n)
    return t, w


def test_vmf_density_integrates_to_one():
    mean = np.array([0.0, 0.0, 1.0])
    t, w = _t_grid()
    for kappa in (1e-4, 0.5, 2.0, 10.0, 100.0):
        pdf = np.array([O.vmf_pdf(mean, kappa, [math.sqrt(max(0.0, 1 - x * x)), 0.0, x]) for x in t])
        total = 2 * math.pi * float((pdf * w).sum())
        assert abs(total - 1.0) < 1e-6, (kappa, total)


def test_vmf_sampler_has_the_lobes_mean_resultant_length():
    rng = np.random.default_rng(3)
    axis = rng.normal(size=3)
    axis /= np.linalg.norm(axis)
    fine = (np.arange(8192) + 0.5) / 8192
    grid = (np.arange(64) + 0.5) / 64
    for kappa in (0.5, 3.0, 20.0, 200.0):
        # the cosine to the axis depends on ux alone: its mean over ux
        d = np.array([O.vmf_sample(axis, kappa, ux, 0.37) for ux in fine])
        assert np.allclose(np.linalg.norm(d, axis=1), 1.0, atol=1e-12)
        want = 1.0 / math.tanh(kappa) - 1.0 / kappa
        assert abs((d @ axis).mean() - want) < 5e-5, (kappa, (d @ axis).mean(), want)
        # the spread around the axis has no preferred side
        d = np.array([O.vmf_sample(axis, kappa, ux, uy) for ux in grid for uy in grid])
        side = d - np.outer(d @ axis, axis)
        assert np.linalg.norm(side.mean(axis=0)) < 1e-3


def test_guide_fit_recovers_a_lobe_with_banerjees_kappa():
    rng = np.random.default_rng(5)
    mu = np.array([0.3, -0.5, 0.8])
    mu /= np.linalg.norm(mu)
    kappa = 8.0
    d = np.array([O.vmf_sample(mu, kappa, a, b) for a, b in rng.uniform(size=(20000, 2))])
    stats = np.zeros((4, 5), np.float32)
    stats[1, 0] = len(d)
    stats[1, 1:4] = d.sum(axis=0)
    stats[1, 4] = len(d)
    stats[3] = [2.0, 0.0, 0.0, 2.0, 2.0]  # two samples along +z: kappa at its cap
    before = np.zeros((4, 4), np.float32)
    before[0] = [1, 0, 0, 5]
    lobes, weights, after = O.guide_fit_cell(stats, np.concatenate([before.reshape(16), np.zeros(4)]))
    r = np.linalg.norm(d.sum(axis=0)) / len(d)
    banerjee = r * (3 - r * r) / (1 - r * r)
    assert np.allclose(lobes[1, :3], d.sum(axis=0) / np.linalg.norm(d.sum(axis=0)), atol=1e-6)
    assert abs(lobes[1, 3] - banerjee) < 1e-3 * banerjee
    assert abs(lobes[1, 3] / kappa - 1.0) < 0.1  # the approximation is close at kappa 8
    assert lobes[3, 3] == np.float32(1000.0) and np.allclose(lobes[3, :3], [0, 0, 1])
    assert np.array_equal(lobes[0], before[0])  # a lobe without samples keeps its fit
    assert np.array_equal(weights, stats[:, 0])
    assert np.allclose(after, stats * np.float32(0.9), rtol=1e-7)


def test_guide_key_follows_levels_and_octants():
    from luce_render.loading import Loaded, RenderSettings
    from luce_render import render, scenes
    sc = render.compile_scene(scenes.cornell_render_set())
    st = RenderSettings(96, 96, 4, spectral=False, guiding=1)
    frame = render.ray_frame(sc.cameras[0], 96, 96)
    ld = Loaded(sc, st, frame)
    f16 = frame.pack()
    cam = f16[0:3].astype(np.float64)
    minimum, per_unit = float(ld.constants[52]), float(ld.constants[49])
    up = np.array([0.0, 1.0, 0.0])
    for dist in (0.5, 1.7, 3.9):
        wanted = max(minimum, dist * per_unit)
        size = minimum * 2 ** math.ceil(math.log2(wanted / minimum))
        # a point in the middle of a cell at that distance, then across its faces
        base = cam + np.array([0.0, 0.0, -dist])
        centre = (np.floor(base / size) + 0.5) * size
        k = O.guide_key(ld.constants, f16, centre, up)
        assert O.guide_key(ld.constants, f16, centre + 0.3 * size * np.array([1, -1, 1]), up) == k
        for axis in range(3):
            step = np.zeros(3)
            step[axis] = size
            assert O.guide_key(ld.constants, f16, centre + step, up) != k
        assert O.guide_key(ld.constants, f16, centre, -up) != k
        assert O.guide_key(ld.constants, f16, centre, np.array([0.1, 0.99, 0.0])) == k
