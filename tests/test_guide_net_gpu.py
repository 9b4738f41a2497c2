"""The learned guide's matrix-core kernels against the CPU restatement
(oracle/guide_net.py, itself pinned by finite differences): mixtures of
points, the gradient of a batch, and that Adam steps learn a direction field.
Tolerances: FP16 operands with FP32 accumulation, so mixtures within 1e-2
(directions, log-concentrations, weights) and the gradient within 2% of its
norm per layer."""
import numpy as np
import pytest

from oracle import guide_net as G

pytestmark = pytest.mark.gpu

LO = np.array([-1.0, -0.5, -1.0])
HI = np.array([1.0, 1.5, 1.0])


def say(*a):
    print(*a, flush=True)


def _points(rng, n):
    p = rng.uniform(LO, HI, (n, 3))
    nn = rng.normal(size=(n, 3))
    return p, nn / np.linalg.norm(nn, axis=1, keepdims=True)


def _params(rng):
    return (rng.normal(0, 1.0, G.PARAMS) * np.sqrt(2.0 / 64)).astype(np.float32)


@pytest.fixture
def net(device):
    from luce_render.guide_net import GuideNet
    n = GuideNet(device, LO, HI, seed=3)
    yield n
    n.destroy()


def test_parameters_round_trip(net):
    rng = np.random.default_rng(1)
    p = _params(rng)
    net.set_params(p)
    assert np.array_equal(net.params(), p)


def test_mixtures_match_oracle(net):
    rng = np.random.default_rng(2)
    params = _params(rng)
    net.set_params(params)
    p, n = _points(rng, 1000)
    got = net.mixtures(np.concatenate([p, n], 1))
    x = G.encode(LO, HI, p, n)
    _, out = G.forward(params.astype(np.float64), x, fp16=True)
    mu, kappa, w = G.mixture(out)
    gmu = got[:, :16].reshape(-1, 4, 4)[:, :, :3]
    gk = got[:, :16].reshape(-1, 4, 4)[:, :, 3]
    gw = got[:, 16:20]
    say(f"mixtures: max |dmu| {np.abs(gmu - mu).max():.2e}, max |dlog kappa| "
        f"{np.abs(np.log(gk) - np.log(kappa)).max():.2e}, max |dw| {np.abs(gw - w).max():.2e}")
    assert np.abs(gmu - mu).max() < 1e-2
    assert np.abs(np.log(gk) - np.log(kappa)).max() < 1e-2
    assert np.abs(gw - w).max() < 1e-2


def test_gradient_matches_oracle(net):
    rng = np.random.default_rng(4)
    params = _params(rng)
    net.set_params(params)
    m = 2000
    p, n = _points(rng, m)
    d = rng.normal(size=(m, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    w = rng.exponential(1.0, m)
    w[::9] = 0.0
    s = np.concatenate([p, n, d, w[:, None]], 1)
    grad, loss = net.gradient(s)
    ref, ref_loss = G.gradient(params.astype(np.float64), LO, HI, s.astype(np.float32).astype(np.float64))
    say(f"loss gpu {loss:.6f} oracle {ref_loss:.6f}")
    assert abs(loss - ref_loss) <= 1e-3 * abs(ref_loss) + 1e-4
    bounds = [(l * 64 * 64, l * 64 * 64 + 64 * G.LAYER_OUT[l]) for l in range(4)] + \
        [(G.BIAS + l * 64, G.BIAS + l * 64 + G.LAYER_OUT[l]) for l in range(4)]
    for a, b in bounds:
        err = np.linalg.norm(grad[a:b] - ref[a:b]) / max(np.linalg.norm(ref[a:b]), 1e-12)
        say(f"  params {a}..{b}: relative error {err:.2e}")
        assert err < 2e-2


def test_training_learns_a_direction_field(net):
    """Samples whose directions cluster about a field that turns with the
    position (d ~ vMF(normalize(x, 1, z), 50)): after 600 Adam steps the
    mean -log q drops below what the best single fixed lobe could reach."""
    rng = np.random.default_rng(6)

    def batch(m):
        p, n = _points(rng, m)
        axis = np.stack([p[:, 0], np.ones(m), p[:, 2]], 1)
        axis /= np.linalg.norm(axis, axis=1, keepdims=True)
        d = axis + rng.normal(0, 0.14, (m, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        return np.concatenate([p, n, d, np.ones((m, 1))], 1)

    first = net.step(batch(4096), rate=5e-3)
    for _ in range(600):
        last = net.step(batch(4096), rate=5e-3)
    _, held = net.gradient(batch(8192))
    say(f"training: loss {first:.3f} -> {last:.3f}, held-out {held:.3f}")
    # a lobe of concentration 50 about the right axis scores about -log(50/2pi) = -2.07
    assert held < -1.0
    assert held < first - 2.0
