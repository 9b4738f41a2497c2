"""The learned guide's CPU restatement (oracle/guide_net.py) pinned by what
the mathematics fixes: its analytic gradient equals central finite
differences of its own f64 loss, the mixture is a normalized density, and a
planted mistake in the gradient is caught."""
import numpy as np

from oracle import guide_net as G


def _setup(seed=3, n=48):
    rng = np.random.default_rng(seed)
    params = rng.normal(0, 0.25, G.PARAMS)
    lo, hi = np.array([-1.0, -1.0, -1.0]), np.array([1.0, 2.0, 1.0])
    s = np.zeros((n, 10))
    s[:, 0:3] = rng.uniform(lo, hi, (n, 3))
    nn = rng.normal(size=(n, 3))
    s[:, 3:6] = nn / np.linalg.norm(nn, axis=1, keepdims=True)
    d = rng.normal(size=(n, 3))
    s[:, 6:9] = d / np.linalg.norm(d, axis=1, keepdims=True)
    s[:, 9] = rng.exponential(1.0, n)
    s[::7, 9] = 0.0  # some invalid samples
    return params, lo, hi, s


def _fd(params, lo, hi, s, i, h=1e-5):
    a, b = params.copy(), params.copy()
    a[i] += h
    b[i] -= h
    return (G.batch_loss(a, lo, hi, s) - G.batch_loss(b, lo, hi, s)) / (2 * h)


def test_gradient_equals_finite_differences():
    params, lo, hi, s = _setup()
    grad, _ = G.gradient(params, lo, hi, s, fp16=False)
    rng = np.random.default_rng(9)
    picks = list(rng.choice(G.W3, 12, replace=False)) + list(G.W3 + rng.choice(64 * 32, 8, replace=False)) + \
        list(G.BIAS + rng.choice(3 * 64 + 32, 8, replace=False))
    for i in picks:
        fd = _fd(params, lo, hi, s, i)
        assert abs(fd - grad[i]) <= 1e-6 + 1e-4 * abs(fd), (i, fd, grad[i])


def test_planted_gradient_mistake_is_caught():
    """Dropping the softmax term's -w_k from the weight-logit gradient makes
    it disagree with the finite differences."""
    params, lo, hi, s = _setup(seed=4)
    real = G.loss_grad

    def wrong(o, d, c):
        loss, g = real(o, d, c)
        _, _, w = G.mixture(o)
        g = g.copy()
        g[:, 4:20:5] -= np.asarray(c)[:, None] * w
        return loss, g

    G.loss_grad = wrong
    try:
        grad, _ = G.gradient(params, lo, hi, s, fp16=False)
    finally:
        G.loss_grad = real
    i = G.BIAS + 3 * 64 + 4  # b3 entry of lobe 0's weight logit
    fd = _fd(params, lo, hi, s, i)
    assert abs(fd - grad[i]) > 1e-3 * max(1.0, abs(fd))


def test_mixture_is_a_density():
    """The vMF mixture integrates to 1 over the sphere (Fibonacci quadrature)."""
    rng = np.random.default_rng(5)
    o = rng.normal(0, 1.0, (3, 32))
    m = 20000
    k = np.arange(m) + 0.5
    z = 1.0 - 2.0 * k / m
    phi = np.pi * (1.0 + 5 ** 0.5) * k
    r = np.sqrt(1.0 - z * z)
    dirs = np.stack([r * np.cos(phi), r * np.sin(phi), z], 1)
    for row in o:
        q = G.pdf(np.tile(row, (m, 1)), dirs)
        assert abs(q.mean() * 4 * np.pi - 1.0) < 2e-3
