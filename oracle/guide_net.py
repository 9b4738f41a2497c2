"""CPU restatement of the learned guide (luce-render_amd/csrc/guide_net.h,
guide_net_kernels.h) in numpy float64: the one-blob encoding, the 64-64-64-64-32
MLP, the four-lobe vMF mixture it outputs, the loss -c log q(d) and its
gradient by backpropagation. The network has no counterpart in the reference
(which guides with its hash grid, shaders/guide.glsl); this restatement is the
checker for the GPU kernels, pinned itself by central finite differences of
the loss (tests/test_oracle_guide_net.py).

With fp16=True it rounds where the GPU kernels do: weights, the layer inputs
(encoding and ReLU outputs) and the backpropagated dZ enter the matrix
products as IEEE half; products accumulate in higher precision.

Test infrastructure only: never imported by the product.
"""
import numpy as np

IN, HIDDEN, OUT, LOBES = 64, 64, 32, 4
POS_BINS, NORMAL_BINS = 16, 8
LAYER_OUT = (HIDDEN, HIDDEN, HIDDEN, OUT)
W3 = 3 * IN * HIDDEN
BIAS = W3 + HIDDEN * OUT
PARAMS = BIAS + 3 * HIDDEN + OUT
LOG_KAPPA_LO, LOG_KAPPA_HI = -4.0, 7.0
MAX_SCALE = 1024.0


def unpack(params):
    """(weights [in, out] per layer, biases per layer) from the flat vector."""
    params = np.asarray(params, np.float64)
    ws, bs = [], []
    for l in range(4):
        o = LAYER_OUT[l]
        ws.append(params[l * IN * HIDDEN: l * IN * HIDDEN + IN * o].reshape(IN, o))
        bs.append(params[BIAS + l * HIDDEN: BIAS + l * HIDDEN + o])
    return ws, bs


def _half(x, on):
    return np.asarray(x, np.float16).astype(np.float64) if on else np.asarray(x, np.float64)


def octahedral_uv(n):
    n = np.asarray(n, np.float64)
    s = np.abs(n).sum(-1, keepdims=True)
    s = np.where(s > 0, s, 1.0)
    x, y, z = n[..., 0] / s[..., 0], n[..., 1] / s[..., 0], n[..., 2]
    ox = (1.0 - np.abs(y)) * np.where(x >= 0, 1.0, -1.0)
    oy = (1.0 - np.abs(x)) * np.where(y >= 0, 1.0, -1.0)
    x, y = np.where(z < 0, ox, x), np.where(z < 0, oy, y)
    return 0.5 * x + 0.5, 0.5 * y + 0.5


def encode(bounds_lo, bounds_hi, p, n):
    """One-blob features (N, 64): x, y, z over 16 bins each (position scaled
    to [0, 1] by the bounds and clamped), the normal's octahedral u, v over 8."""
    p = np.asarray(p, np.float64)
    lo = np.asarray(bounds_lo, np.float64)
    ext = np.asarray(bounds_hi, np.float64) - lo
    inv = np.where(ext > 0, 1.0 / np.where(ext > 0, ext, 1.0), 1.0)
    u = np.clip((p - lo) * inv, 0.0, 1.0)
    f = np.zeros((len(p), IN))
    for d in range(3):
        c = (np.arange(POS_BINS) + 0.5) / POS_BINS
        f[:, 16 * d:16 * d + 16] = np.exp(-0.5 * ((u[:, d:d + 1] - c) * POS_BINS) ** 2)
    ou, ov = octahedral_uv(n)
    c = (np.arange(NORMAL_BINS) + 0.5) / NORMAL_BINS
    f[:, 48:56] = np.exp(-0.5 * ((ou[:, None] - c) * NORMAL_BINS) ** 2)
    f[:, 56:64] = np.exp(-0.5 * ((ov[:, None] - c) * NORMAL_BINS) ** 2)
    return f


def forward(params, x, fp16=True):
    """Pre-activations of the hidden layers and the 32 outputs."""
    ws, bs = unpack(params)
    h = _half(x, fp16)
    zs = []
    for l in range(3):
        z = h @ _half(ws[l], fp16) + bs[l]
        zs.append(z)
        h = _half(np.maximum(z, 0.0), fp16)
    out = h @ _half(ws[3], fp16) + bs[3]
    return zs, out


def mixture(o):
    """Per point: directions (N, 4, 3), concentrations (N, 4), weights (N, 4)."""
    o = np.asarray(o, np.float64)
    v = o[:, :20].reshape(-1, LOBES, 5)
    vec = v[:, :, :3]
    ln = np.linalg.norm(vec, axis=-1, keepdims=True)
    small = ln < 1e-6
    mu = np.where(small, np.array([0.0, 0.0, 1.0]), vec / np.where(small, 1.0, ln))
    kappa = np.exp(np.clip(v[:, :, 3], LOG_KAPPA_LO, LOG_KAPPA_HI))
    logits = v[:, :, 4]
    e = np.exp(logits - logits.max(-1, keepdims=True))
    return mu, kappa, e / e.sum(-1, keepdims=True)


def vmf(kappa, cos_mu):
    return kappa / (2.0 * np.pi * -np.expm1(-2.0 * kappa)) * np.exp(kappa * (cos_mu - 1.0))


def pdf(o, d):
    mu, kappa, w = mixture(o)
    c = np.einsum("nkj,nj->nk", mu, np.asarray(d, np.float64))
    return (w * vmf(kappa, c)).sum(-1)


def loss_grad(o, d, c):
    """Per-sample loss -c log q(d) and dL/do (N, 32)."""
    o = np.asarray(o, np.float64)
    d = np.asarray(d, np.float64)
    mu, kappa, w = mixture(o)
    cosk = np.einsum("nkj,nj->nk", mu, d)
    dens = w * vmf(kappa, cosk)
    q = dens.sum(-1)
    ok = (q > 1e-30) & (c != 0)
    qs = np.where(ok, q, 1.0)
    resp = dens / qs[:, None]
    g = np.zeros_like(o)
    cc = np.where(ok, c, 0.0)[:, None]
    v = o[:, :20].reshape(-1, LOBES, 5)
    g5 = np.zeros((len(o), LOBES, 5))
    g5[:, :, 4] = -cc * (resp - w)
    lk = v[:, :, 3]
    inside = (lk > LOG_KAPPA_LO) & (lk < LOG_KAPPA_HI)
    with np.errstate(over="ignore"):
        t = 2.0 * kappa / np.expm1(2.0 * kappa)
    g5[:, :, 3] = np.where(inside, -cc * resp * (1.0 - t + kappa * (cosk - 1.0)), 0.0)
    vec = v[:, :, :3]
    ln = np.linalg.norm(vec, axis=-1)
    big = ln >= 1e-6
    s = np.where(big, -cc * resp * kappa / np.where(big, ln, 1.0), 0.0)
    g5[:, :, :3] = s[:, :, None] * (d[:, None, :] - cosk[:, :, None] * mu)
    g[:, :20] = g5.reshape(-1, 20)
    loss = np.where(ok, -np.where(ok, c, 0.0) * np.log(qs), 0.0)
    return loss, g


def scales(weights):
    """Per-sample loss weights c = min(w / mean(w), 1024) (invalid: w <= 0 -> 0)."""
    w = np.asarray(weights, np.float64)
    valid = w > 0
    mean = np.where(valid, w, 0.0).sum() / max(len(w), 1)
    if mean <= 0:
        return np.zeros_like(w)
    return np.where(valid, np.minimum(w / mean, MAX_SCALE), 0.0)


def gradient(params, bounds_lo, bounds_hi, samples, fp16=True):
    """Mean over the batch of c_s dL_s/dparams (the library's
    lr_guide_net_gradient), and the mean loss. samples: (N, 10)."""
    samples = np.asarray(samples, np.float64)
    n = len(samples)
    x = encode(bounds_lo, bounds_hi, samples[:, 0:3], samples[:, 3:6])
    c = scales(samples[:, 9])
    ws, _ = unpack(params)
    zs, out = forward(params, x, fp16)
    loss, dz = loss_grad(out, samples[:, 6:9], c)
    grads_w, grads_b = [None] * 4, [None] * 4
    hs = [_half(x, fp16)] + [_half(np.maximum(z, 0.0), fp16) for z in zs]
    for l in range(3, -1, -1):
        grads_w[l] = hs[l].T @ _half(dz, fp16)
        grads_b[l] = dz.sum(0)
        if l > 0:
            dh = _half(dz, fp16) @ _half(ws[l], fp16).T
            dz = np.where(zs[l - 1] > 0, dh, 0.0)
    flat = np.zeros(PARAMS)
    for l in range(4):
        o = LAYER_OUT[l]
        flat[l * IN * HIDDEN: l * IN * HIDDEN + IN * o] = grads_w[l].reshape(-1)
        flat[BIAS + l * HIDDEN: BIAS + l * HIDDEN + o] = grads_b[l]
    return flat / n, float(loss.sum() / n)


def batch_loss(params, bounds_lo, bounds_hi, samples):
    """The f64 mean loss (for finite differences)."""
    samples = np.asarray(samples, np.float64)
    x = encode(bounds_lo, bounds_hi, samples[:, 0:3], samples[:, 3:6])
    c = scales(samples[:, 9])
    _, out = forward(params, x, fp16=False)
    q = pdf(out, samples[:, 6:9])
    ok = (q > 1e-30) & (c != 0)
    return float(np.where(ok, -c * np.log(np.where(ok, q, 1.0)), 0.0).sum() / len(samples))
