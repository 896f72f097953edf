"""Oracle (test infrastructure only): FP64 reference of the path-guiding MLP.

The reference crate has no path guiding; this module is the specification
the HIP kernels are checked against. It is pinned by tests/test_guide_oracle.py
(finite differences for every gradient, quadrature for the vMF normalisation,
the closed-form Adam step, sampling histograms, bfloat16 rounding on exactly
representable and tie values, and the bfloat16 variants reducing to the FP64
ones when the inputs are already bfloat16 and the network is linear enough).

Network: a 64-dim encoding of (position, outgoing direction, normal) ->
4 hidden layers of 64 ReLU units -> 40 outputs = 8 von Mises-Fisher lobes x
(mixture logit, 3-vector direction, concentration parameter s with
log kappa = 1.5 + 5.5 tanh((s - 1.5) / 5.5), i.e. kappa in (e^-4, e^7)).
"""
from __future__ import annotations

import numpy as np

N_IN = 64
N_HID = 64
N_LOBES = 8
N_OUT = 5 * N_LOBES
LAYERS = [(N_HID, N_IN), (N_HID, N_HID), (N_HID, N_HID), (N_HID, N_HID), (N_OUT, N_HID)]
N_PARAMS = sum(o * i + o for o, i in LAYERS)  # 19,240
LOG_KAPPA_MIN, LOG_KAPPA_MAX = -4.0, 7.0
# log kappa = C + R tanh((s - C) / R): the identity near C, saturating smoothly at (MIN, MAX), so no
# lobe's concentration is ever stuck where its gradient vanishes
LOG_KAPPA_C = 0.5 * (LOG_KAPPA_MIN + LOG_KAPPA_MAX)
LOG_KAPPA_R = 0.5 * (LOG_KAPPA_MAX - LOG_KAPPA_MIN)
N_FREQ = 8


def encode(p, wo, n, bmin, bmax):
    """(S, 64) encoding: for each position axis i and k < 8, sin and cos of 2^k * pi * q_i with q the
    position normalised to the scene box (48), then wo (3), n (3), sin and cos of pi * wo_i (6), zeros (4)."""
    p = np.asarray(p, np.float64)
    q = (p - np.asarray(bmin, np.float64)) / (np.asarray(bmax, np.float64) - np.asarray(bmin, np.float64))
    S = p.shape[0]
    e = np.zeros((S, N_IN), np.float64)
    for i in range(3):
        for k in range(N_FREQ):
            a = (2.0 ** k) * np.pi * q[:, i]
            e[:, 2 * (N_FREQ * i + k)] = np.sin(a)
            e[:, 2 * (N_FREQ * i + k) + 1] = np.cos(a)
    e[:, 48:51] = wo
    e[:, 51:54] = n
    for i in range(3):
        e[:, 54 + 2 * i] = np.sin(np.pi * np.asarray(wo)[:, i])
        e[:, 55 + 2 * i] = np.cos(np.pi * np.asarray(wo)[:, i])
    return e


def unpack(flat):
    """Flat parameter vector -> [(W (out, in), b (out,))] in layer order (W row-major, then b)."""
    out, o = [], 0
    for no, ni in LAYERS:
        W = flat[o:o + no * ni].reshape(no, ni)
        o += no * ni
        b = flat[o:o + no]
        o += no
        out.append((W, b))
    return out


def forward(flat, X):
    """Returns (out (S, 40), cache)."""
    layers = unpack(np.asarray(flat, np.float64))
    h = np.asarray(X, np.float64)
    acts = [h]
    pre = []
    for li, (W, b) in enumerate(layers):
        z = h @ W.T + b
        pre.append(z)
        h = np.maximum(z, 0.0) if li < len(layers) - 1 else z
        acts.append(h)
    return h, (acts, pre)


def mixture(out):
    """out (S, 40) -> (pi (S, 8), mu (S, 8, 3), kappa (S, 8), norm |m| (S, 8))."""
    o = out.reshape(-1, N_LOBES, 5)
    a = o[..., 0]
    a = a - a.max(axis=1, keepdims=True)
    pi = np.exp(a)
    pi /= pi.sum(axis=1, keepdims=True)
    m = o[..., 1:4]
    nm = np.sqrt((m * m).sum(-1) + 1e-12)
    mu = m / nm[..., None]
    s = LOG_KAPPA_C + LOG_KAPPA_R * np.tanh((o[..., 4] - LOG_KAPPA_C) / LOG_KAPPA_R)
    kappa = np.exp(s)
    return pi, mu, kappa, nm


def vmf_pdf(mu, kappa, w):
    """kappa / (2 pi (1 - e^{-2 kappa})) * exp(kappa (mu.w - 1)); mu (..., 3), kappa (...), w (..., 3)."""
    c = kappa / (2.0 * np.pi * (-np.expm1(-2.0 * kappa)))
    return c * np.exp(kappa * ((mu * w).sum(-1) - 1.0))


def mixture_pdf(out, w):
    pi, mu, kappa, _ = mixture(out)
    return (pi * vmf_pdf(mu, kappa, np.asarray(w)[:, None, :])).sum(1)


def loss(out, w, weight):
    """Batch mean of -weight * log p(w)."""
    return float(np.mean(-np.asarray(weight) * np.log(mixture_pdf(out, w))))


def loss_grad_out(out, w, weight):
    """d loss / d out for loss(): (S, 40)."""
    S = out.shape[0]
    o = out.reshape(S, N_LOBES, 5)
    pi, mu, kappa, nm = mixture(out)
    w = np.asarray(w, np.float64)
    f = vmf_pdf(mu, kappa, w[:, None, :])
    p = (pi * f).sum(1, keepdims=True)
    r = pi * f / p                                     # responsibilities
    dlogp = np.zeros_like(o)
    dlogp[..., 0] = r - pi                             # softmax logits
    muw = (mu * w[:, None, :]).sum(-1)
    dlogk = 1.0 / kappa - 2.0 * np.exp(-2.0 * kappa) / (-np.expm1(-2.0 * kappa)) + (muw - 1.0)
    th = np.tanh((o[..., 4] - LOG_KAPPA_C) / LOG_KAPPA_R)
    dlogp[..., 4] = r * dlogk * kappa * (1.0 - th * th)  # kappa = exp(C + R tanh((s - C) / R))
    dmu = kappa[..., None] * w[:, None, :]             # d log f / d mu
    proj = dmu - (dmu * mu).sum(-1, keepdims=True) * mu
    dlogp[..., 1:4] = r[..., None] * proj / nm[..., None]
    g = -np.asarray(weight, np.float64)[:, None, None] * dlogp / S
    return g.reshape(S, N_OUT)


def backward(flat, cache, dout):
    """Gradient of the loss w.r.t. the flat parameters given d loss / d out."""
    acts, pre = cache
    layers = unpack(np.asarray(flat, np.float64))
    grads = []
    d = dout
    for li in range(len(layers) - 1, -1, -1):
        W, _ = layers[li]
        if li < len(layers) - 1:
            d = d * (pre[li] > 0)
        gW = d.T @ acts[li]
        gb = d.sum(0)
        grads.append((gW, gb))
        d = d @ W
    grads.reverse()
    return np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])


def bf16(x):
    """Rounds to the nearest bfloat16 (ties to even), returned as float64. NaN stays NaN."""
    f = np.asarray(x, np.float64).astype(np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32).astype(np.float64)
    return np.where(np.isnan(f), np.nan, out)


def forward_bf16(flat, X):
    """forward() with the rounding points of the matrix-core kernels: the encoding, every weight
    and every hidden activation (after bias and ReLU) are rounded to bfloat16; products and
    biases accumulate exactly (the kernels accumulate in FP32); outputs are not rounded."""
    layers = unpack(np.asarray(flat, np.float64))
    h = bf16(X)
    acts, pre = [h], []
    for li, (W, b) in enumerate(layers):
        z = h @ bf16(W).T + b
        pre.append(z)
        h = bf16(np.maximum(z, 0.0)) if li < len(layers) - 1 else z
        acts.append(h)
    return h, (acts, pre)


def backward_bf16(flat, cache, dout):
    """backward() with the kernels' rounding points: d out and every back-propagated pre-activation
    gradient are rounded to bfloat16, products use the bfloat16 weights, the ReLU mask is taken
    from the rounded activations (X_l > 0), and dW, db accumulate exactly."""
    acts, _ = cache
    layers = unpack(np.asarray(flat, np.float64))
    grads = []
    d = bf16(dout)
    for li in range(len(layers) - 1, -1, -1):
        W, _ = layers[li]
        grads.append((d.T @ acts[li], d.sum(0)))
        if li > 0:
            d = bf16((d @ bf16(W)) * (acts[li] > 0))
    grads.reverse()
    return np.concatenate([np.concatenate([gW.ravel(), gb]) for gW, gb in grads])


def adam(flat, grad, m, v, t, lr=1e-2, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step (Kingma & Ba 2015, bias-corrected); t is the 1-based step."""
    m = b1 * m + (1 - b1) * grad
    v = b2 * v + (1 - b2) * grad * grad
    mh = m / (1 - b1 ** t)
    vh = v / (1 - b2 ** t)
    return flat - lr * mh / (np.sqrt(vh) + eps), m, v


def init_params(seed: int) -> np.ndarray:
    """He-uniform weights and zero biases from a seeded numpy generator (the library accepts any
    initial parameters; tests pass these)."""
    rng = np.random.default_rng(seed)
    parts = []
    for no, ni in LAYERS:
        lim = np.sqrt(6.0 / ni)
        parts.append(rng.uniform(-lim, lim, no * ni))
        parts.append(np.zeros(no))
    return np.concatenate(parts)


def frame(mu):
    """Orthonormal (s, t) around each mu (vec3.rs:25-33 coordinate_system)."""
    s = np.where(np.abs(mu[:, 0:1]) > np.abs(mu[:, 1:2]),
                 np.stack([-mu[:, 2], np.zeros(len(mu)), mu[:, 0]], 1) / np.sqrt(mu[:, 0:1] ** 2 + mu[:, 2:3] ** 2),
                 np.stack([np.zeros(len(mu)), mu[:, 2], -mu[:, 1]], 1) / np.sqrt(mu[:, 1:2] ** 2 + mu[:, 2:3] ** 2))
    t = np.cross(mu, s)
    return s, t


def sample_vmf(mu, kappa, u1, u2):
    """Inverse-CDF vMF sample: cos = 1 + log(u1 + (1 - u1) e^{-2 kappa}) / kappa, phi = 2 pi u2."""
    c = 1.0 + np.log(u1 + (1.0 - u1) * np.exp(-2.0 * kappa)) / kappa
    c = np.clip(c, -1.0, 1.0)
    sn = np.sqrt(np.maximum(0.0, 1.0 - c * c))
    phi = 2.0 * np.pi * u2
    s, t = frame(mu)
    return (sn * np.cos(phi))[:, None] * s + (sn * np.sin(phi))[:, None] * t + c[:, None] * mu
