"""Pins the FP64 guiding-MLP oracle (oracle/guide_mlp.py): finite differences
for every gradient, quadrature for the vMF normalisation, the closed-form
Adam step, and sample histograms against the pdf."""
import sys
from pathlib import Path

import numpy as np

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
import guide_mlp as G  # noqa: E402


def fib_sphere(n):
    i = np.arange(n) + 0.5
    z = 1 - 2 * i / n
    r = np.sqrt(1 - z * z)
    phi = np.pi * (1 + 5 ** 0.5) * i
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], 1)


def test_param_count():
    assert G.N_PARAMS == 19240  # 4 * 64^2 + 64 * 40 + biases


def test_vmf_normalises():
    w = fib_sphere(200_000)
    for kappa in (0.05, 1.0, 10.0, 200.0):
        mu = np.array([0.3, -0.5, 0.81])
        mu /= np.linalg.norm(mu)
        p = G.vmf_pdf(mu[None], np.array([kappa]), w)
        integral = p.mean() * 4 * np.pi
        assert abs(integral - 1) < 2e-3, (kappa, integral)


def test_mixture_normalises_and_samples_match_pdf():
    rng = np.random.default_rng(1)
    out = rng.standard_normal((1, 40))
    w = fib_sphere(100_000)
    p = G.mixture_pdf(np.repeat(out, len(w), 0), w)
    assert abs(p.mean() * 4 * np.pi - 1) < 2e-3
    # sample: pick lobes by pi, then vMF; histogram over z vs the pdf's z-marginal
    pi, mu, kappa, _ = G.mixture(out)
    n = 200_000
    lobes = rng.choice(8, n, p=pi[0])
    d = G.sample_vmf(mu[0, lobes], kappa[0, lobes], rng.random(n), rng.random(n))
    np.testing.assert_allclose(np.linalg.norm(d, axis=1), 1, atol=1e-9)
    bins = np.linspace(-1, 1, 11)
    h, _ = np.histogram(d[:, 2], bins)
    ph, _ = np.histogram(w[:, 2], bins, weights=p * 4 * np.pi / len(w))
    np.testing.assert_allclose(h / n, ph, atol=6e-3)


def numeric_grad(f, x, idx, h=1e-6):
    g = np.zeros(len(idx))
    for j, i in enumerate(idx):
        xp = x.copy()
        xm = x.copy()
        xp[i] += h
        xm[i] -= h
        g[j] = (f(xp) - f(xm)) / (2 * h)
    return g


def batch(seed, S=16):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((S, 64))
    w = rng.standard_normal((S, 3))
    w /= np.linalg.norm(w, axis=1, keepdims=True)
    weight = rng.uniform(0.1, 2.0, S)
    return X, w, weight


def test_output_gradient_matches_finite_differences():
    rng = np.random.default_rng(2)
    out = rng.standard_normal((6, 40)) * 0.7
    _, w, weight = batch(3, 6)
    g = G.loss_grad_out(out, w, weight)
    f = lambda flat: G.loss(flat.reshape(6, 40), w, weight)  # noqa: E731
    idx = np.arange(240)
    num = numeric_grad(f, out.ravel(), idx)
    np.testing.assert_allclose(g.ravel()[idx], num, rtol=1e-5, atol=1e-8)


def test_parameter_gradient_matches_finite_differences():
    flat = G.init_params(4)
    X, w, weight = batch(5)
    out, cache = G.forward(flat, X)
    g = G.backward(flat, cache, G.loss_grad_out(out, w, weight))

    def f(p):
        return G.loss(G.forward(p, X)[0], w, weight)

    rng = np.random.default_rng(6)
    idx = np.concatenate([rng.choice(G.N_PARAMS, 60, replace=False),
                          [0, 4095, 4096 + 63, G.N_PARAMS - 1, G.N_PARAMS - 41]])
    num = numeric_grad(f, flat, idx)
    np.testing.assert_allclose(g[idx], num, rtol=1e-4, atol=1e-8)


def test_adam_closed_form():
    # first step with bias correction moves every parameter by lr * sign(g) (up to eps)
    p = np.zeros(5)
    g = np.array([1.0, -2.0, 0.5, 1e-3, -4.0])
    p1, m, v = G.adam(p, g, np.zeros(5), np.zeros(5), 1, lr=1e-2)
    np.testing.assert_allclose(p1, -1e-2 * np.sign(g), rtol=1e-4)
    np.testing.assert_allclose(m, 0.1 * g)
    np.testing.assert_allclose(v, 0.001 * g * g)


def test_training_reduces_loss_on_a_fixed_target():
    # a single bright direction per position half: training must concentrate the mixture there
    rng = np.random.default_rng(7)
    S = 256
    p = rng.random((S, 3))
    wo = np.tile([0.0, 0.0, 1.0], (S, 1))
    n = np.tile([0.0, 1.0, 0.0], (S, 1))
    X = G.encode(p, wo, n, np.zeros(3), np.ones(3))
    target = np.where(p[:, :1] < 0.5, [[0.0, 1.0, 0.0]], [[1.0, 0.0, 0.0]])
    flat = G.init_params(8)
    m = np.zeros_like(flat)
    v = np.zeros_like(flat)
    first = None
    for t in range(1, 201):
        wdir = target + 0.05 * rng.standard_normal((S, 3))
        wdir /= np.linalg.norm(wdir, axis=1, keepdims=True)
        out, cache = G.forward(flat, X)
        lval = G.loss(out, wdir, np.ones(S))
        first = lval if first is None else first
        g = G.backward(flat, cache, G.loss_grad_out(out, wdir, np.ones(S)))
        flat, m, v = G.adam(flat, g, m, v, t, lr=1e-2)
    out, _ = G.forward(flat, X)
    assert G.loss(out, target, np.ones(S)) < first - 2.0


def test_bf16_rounding():
    e7 = 2.0 ** -7
    vals = np.array([1.0, 1 + e7 / 2, 1 + 3 * e7 / 2, 1 + e7 / 2 + 2.0 ** -20, -(1 + 3 * e7 / 2), 3.0e38, 1e-40, 0.0])
    exp = np.array([1.0, 1.0, 1 + 2 * e7, 1 + e7, -(1 + 2 * e7), 2.9999e38, None, 0.0], dtype=object)
    got = G.bf16(vals)
    for g, e in zip(got[:5], exp[:5]):
        assert g == e
    assert abs(got[5] / 3.0e38 - 1) < 2.0 ** -8
    assert got[7] == 0.0 and np.isnan(G.bf16(np.nan))
    rng = np.random.default_rng(0)
    x = rng.standard_normal(10000) * 10.0 ** rng.uniform(-20, 20, 10000)
    r = G.bf16(x)
    assert np.all(np.abs(r / x - 1) <= 2.0 ** -8)
    np.testing.assert_array_equal(G.bf16(r), r)  # idempotent


def test_bf16_variants_reduce_to_fp64_without_rounding(monkeypatch):
    flat = G.init_params(4)
    X, w, weight = batch(5, 24)
    out, cache = G.forward(flat, X)
    ref = G.backward(flat, cache, G.loss_grad_out(out, w, weight))
    monkeypatch.setattr(G, "bf16", lambda x: np.asarray(x, np.float64))
    out2, cache2 = G.forward_bf16(flat, X)
    np.testing.assert_allclose(out2, out, rtol=1e-12, atol=1e-12)
    got = G.backward_bf16(flat, cache2, G.loss_grad_out(out2, w, weight))
    np.testing.assert_allclose(got, ref, rtol=1e-10, atol=1e-14)
