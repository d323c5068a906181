"""Learned-proposal oracle pinned by mathematics.

- the backward pass equals central finite differences of the loss;
- the loss (reverse KL to the Boltzmann conditional vMF(u/|u|, |u|)) is zero,
  with zero gradient, when the network outputs exactly that law, and positive
  otherwise;
- the vMF log density integrates to 1 over the sphere; its sampler reproduces
  the mean cosine A(k) = coth k - 1/k and is isotropic about the axis.
"""
import math

import numpy as np
import pytest

from oracle import oracle

rng = np.random.default_rng(42)


def envs(n, beta_scale=3.0):
    u = rng.normal(size=(n, 3)) * beta_scale
    s = rng.normal(size=(n, 3))
    s /= np.linalg.norm(s, axis=1, keepdims=True)
    return np.concatenate([u, s], axis=1)


def test_param_count():
    assert oracle.mlp_nparams() == 6 * 64 + 64 + 64 * 64 + 64 + 4 * 64 + 4 == 4868


def test_gradient_matches_finite_differences():
    w = oracle.mlp_init(1)
    w[-1] = 0.5  # kappa away from the clamp
    e = envs(16)
    loss, g = oracle.mlp_loss_grad(w, e)
    idx = np.concatenate([rng.choice(384, 20, replace=False),            # W1
                          384 + rng.choice(64, 5, replace=False),        # b1
                          448 + rng.choice(4096, 30, replace=False),     # W2
                          4544 + rng.choice(64, 5, replace=False),       # b2
                          4608 + rng.choice(256, 20, replace=False),     # W3
                          [4864, 4865, 4866, 4867]])                     # b3
    for p in idx:
        h = 1e-6
        wp, wm = w.copy(), w.copy()
        wp[p] += h
        wm[p] -= h
        fd = (oracle.mlp_loss_grad(wp, e, False)[0] - oracle.mlp_loss_grad(wm, e, False)[0]) / (2 * h)
        assert abs(fd - g[p]) <= 1e-6 + 1e-4 * abs(fd), (p, fd, g[p])


def exact_net_for(u):
    """A network whose output is (u/|u|, log|u|) for every input: W3 = 0, b3 set."""
    w = oracle.mlp_init(2)
    w[4608:4864] = 0.0
    r = np.linalg.norm(u)
    w[4864:4867] = u / r
    w[4867] = math.log(r)
    return w


@pytest.mark.parametrize("scale", [0.05, 1.0, 40.0])
def test_loss_zero_at_heat_bath(scale):
    u = rng.normal(size=3) * scale
    s = np.array([0.0, 0.0, 1.0])
    w = exact_net_for(u)
    loss, g = oracle.mlp_loss_grad(w, np.concatenate([u, s])[None, :])
    assert abs(loss) < 1e-9
    assert np.abs(g[4864:]).max() < 1e-7
    w2 = w.copy()
    w2[4867] += 0.3
    assert oracle.mlp_loss_grad(w2, np.concatenate([u, s])[None, :], False)[0] > 1e-4
    w3 = w.copy()
    w3[4864] += 0.5
    assert oracle.mlp_loss_grad(w3, np.concatenate([u, s])[None, :], False)[0] > 1e-6


@pytest.mark.parametrize("k", [1e-3, 0.7, 5.0, 300.0])
def test_vmf_density_normalised(k):
    mu = np.array([0.0, 0.6, 0.8])
    # integrate over the cosine to the axis: 2 pi int_{-1}^{1} q(c) dc
    c = np.linspace(-1, 1, 200001)
    lq = np.array([oracle.vmf_log_norm(k)]) + k * c
    integral = 2 * np.pi * np.trapezoid(np.exp(lq), c)
    assert abs(integral - 1) < 1e-6
    y = np.array([0.0, 0.0, 1.0])
    assert oracle.vmf_log_density(mu, k, y) == pytest.approx(oracle.vmf_log_norm(k) + 0.8 * k)


@pytest.mark.parametrize("k", [0.01, 2.0, 50.0])
def test_vmf_sampler_moments(k):
    mu = rng.normal(size=3)
    mu /= np.linalg.norm(mu)
    if k == 50.0:
        mu = np.array([0.0, 0.0, -1.0])  # exercise the other reflection pole
    n = 20000
    u = rng.random((n, 2))
    ys = np.array([oracle.vmf_sample(mu, k, a, b) for a, b in u])
    assert np.allclose(np.linalg.norm(ys, axis=1), 1, atol=1e-12)
    c = ys @ mu
    A = 1 / math.tanh(k) - 1 / k
    var_c = 1 - 2 * A / k - A * A  # variance of the cosine under vMF on S^2
    assert abs(c.mean() - A) < 5 * math.sqrt(max(var_c, 1e-12) / n) + 1e-9
    perp = ys - np.outer(c, mu)
    assert np.linalg.norm(perp.mean(axis=0)) < 6 * math.sqrt((1 - A * A + 1e-9) / n) + 1e-9


def test_features():
    u = np.array([3.0, 0.0, 4.0])
    s = np.array([0.0, 1.0, 0.0])
    x = oracle.mlp_features(u, s)
    assert np.allclose(x[:3], u / 5 * math.log1p(5))
    assert np.allclose(x[3:], s)
    assert np.allclose(oracle.mlp_features(np.zeros(3), s)[:3], 0)
