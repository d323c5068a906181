"""Oracle arithmetic pinned against libm and against closed forms.

exp and sin/cos(pi x) are the oracle's own polynomial evaluations (the GPU
evaluates the same sequence); here they are held to the double-precision
library functions. The proposal, the Metropolis test and over-relaxation are
held to what the mathematics fixes: cap uniformity, unit length, involution,
energy conservation, and the Langevin law of a single spin in a field.
"""
import math

import numpy as np
import pytest

from oracle import oracle

rng = np.random.default_rng(1234)


def ulp_err(got, ref):
    return abs(got - ref) / np.spacing(np.float32(abs(ref)))


def test_exp_neg_against_libm():
    xs = np.concatenate([np.linspace(-85.9, 0.0, 4001), -rng.random(4000) * 20,
                         [-1e-8, -1e-30, 0.0, -0.34657, -0.34658, -0.6931472]])
    worst = 0.0
    for x in xs.astype(np.float32):
        got = oracle.exp_neg(float(x))
        ref = math.exp(float(x))
        worst = max(worst, ulp_err(got, ref))
    assert worst <= 2.0, worst


def test_exp_neg_limits():
    assert oracle.exp_neg(0.0) == 1.0
    assert oracle.exp_neg(-86.5) == 0.0
    assert oracle.exp_neg(-1000.0) == 0.0
    assert oracle.exp_neg(float("-inf")) == 0.0
    # Monotone on a fine grid: the Metropolis test relies on it.
    xs = np.linspace(-20, 0, 20001).astype(np.float32)
    ys = np.array([oracle.exp_neg(float(x)) for x in xs])
    assert np.all(np.diff(ys) >= 0)


def test_sincospi_against_libm():
    As = np.concatenate([np.linspace(0, 2, 8001), rng.random(4000) * 2,
                         [0.25, 0.5, 0.75, 1.0, 1.25, 1.5, 1.75, 2.0]]).astype(np.float32)
    for a in As:
        s, c = oracle.sincospi(float(a))
        assert abs(s - math.sin(math.pi * float(a))) < 4e-7
        assert abs(c - math.cos(math.pi * float(a))) < 4e-7
    s, c = oracle.sincospi(0.5)
    assert (s, c) == (1.0, 0.0) or abs(c) < 1e-7


def random_unit(n):
    v = rng.normal(size=(n, 3))
    return (v / np.linalg.norm(v, axis=1, keepdims=True)).astype(np.float32)


@pytest.mark.parametrize("cos_cone", [-1.0, 0.0, 0.7, 0.99])
def test_cone_proposal_uniform_on_cap(cos_cone):
    s = random_unit(1)[0]
    n = 20000
    u = rng.random((n, 2)).astype(np.float32)
    out = np.array([oracle.cone_proposal(s, cos_cone, a, b) for a, b in u])
    norms = np.linalg.norm(out.astype(np.float64), axis=1)
    assert np.all(np.abs(norms - 1) < 3e-7)
    c = out.astype(np.float64) @ s.astype(np.float64)
    assert c.min() >= cos_cone - 1e-6
    # cos(theta) uniform on [cos_cone, 1]: mean and variance of a uniform law.
    lo = cos_cone
    assert abs(c.mean() - (1 + lo) / 2) < 5 * (1 - lo) / np.sqrt(12 * n)
    assert abs(c.var() - (1 - lo) ** 2 / 12) < 0.05 * (1 - lo) ** 2 / 12 + 1e-9
    # Isotropic about s: the mean transverse component vanishes.
    perp = out - np.outer(c, s)
    assert np.linalg.norm(perp.mean(axis=0)) < 6 * np.sqrt((1 - lo) / n) + 1e-6


def test_cone_proposal_at_south_pole():
    """The reference's s_z < -0.99995 branch is wrong (Math.cu:254-260); ours is exact."""
    s = np.array([0, 0, -1], np.float32)
    for u in rng.random((500, 2)).astype(np.float32):
        o = oracle.cone_proposal(s, 0.9, u[0], u[1])
        assert float(o @ s) >= 0.9 - 1e-6


def test_overrelax_properties():
    for _ in range(2000):
        h = rng.normal(size=3).astype(np.float32) * rng.uniform(0.1, 6)
        s = random_unit(1)[0]
        sp = oracle.overrelax(h, s)
        assert abs(np.linalg.norm(sp.astype(np.float64)) - 1) < 3e-7
        # Energy -h.s conserved; the component along h is kept, the rest reversed.
        assert abs(float(h @ sp) - float(h @ s)) < 2e-6 * np.linalg.norm(h)
        spp = oracle.overrelax(h, sp)
        assert np.allclose(spp, s, atol=2e-6)
    s = random_unit(1)[0]
    assert np.array_equal(oracle.overrelax(np.zeros(3, np.float32), s), s)


def test_metropolis_accept_rule():
    h = np.array([0, 0, 2.0], np.float32)
    up = np.array([0, 0, 1.0], np.float32)
    dn = np.array([0, 0, -1.0], np.float32)
    # Downhill (lower local energy -h.s) is always accepted.
    assert oracle.metropolis_accept(h, dn, up, 1.0, 0.999999)
    # Uphill by dE = 4 at beta = 0.5: accepted iff u < e^-2.
    assert oracle.metropolis_accept(h, up, dn, 0.5, 0.1353)
    assert not oracle.metropolis_accept(h, up, dn, 0.5, 0.1354)


def test_single_spin_langevin_law():
    """Metropolis with the cone proposal in a fixed field h samples
    p(s) ~ exp(beta h.s): <s.h^> = coth(k) - 1/k with k = beta |h|."""
    h = np.array([0.3, -0.4, 1.2], np.float32)
    beta = np.float32(1.5)
    k = float(beta) * float(np.linalg.norm(h))
    expect = 1 / math.tanh(k) - 1 / k
    s = np.array([1, 0, 0], np.float32)
    hh = h / np.linalg.norm(h)
    u = rng.random((60000, 3)).astype(np.float32)
    vals = []
    for i, (a, b, c) in enumerate(u):
        sp = oracle.cone_proposal(s, 0.3, a, b)
        if oracle.metropolis_accept(h, s, sp, beta, c):
            s = sp
        if i >= 2000:
            vals.append(float(s @ hh))
    vals = np.array(vals)
    # Batch means for the error bar.
    bm = vals[: len(vals) // 100 * 100].reshape(100, -1).mean(axis=1)
    err = bm.std(ddof=1) / np.sqrt(len(bm))
    assert abs(vals.mean() - expect) < 4 * err + 1e-3, (vals.mean(), expect, err)


def test_renormalise_near_unit_vectors():
    """The near-unit renormalisation (one Newton step for 1/sqrt from 1) used
    after every reflection and rotation: on vectors whose squared length is
    within 1e-5 of 1 the result has unit length to 2.5e-7 and the direction
    of the double-precision normalisation to 2e-7 (a few FP32 ulp), with no
    systematic bias of the length (mean deviation below 2e-9), also for the
    few-ulp deviations that rounding leaves."""
    rng = np.random.default_rng(11)
    v = rng.normal(size=(20000, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    v *= np.sqrt(1.0 + rng.uniform(-1e-5, 1e-5, size=(20000, 1)))
    v = v.astype(np.float32)
    out = np.array([oracle.renormalise(x) for x in v]).astype(np.float64)
    n = np.linalg.norm(out, axis=1)
    exact = v.astype(np.float64) / np.linalg.norm(v.astype(np.float64), axis=1, keepdims=True)
    assert np.abs(n - 1).max() < 2.5e-7
    assert np.abs(out - exact).max() < 2e-7
    assert abs((n - 1).mean()) < 2e-9
    # at the deviations rounding leaves (a few ulp) the length is unbiased too
    w = rng.normal(size=(20000, 3))
    w /= np.linalg.norm(w, axis=1, keepdims=True)
    w = (w * (1 + rng.uniform(-3e-7, 3e-7, size=(20000, 1)))).astype(np.float32)
    m = np.linalg.norm(np.array([oracle.renormalise(x) for x in w]).astype(np.float64), axis=1)
    assert abs((m - 1).mean()) < 1e-8 and np.abs(m - 1).max() < 1.5e-7
