"""GPU tests of the learned local proposal, through the C ABI.

- the BF16-MFMA forward pass matches the FP64 oracle forward (BF16 tolerance);
- the Metropolis-Hastings log ratio computed on the device equals the
  closed-form vMF ratio evaluated in FP64 on the device's own outputs;
- FP32 training gradients match the FP64 oracle gradients;
- the device vMF sampler reproduces the mean cosine coth k - 1/k;
- training lowers the reverse KL and raises the acceptance;
- exactness: learned sweeps, with a trained or an untrained network, sample
  the same equilibrium energy as the plain Metropolis sweeps.
"""
import math
import sys

import numpy as np
import pytest

from oracle import oracle

pytestmark = pytest.mark.gpu

rng = np.random.default_rng(7)


def say(msg):
    print(f"[gpu-learned] {msg}", flush=True)


def ctx(L=4, temps=(1.0,), **kw):
    from cudacode_amd import pyro
    say(f"create L={L} R={len(temps)} {kw}")
    return pyro.Pyrochlore(L, list(temps), seed=kw.pop("seed", 99), **kw)


def random_weights(seed=3, out_scale=1.0):
    w = oracle.mlp_init(seed)
    w[4608:4864] *= out_scale
    return w


def env_rows(n, scale=3.0):
    u = rng.normal(size=(n, 3)) * scale
    s = rng.normal(size=(n, 3))
    s /= np.linalg.norm(s, axis=1, keepdims=True)
    return np.concatenate([u, s], axis=1).astype(np.float32)


def test_mfma_forward_matches_oracle():
    w = random_weights().astype(np.float32)
    e = env_rows(256)
    sp = env_rows(256)[:, 3:]
    rows = np.concatenate([e, sp], axis=1)
    with ctx() as g:
        g.mlp_set_weights(w)
        out = g.mlp_eval(rows)
    ref1 = np.array([oracle.mlp_forward(w.astype(np.float64),
                                        oracle.mlp_features(r[:3].astype(np.float64), r[3:6].astype(np.float64)))
                     for r in rows])
    ref2 = np.array([oracle.mlp_forward(w.astype(np.float64),
                                        oracle.mlp_features(r[:3].astype(np.float64), r[6:9].astype(np.float64)))
                     for r in rows])
    scale = np.sqrt(np.mean(ref1 ** 2))
    assert np.abs(out[:, :4] - ref1).max() < 0.05 * scale, np.abs(out[:, :4] - ref1).max() / scale
    assert np.abs(out[:, 4:8] - ref2).max() < 0.05 * scale
    # every lane position gives the same answer for the same input
    same = np.repeat(rows[:1], 128, axis=0)
    with ctx() as g:
        g.mlp_set_weights(w)
        o = g.mlp_eval(same)
    assert np.all(o == o[0])


def host_log_ratio(o1, o2, u, s, sp):
    def dec(o):
        o = o.astype(np.float64)
        n = np.linalg.norm(o[:3])
        mu = o[:3] / n
        k = math.exp(min(max(o[3], math.log(0.01)), math.log(4096)))
        return mu, k

    mu1, k1 = dec(o1)
    mu2, k2 = dec(o2)
    u, s, sp = (np.asarray(v, np.float64) for v in (u, s, sp))
    return (u @ sp - u @ s) + oracle.vmf_log_density(mu2, k2, s) - oracle.vmf_log_density(mu1, k1, sp)


def test_mh_log_ratio_formula():
    w = random_weights(out_scale=0.3).astype(np.float32)
    rows = np.concatenate([env_rows(200, 1.0), env_rows(200)[:, 3:]], axis=1)
    with ctx() as g:
        g.mlp_set_weights(w)
        out = g.mlp_eval(rows)
    for r, o in zip(rows, out):
        ref = host_log_ratio(o[:4], o[4:8], r[:3], r[3:6], r[6:9])
        k = max(math.exp(min(o[3], 8.3)), math.exp(min(o[7], 8.3)))
        assert abs(o[8] - ref) < 1e-4 * (1 + k + abs(ref)), (o[8], ref)


def test_training_gradient_matches_oracle():
    w = random_weights(5).astype(np.float32)
    w[-1] = 0.3
    e = env_rows(512)
    with ctx() as g:
        g.mlp_set_weights(w)
        loss, grad = g.mlp_loss_grad(e)
    ref_loss, ref_grad = oracle.mlp_loss_grad(w.astype(np.float64), e.astype(np.float64))
    assert loss == pytest.approx(ref_loss, rel=1e-5)
    scale = np.abs(ref_grad).max()
    assert np.abs(grad - ref_grad).max() < 1e-4 * scale, np.abs(grad - ref_grad).max() / scale


@pytest.mark.parametrize("k", [0.05, 3.0, 200.0])
def test_device_vmf_sampler(k):
    n = 40000
    mu = np.array([0.3, -0.5, 0.8])
    mu /= np.linalg.norm(mu)
    if k == 200.0:
        mu = -mu
    rows = np.tile(np.concatenate([mu, [k]]).astype(np.float32), (n, 1))
    with ctx() as g:
        ys = g.vmf_draw(rows).astype(np.float64)
    c = ys @ mu
    A = 1 / math.tanh(k) - 1 / k
    var_c = max(1 - 2 * A / k - A * A, 1e-12)
    assert abs(c.mean() - A) < 5 * math.sqrt(var_c / n) + 2e-6
    assert np.abs(np.linalg.norm(ys, axis=1) - 1).max() < 1e-6


def equilibrium_energy(g, sweeps):
    es = []
    for _ in range(sweeps):
        g.Mc(1, use_graph=False)
        es.append(g.TotalEnergyv2())
    return np.array(es) / g.N  # (sweeps, R)


def replica_mean(es, burn):
    m = es[burn:].mean(axis=0)  # one mean per independent replica
    return m.mean(), m.std(ddof=1) / math.sqrt(len(m))


@pytest.mark.parametrize("trained,L", [(False, 4), (True, 4), (True, 6)])
def test_learned_sweeps_sample_the_boltzmann_law(trained, L):
    """16 independent replicas at T = 0.35: the mean energy of learned-only
    sweeps equals that of plain Metropolis sweeps within 4 sigma. With an
    untrained network the acceptance is low but the law is the same: that is
    the exactness of the Metropolis-Hastings correction. L=6 (216 cells) runs
    the runtime-size kernel with a partly filled last wave of cells."""
    T, R = 0.35, 16
    with ctx(L, [T] * R, n_metro=1, n_or=0, cos_cone=0.5, seed=11) as plain:
        say("plain reference")
        es_p = equilibrium_energy(plain, 300)
    with ctx(L, [T] * R, n_metro=0, n_or=0, seed=12) as g:
        if trained:
            g.mlp_set_weights(random_weights(9).astype(np.float32))
            g.set_passes(1, 0)
            g.Mc(100)
            g.gather_environments(2048, 1, 200000)
            g.Mc(20)
            g.gather_environments(2048, 2, 200000)
            hist = g.mlp_train(300, batch=2048, lr=3e-3, lr_final=3e-4)
            assert hist[-20:].mean() < 0.25 * hist[:5].mean()
            g.set_passes(0, 0)
        else:
            g.mlp_set_weights(random_weights(10, out_scale=0.2).astype(np.float32))
        g.set_learned_passes(1)
        say(f"learned sweeps (trained={trained})")
        es_l = equilibrium_energy(g, 300)
        a, p = g.acceptance_learned()
        acc = a.sum() / p.sum()
    mp, ep = replica_mean(es_p, 100)
    ml, el = replica_mean(es_l, 100)
    say(f"E/N plain {mp:.5f}+-{ep:.5f} learned {ml:.5f}+-{el:.5f} acceptance {acc:.3f}")
    assert abs(mp - ml) < 4 * math.hypot(ep, el)
    if trained:
        assert acc > 0.8
    else:
        assert 0.0 < acc < 0.9


def test_learned_then_untracked_overrelaxation_keeps_the_energy_books():
    """Learned proposals followed by over-relaxation passes without FP64
    bookkeeping (the fast default) and no exchanges (which re-anchor): the
    incremental energy stays within FP32 rounding of the recomputed one over
    200 sweeps on 4 x 65,536 spins (measured about 1.4e-6 relative). The
    over-relaxation renormalises each spin, so a learned spin stored with a
    biased norm shows here as a steady drift (a hardware rsq normalisation
    gave about 2e-5)."""
    from cudacode_amd import pyro
    temps = [0.02, 0.1, 0.4, 1.0]
    with ctx(16, temps, n_metro=1, n_or=10, pt_every=0, cos_cone=0.5) as g:
        g.Mc(20)
        g.gather_environments(4096, 1, 65536)
        g.Mc(5)
        g.gather_environments(4096, 2, 65536)
        g.mlp_set_weights(pyro.mlp_init(3))
        g.mlp_train(300, batch=2048, lr=3e-3, lr_final=3e-4)
        g.set_passes(0, 10)
        g.set_learned_passes(1)
        g.set_spins(g.get_spins())  # re-anchor: the books start exact here
        g.Mc(200)
        a, p = g.acceptance_learned()
        assert a.sum() > 0.7 * p.sum()
        E, inc = g.TotalEnergyv2(), g.energy_incremental()
        say(f"inc - E {inc - E}")
        assert np.allclose(inc, E, rtol=4e-6, atol=1e-3)


def test_trained_network_gains_over_plain_metropolis():
    """What the trained network buys: one sweep of learned proposals moves the
    spins much further than one sweep of the plain cone Metropolis (adapted
    to 50% acceptance), while being accepted > 80% of the time. Overlap
    q = <s_i(t) . s_i(t+1)> over all sites, averaged over 10 sweeps."""
    from cudacode_amd import learned
    L, temps = 8, [0.1, 0.4]
    with ctx(L, temps, n_metro=1, n_or=0, cos_cone=0.5, seed=31) as g:
        g.CalculateSolidAngle(0.5, 10, 5)
        learned.train(g, snapshots=8, per_replica=2048, steps=400, batch=2048, init_seed=4)

        def overlap():
            qs = []
            for _ in range(10):
                a = g.get_spins().reshape(len(temps), 3, -1)
                g.Mc(1, use_graph=False)
                b = g.get_spins().reshape(len(temps), 3, -1)
                qs.append((a * b).sum(axis=1).mean(axis=1))
            return np.mean(qs, axis=0)

        q_plain = overlap()
        prev = learned.use_learned(g, 1)
        g.reset_acceptance()
        q_learned = overlap()
        a, p = g.acceptance_learned()
        learned.use_plain(g, prev)
    say(f"overlap plain {q_plain} learned {q_learned} acceptance {a.sum() / p.sum():.3f}")
    assert a.sum() > 0.8 * p.sum()
    assert np.all(q_learned < q_plain - 0.02)


def test_adam_step_closed_form_and_descent():
    """With a constant gradient g every bias-corrected Adam step moves each
    weight by -lr g / (|g| + eps); with the loss gradient of a batch, steps
    lower that batch's loss."""
    w0 = random_weights(7).astype(np.float32)
    g = rng.normal(size=w0.size).astype(np.float32)
    lr = 1e-3
    with ctx() as c:
        c.mlp_set_weights(w0)
        for _ in range(3):
            c.mlp_adam_step(g, lr)
        w3 = c.mlp_get_weights()
        expect = w0.astype(np.float64) - 3 * lr * g / (np.abs(g) + 1e-8)
        assert np.allclose(w3, expect, rtol=1e-5, atol=1e-6)
        c.mlp_set_weights(w0)
        e = env_rows(1024)
        loss0, grad = c.mlp_loss_grad(e)
        for _ in range(20):
            _, grad = c.mlp_loss_grad(e)
            c.mlp_adam_step(grad, 3e-3)
        loss1, _ = c.mlp_loss_grad(e)
        assert loss1 < 0.8 * loss0


@pytest.mark.parametrize("graph,streams,groups", [
    (True, 2, 0), (False, 2, 0), (True, 4, 0), (True, 1, 3), (False, 2, 5)])
def test_learned_sweeps_on_split_streams_identical(graph, streams, groups):
    """Learned sweeps whose replica groups run on 2 or 4 streams, or
    cache-blocked (each exchange interval group after group, one or two in
    flight), equal one-stream learned sweeps bit for bit (spins, learned
    acceptance, incremental energies to FP64 rounding: per-block atomic
    sums), with exchanges."""
    from cudacode_amd import pyro
    temps = [0.05, 0.2, 0.5, 1.0, 2.0]
    kw = dict(n_metro=0, n_or=2, pt_every=3, cos_cone=0.5)
    w = random_weights(11, out_scale=0.3).astype(np.float32)
    with ctx(8, temps, **kw) as a, ctx(8, temps, **kw) as b:
        for g, n in ((a, streams), (b, 1)):
            g.set_streams(n)
            g.set_group_blocking(groups if g is a else 0)
            g.mlp_set_weights(w)
            g.set_learned_passes(1)
        a.Mc(12, use_graph=graph)
        b.Mc(12, use_graph=graph)
        assert a.streams() == streams and b.streams() == 1
        assert np.array_equal(a.get_spins(), b.get_spins())
        assert np.array_equal(a.acceptance_learned()[0], b.acceptance_learned()[0])
        assert a.acceptance_learned()[0].sum() > 0
        assert np.allclose(a.energy_incremental(), b.energy_incremental(), rtol=1e-11, atol=0)
        assert np.array_equal(a.temperatures()[1], b.temperatures()[1])


def check_trained_network(g, n_rows=4096, seed=0):
    """The library's forward pass and MH log ratio on n_rows of the
    environments the ctx gathered (the trained network's real inputs), each
    with a uniform proposal s': the BF16 MFMA outputs within BF16 tolerance of
    the FP64 oracle forward, the device log ratio equal to the FP64 host
    formula evaluated on the device's own outputs."""
    env = g.environments()
    assert env.shape[0] >= n_rows
    pick = np.random.default_rng(seed).choice(env.shape[0], n_rows, replace=False)
    e = env[pick].astype(np.float32)
    sp = env_rows(n_rows)[:, 3:]
    rows = np.concatenate([e, sp], axis=1)
    out = g.mlp_eval(rows)
    w = g.mlp_get_weights().astype(np.float64)
    ref1 = np.array([oracle.mlp_forward(w, oracle.mlp_features(r[:3].astype(np.float64),
                                                               r[3:6].astype(np.float64)))
                     for r in rows])
    ref2 = np.array([oracle.mlp_forward(w, oracle.mlp_features(r[:3].astype(np.float64),
                                                               r[6:9].astype(np.float64)))
                     for r in rows])
    scale = np.sqrt(np.mean(ref1 ** 2))
    err = max(np.abs(out[:, :4] - ref1).max(), np.abs(out[:, 4:8] - ref2).max()) / scale
    say(f"trained forward vs FP64 oracle on {n_rows} gathered environments: max err {err:.4f} x rms")
    assert err < 0.05
    worst = 0.0
    for r, o in zip(rows, out):
        ref = host_log_ratio(o[:4], o[4:8], r[:3], r[3:6], r[6:9])
        k = max(math.exp(min(o[3], 8.3)), math.exp(min(o[7], 8.3)))
        d = abs(o[8] - ref) / (1 + k + abs(ref))
        worst = max(worst, d)
    say(f"device MH log ratio vs FP64 host formula: max scaled err {worst:.2e}")
    assert worst < 1e-4


def check_books(g, tag):
    n_bad, first = g.check_spins(1e-5)
    E, inc = g.TotalEnergyv2(), g.energy_incremental()
    rel = np.abs(inc - E).max() / np.abs(E).max()
    say(f"{tag}: non-unit spins {n_bad}, max |incremental - recomputed| / |E| = {rel:.2e}")
    assert n_bad == 0, first
    assert np.allclose(inc, E, rtol=1e-6, atol=1e-3)


def test_c4_workload_learned_proposal():
    """The learned-proposal workload at its size: pyrochlore L=16 x 64
    replicas on the geometric ladder 0.01..1, the network trained by the
    library on 1,048,576 environments (beta h, s) gathered from the plain
    sampler across the ladder, then 46 to 55 sweeps of 1 learned + 10
    over-relaxation passes with five exchanges (every 10 sweeps). The trained
    forward pass on 4,096 of the gathered environments matches the FP64
    oracle within BF16 tolerance, the device MH log ratio equals the FP64
    formula, acceptance exceeds 0.8, spins stay unit and the incremental
    energy equals the recomputed one (5 sweeps after the last re-anchoring
    exchange). The move replaced is the Metropolis phase of
    Pyrochlore.cu:829-885."""
    from cudacode_amd import learned, pyro
    temps = pyro.geometric_ladder(0.01, 1.0, 64)
    with ctx(16, temps, n_metro=1, n_or=10, pt_every=10, cos_cone=0.5, seed=56081376) as g:
        g.CalculateSolidAngle(0.5, 5, 10)
        g.Mc(100)
        info = learned.train(g, snapshots=16, per_replica=1024, steps=2000, init_seed=1)
        say(f"trained: {info}")
        assert info["environments"] == 1 << 20 and g.environments().shape[0] == 1 << 20
        assert info["loss_last"] < 0.1 * info["loss_first"]
        check_trained_network(g)
        learned.use_learned(g, 1)
        g.reset_acceptance()
        s0, r0 = g.sweep_counters()
        g.Mc(55 - s0 % 10)  # 5 exchanges, ending 5 sweeps after the last
        a, p = g.acceptance_learned()
        acc = a.sum() / p.sum()
        say(f"c4 learned sweeps: acceptance {acc:.3f}, counters {g.sweep_counters()}")
        assert g.sweep_counters()[1] - r0 == 5
        assert acc > 0.8
        assert np.all(a > 0.5 * p)  # every rung
        check_books(g, "c4")


def test_c5_workload_learned_pass_interleaved():
    """The large-lattice workload at its size: L=64 (4,194,304 spins) x 32
    replicas on the geometric ladder 0.02..0.5 (1.61 GB of spins), the
    network trained by the library on 1,048,576 environments of this
    ensemble, then 15 to 24 sweeps of 1 Metropolis + 1 learned + 10
    over-relaxation passes with two exchanges (every 10 sweeps): the trained
    forward pass and MH log ratio as for the L=16 workload, learned
    acceptance > 0.8, unit spins, incremental == recomputed energy."""
    from cudacode_amd import learned, pyro
    temps = pyro.geometric_ladder(0.02, 0.5, 32)
    with ctx(64, temps, n_metro=1, n_or=10, pt_every=10, cos_cone=0.5, seed=56081376) as g:
        g.CalculateSolidAngle(0.5, 3, 5)
        g.Mc(20)
        info = learned.train(g, snapshots=16, per_replica=2048, steps=2000, init_seed=1)
        say(f"trained: {info}")
        assert info["environments"] == 1 << 20
        check_trained_network(g, seed=1)
        g.set_learned_passes(1)  # interleaved: Metropolis, learned, over-relaxation
        g.reset_acceptance()
        s0, r0 = g.sweep_counters()
        g.Mc(25 - s0 % 10)  # 2 exchanges, ending 5 sweeps after the last
        a, p = g.acceptance_learned()
        am, pm = g.acceptance()
        say(f"c5 interleaved sweeps: learned acceptance {a.sum() / p.sum():.3f}, "
            f"Metropolis {am.sum() / pm.sum():.3f}, counters {g.sweep_counters()}")
        assert g.sweep_counters()[1] - r0 == 2
        assert a.sum() > 0.8 * p.sum()
        assert am.sum() > 0
        check_books(g, "c5")


def test_learned_books_on_every_launch_path():
    """The learned phases book their energy change and acceptance through
    per-block slots that a fold moves into the books once per sweep. Every
    path keeps the books exact to FP32 rounding of the moves and the
    acceptance counts whole: two learned passes a sweep (the second pass adds
    to the slots the first wrote), a single-phase call, and eager timed
    launches (folded after the timed region). With over-relaxation tracked
    in FP64 and no exchanges, the incremental energy equals the recomputed
    one within the learned moves' FP32 rounding."""
    temps = [0.05, 0.3, 1.0]
    w = random_weights(5, out_scale=0.3).astype(np.float32)
    with ctx(8, temps, n_metro=0, n_or=2, pt_every=0) as g:
        g.set_kernel_options(1, True)
        g.mlp_set_weights(w)
        g.set_learned_passes(2)
        g.set_spins(g.get_spins())  # re-anchor
        g.Mc(6)
        a1, p1 = g.acceptance_learned()
        assert np.all(p1 == 6 * 2 * 16 * 8 ** 3)
        assert np.all(a1 > 0) and np.all(a1 <= p1)
        E, inc = g.TotalEnergyv2(), g.energy_incremental()
        say(f"two passes: inc - E {inc - E}")
        assert np.allclose(inc, E, rtol=1e-6, atol=1e-4)
        g.learned_phase(2, 6)
        avg_ms, _ = g.time_phase(2, 5)
        assert avg_ms > 0
        a2, p2 = g.acceptance_learned()
        assert np.all(p2 == p1 + 6 * 4 * 8 ** 3)
        assert np.all(a2 > a1)
        E, inc = g.TotalEnergyv2(), g.energy_incremental()
        say(f"single phase + timed launches: inc - E {inc - E}")
        assert np.allclose(inc, E, rtol=1e-6, atol=1e-4)
