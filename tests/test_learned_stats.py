"""The learned-vs-plain verdict (cudacode_amd/learned.py) on synthetic legs.

Energy series per rung are AR(1) processes x_t = phi x_{t-1} + e_t, whose
integrated autocorrelation time in Sokal's convention (tau = 1/2 +
sum_{t>=1} rho(t)) is (1 + phi) / (2 (1 - phi)) in closed form. With
known sweeps/s for each sampler, the samples/s ratio of the two samplers is
(sweeps/s_L / sweeps/s_P) * (tau_P / tau_L) exactly, which pins pool(),
ratio_interval() and judge_paired() against mistakes such as an inverted
ratio, a tau taken from the wrong sampler or a dropped factor of 2; planted
mean shifts pin the agreement statistics, and choose() is checked on both
outcomes.
"""
import numpy as np
import pytest

from cudacode_amd import learned

R = 6  # rungs


def ar1(phi, n, rng, mean=0.0, scale=1.0):
    """(n, R) AR(1) series, one column per rung, started in equilibrium."""
    phi = np.asarray(phi, np.float64)
    x = np.empty((n, phi.size))
    x[0] = rng.standard_normal(phi.size) / np.sqrt(1.0 - phi ** 2)
    e = rng.standard_normal((n, phi.size))
    for t in range(1, n):
        x[t] = phi * x[t - 1] + e[t]
    return mean + scale * x


def tau_ar1(phi):
    return (1.0 + np.asarray(phi)) / (2.0 * (1.0 - np.asarray(phi)))


def leg(series, sweeps_per_s, acceptance=0.5):
    taus, err = learned._taus(series)
    n = series.shape[0]
    samples = sweeps_per_s / (2.0 * taus)
    return learned.Leg(series, sweeps_per_s, taus, samples, acceptance, err, n / sweeps_per_s)


PHI_P = np.linspace(0.55, 0.85, R)  # plain: tau 1.7 .. 6.2 sweeps
PHI_L = np.linspace(0.10, 0.40, R)  # learned: tau 0.6 .. 1.2 sweeps
SPS_P, SPS_L = 3000.0, 2400.0
MEANS = np.linspace(-2.0, -0.5, R) * 1000.0  # per-rung energies


def legs(n, seed, shift_learned=0.0):
    rng = np.random.default_rng(seed)
    p = [leg(ar1(PHI_P, n, rng, MEANS, 5.0), SPS_P, 0.4) for _ in range(2)]
    q = [leg(ar1(PHI_L, n, rng, MEANS + shift_learned, 5.0), SPS_L, 0.9) for _ in range(2)]
    return p, q


def test_tau_estimator_matches_the_ar1_closed_form():
    rng = np.random.default_rng(1)
    x = ar1(PHI_P, 40000, rng)
    taus, err = learned._taus(x)
    assert np.all(np.abs(taus - tau_ar1(PHI_P)) < 4 * err + 0.02 * tau_ar1(PHI_P))


def test_pool_combines_legs():
    p, _ = legs(3000, seed=2)
    P = learned.pool(p)
    assert P.series.shape == (6000, R)
    assert P.sweeps_per_s == pytest.approx(SPS_P)
    assert np.allclose(P.taus, (p[0].taus + p[1].taus) / 2)
    assert np.allclose(P.tau_err, np.hypot(p[0].tau_err, p[1].tau_err) / 2)
    assert np.allclose(P.samples, SPS_P / (2.0 * P.taus))


def test_ratio_interval_degenerates_without_tau_errors():
    p, q = legs(3000, seed=3)
    P, Q = learned.pool(p), learned.pool(q)
    P.tau_err = np.zeros(R)
    Q.tau_err = np.zeros(R)
    point = np.median((Q.sweeps_per_s / P.sweeps_per_s) * np.maximum(P.taus, 0.5) / np.maximum(Q.taus, 0.5))
    assert np.allclose(learned.ratio_interval(P, Q, n_boot=200), point)


def test_verdict_recovers_the_closed_form_ratio():
    p, q = legs(8000, seed=4)
    v = learned.judge_paired(p, q, 1.0, np.linspace(1.0, 0.1, R))
    exact = np.median((SPS_L / SPS_P) * tau_ar1(PHI_P) / tau_ar1(PHI_L))
    lo, hi = v["samples_ratio_median_95"]
    assert lo <= v["samples_ratio_median"] <= hi
    assert lo <= exact <= hi
    assert v["samples_ratio_median"] == pytest.approx(exact, rel=0.15)
    per_rung = np.array(v["samples_ratio_per_rung"])
    assert np.allclose(per_rung, (SPS_L / SPS_P) * tau_ar1(PHI_P) / tau_ar1(PHI_L), rtol=0.25)
    assert v["sweeps_per_s"] == pytest.approx(SPS_L) and v["plain_sweeps_per_s"] == pytest.approx(SPS_P)
    assert v["series_long_enough"]
    # same law: the samplers agree about as often as each sampler with itself
    assert v["energy_agreement_rungs_within_2sigma"] >= R - 2
    assert v["control_plain_vs_plain_within_2sigma"] >= R - 2
    assert v["control_learned_vs_learned_within_2sigma"] >= R - 2


def test_verdict_flags_short_series_and_a_mean_shift():
    p, q = legs(200, seed=5)
    assert not learned.judge_paired(p, q, 1.0, np.linspace(1.0, 0.1, R))["series_long_enough"]
    p, q = legs(8000, seed=6, shift_learned=3.0)  # 3 units against a standard error ~0.1
    v = learned.judge_paired(p, q, 1.0, np.linspace(1.0, 0.1, R))
    assert v["energy_agreement_rungs_within_2sigma"] == 0
    assert v["energy_agreement_max_z"] > 10 * max(v["control_plain_vs_plain_max_z"], 2.0)
    assert v["control_learned_vs_learned_within_2sigma"] >= R - 2


class FakeCtx:
    def __init__(self):
        self.n_metro, self.n_or, self.n_learned = 1, 10, 0

    def set_passes(self, n_metro, n_or):
        self.n_metro, self.n_or = n_metro, n_or

    def set_learned_passes(self, n):
        self.n_learned = n


@pytest.mark.parametrize("shift, expect", [(0.0, "learned"), (3.0, "plain")])
def test_choose_keeps_or_drops_the_learned_proposal(shift, expect):
    p, q = legs(8000, seed=7, shift_learned=shift)
    v = learned.judge_paired(p, q, 1.0, np.linspace(1.0, 0.1, R))
    g = FakeCtx()
    prev = (g.n_metro, g.n_or)
    assert learned.choose(g, v, prev) == expect
    if expect == "learned":
        assert (g.n_metro, g.n_or, g.n_learned) == (0, 10, 1)
    else:
        assert (g.n_metro, g.n_or, g.n_learned) == (1, 10, 0)


def test_choose_drops_a_slower_sampler():
    p, q = legs(8000, seed=8)
    v = learned.judge_paired(p, q, 1.0, np.linspace(1.0, 0.1, R))
    v = dict(v, samples_ratio_median=0.9)
    g = FakeCtx()
    assert learned.choose(g, v, (1, 10)) == "plain"


def test_observables_verdict_reports_the_lowest_rung():
    """The observables part of the verdict: z-scores of C/N and <|m_A2|^2>
    between two summaries, and the lowest-temperature rung's estimates of
    each sampler copied out unchanged (rungs by descending T)."""
    from cudacode_amd import measure
    rng = np.random.default_rng(5)
    n, T = 2048, np.array([2.0, 1.0, 0.1])
    rows = np.zeros((n, 3, 6))
    for k in range(3):
        rows[:, k, 0] = rng.normal(-500.0 * (k + 1), 4.0, n)
        rows[:, k, 1:] = np.abs(rng.normal(0.1 * (k + 1), 0.01, (n, 5)))
    p = measure.summarise(rows, T, 1024)
    q = measure.summarise(rows[::-1].copy(), T, 1024)  # same samples, other order
    v = learned._obs_verdict(p, q)
    assert v["specific_heat_rungs_within_2sigma"] == 3
    assert v["m_A2_sq_rungs_within_2sigma"] == 3
    low = v["lowest_T"]
    assert low["T"] == 0.1
    assert low["plain_energy_per_spin"] == [p.e[-1, 0], p.e[-1, 1]]
    assert low["energy_per_spin"][0] == pytest.approx(p.e[-1, 0], rel=1e-12)
    assert low["specific_heat"] == [q.c[-1, 0], q.c[-1, 1]]
    assert low["plain_m_A2_sq"] == [p.m2[-1, 0, 0], p.m2[-1, 0, 1]]
