"""Oracle dynamics pinned by closed forms and invariants of the model.

H = jeff * sum_<ij> s_i.s_j on the pyrochlore lattice. Closed forms used:
  ferromagnetic state: 6 bonds per tetrahedron, E/N = 3 jeff;
  all-in-all-out (any state with zero spin sum on every tetrahedron): E/N = -jeff;
  high temperature: E/N = -jeff^2 / T to leading order (3 bonds per spin,
      <s_i.s_j> = -beta jeff / 3);
  low temperature: C/N -> 3/4 (Moessner & Chalker, PRL 80, 2929 (1998)), so
      E/N = -jeff + (3/4) T.
Invariants: incremental FP64 energy equals a full recompute; over-relaxation
conserves the energy; a colour phase touches only its own sites and does not
depend on the order or thread count of its updates; runs are resumable.
"""
import numpy as np
import pytest

from oracle import oracle

AIAO_AXES = np.array([[-1, -1, -1], [1, 1, -1], [-1, 1, 1], [1, -1, 1]], np.float64) / np.sqrt(3)


def state(L, R, fn):
    """spins[r][comp][site][cell] with spin(site) = fn(site)."""
    nc = L ** 3
    s = np.zeros((R, 3, 16, nc), np.float32)
    for site in range(16):
        v = np.asarray(fn(site), np.float32)
        s[:, :, site, :] = v[None, :, None]
    return s.ravel()


@pytest.mark.parametrize("jeff", [1.0, 0.25])
@pytest.mark.parametrize("L", [2, 3, 4])
def test_ferro_and_aiao_energies(L, jeff):
    N = 16 * L ** 3
    ferro = state(L, 1, lambda site: [1, 0, 0])
    obs = oracle.observables(ferro, L, 1, jeff)
    assert obs[0, 0] == pytest.approx(3 * jeff * N, rel=1e-12)
    assert obs[0, 1:4] == pytest.approx([N / 4, 0, 0])
    aiao = state(L, 1, lambda site: AIAO_AXES[site % 4])
    obs = oracle.observables(aiao, L, 1, jeff)
    assert obs[0, 0] == pytest.approx(-jeff * N, rel=1e-6)


def fresh(L, R, seed=7):
    return oracle.init_spins(L, R, seed)


def test_init_spins_unit_and_isotropic():
    L, R = 4, 3
    s = fresh(L, R).reshape(R, 3, -1).astype(np.float64)
    n = np.sqrt((s ** 2).sum(axis=1))
    assert np.all(np.abs(n - 1) < 3e-7)
    m = s.mean(axis=2)
    assert np.all(np.abs(m) < 5 / np.sqrt(s.shape[2] * 3))


def test_phase_touches_only_its_colour():
    L, R = 3, 2
    s0 = fresh(L, R)
    for colour in range(4):
        s = s0.copy()
        oracle.metropolis_phase(s, L, R, colour, 5, 0, [2.0, 0.5], [-1.0, 0.3], 1.0, 99)
        d = (s != s0).reshape(R, 3, 16, -1).any(axis=(0, 1, 3))
        assert d[colour::4].any()
        assert not np.delete(d, np.arange(colour, 16, 4)).any()


def test_phase_independent_of_thread_count():
    L, R = 4, 3
    ref = fresh(L, R)
    outs = []
    for nt in (1, 3, 8):
        oracle.set_num_threads(nt)
        s = ref.copy()
        for c in range(4):
            oracle.metropolis_phase(s, L, R, c, 1, 0, [1.0, 2.0, 3.0], [-1.0] * 3, 1.0, 11)
            oracle.overrelax_phase(s, L, R, c, 1.0)
        outs.append(s)
    oracle.set_num_threads(0)
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])


def test_incremental_energy_matches_recompute():
    L = 3
    run = oracle.OracleRun(L, [0.3, 1.0, 4.0], seed=5, n_metro=1, n_or=3, cos_cone=0.5)
    run.run(50)
    e_full = run.observables()[:, 0]
    assert np.allclose(run.anchor + run.dE, e_full, rtol=1e-11, atol=1e-9)
    assert run.accepted.sum() > 0


def test_exchange_reanchors_incremental_energy():
    """The full energies computed for a swap become the anchor (dE = 0), and
    between swaps the books follow the moves: after 3 sweeps past the swap at
    sweep 20, anchor + dE equals the recomputed energy again."""
    L = 3
    run = oracle.OracleRun(L, [0.3, 1.0, 4.0], seed=5, n_metro=1, n_or=3, pt_every=10,
                           cos_cone=0.5)
    start = run.anchor.copy()
    run.run(20)
    assert np.array_equal(run.dE, np.zeros(3))
    assert not np.array_equal(run.anchor, start)
    assert np.allclose(run.anchor, run.observables()[:, 0], rtol=0, atol=0)
    run.run(3)
    assert np.any(run.dE != 0)
    assert np.allclose(run.anchor + run.dE, run.observables()[:, 0], rtol=1e-11, atol=1e-9)


def test_overrelaxation_conserves_energy():
    L, R = 4, 2
    s = fresh(L, R)
    e0 = oracle.observables(s, L, R, 1.0)[:, 0]
    for _ in range(20):
        for c in range(4):
            oracle.overrelax_phase(s, L, R, c, 1.0)
    e1 = oracle.observables(s, L, R, 1.0)[:, 0]
    assert np.allclose(e0, e1, rtol=0, atol=16 * L ** 3 * 1e-6)
    assert not np.array_equal(s, fresh(L, R))


def test_run_is_resumable():
    """Counter-based RNG: 30 sweeps equal 10 + 20 sweeps, bit for bit, with PT."""
    temps = [0.2, 0.4, 0.8, 1.6]
    a = oracle.OracleRun(2, temps, seed=3, n_metro=1, n_or=2, pt_every=5, cos_cone=0.2)
    b = oracle.OracleRun(2, temps, seed=3, n_metro=1, n_or=2, pt_every=5, cos_cone=0.2)
    a.run(30)
    b.run(10)
    b.run(20)
    assert np.array_equal(a.spins, b.spins)
    assert np.array_equal(a.replica_at, b.replica_at)
    assert a.pt_round.value == b.pt_round.value == 6


def test_pt_swap_detailed_balance():
    """Two replicas with fixed energies: the swap is taken with probability
    min(1, exp((1/T1 - 1/T2)(E1 - E2))) (MonteCarlo.cu:131), independent of
    which replica currently holds which rung."""
    T = np.array([2.0, 1.0])
    E = np.array([-10.0, -9.0])  # the hotter replica holds the lower energy
    p_expect = min(1.0, np.exp((1 / 2.0 - 1 / 1.0) * (-10.0 + 9.0)))
    assert p_expect == 1.0
    E = np.array([-9.0, -10.0])
    p_expect = np.exp((1 / 2.0 - 1 / 1.0) * (-9.0 + 10.0))
    n, acc = 20000, 0
    for rnd in range(n):
        temps = T.copy()
        beta = (1 / temps).astype(np.float32)
        cosc = np.array([0.1, 0.2], np.float32)
        at = np.array([0, 1], np.int32)
        a = oracle.pt_swap(2, 0, rnd, 17, E, temps, beta, cosc, at)
        acc += a
        if a:
            assert list(temps) == [1.0, 2.0] and list(at) == [1, 0]
            assert list(cosc) == pytest.approx([0.2, 0.1])
            assert beta[0] == np.float32(1.0)
    p = acc / n
    assert abs(p - p_expect) < 5 * np.sqrt(p_expect * (1 - p_expect) / n)


def test_pt_pairs_follow_parity():
    R = 6
    temps = np.array([1.0, 2.0, 3.0, 4.0, 5.0, 6.0])
    at = np.argsort(-temps).astype(np.int32)
    E = np.zeros(R)  # every proposal accepted (weight 1)
    beta = (1 / temps).astype(np.float32)
    cosc = np.zeros(R, np.float32)
    assert oracle.pt_swap(R, 0, 0, 1, E, temps, beta, cosc, at) == 3
    assert list(at) == [4, 5, 2, 3, 0, 1]
    assert oracle.pt_swap(R, 1, 1, 1, E, temps, beta, cosc, at) == 2
    assert list(at) == [4, 2, 5, 0, 3, 1]
    # the temperature held by each replica follows its rung
    for rung, rep in enumerate(at):
        assert temps[rep] == 6.0 - rung


def equilibrium_energy(L, T, sweeps, therm, start, cos_cone, n_or=3, seed=21):
    run = oracle.OracleRun(L, [T], seed=seed, n_metro=1, n_or=n_or, cos_cone=cos_cone)
    if start is not None:
        run.spins[:] = start
        run.anchor = run.observables()[:, 0].copy()
    run.run(therm)
    es = []
    for _ in range(sweeps):
        run.run(1)
        es.append(run.observables()[0, 0])
    es = np.array(es) / (16 * L ** 3)
    bm = es.reshape(20, -1).mean(axis=1)
    return es.mean(), bm.std(ddof=1) / np.sqrt(len(bm))


def test_high_temperature_energy():
    T = 40.0
    e, err = equilibrium_energy(2, T, 4000, 200, None, -1.0, n_or=0)
    assert abs(e - (-1.0 / T)) < 4 * err + 3e-4, (e, err)


def test_low_temperature_specific_heat_three_quarters():
    L, T = 2, 0.01
    aiao = state(L, 1, lambda site: AIAO_AXES[site % 4])
    e, err = equilibrium_energy(L, T, 2000, 500, aiao, 0.995, n_or=5)
    c_est = (e + 1.0) / T
    assert abs(c_est - 0.75) < 4 * err / T + 0.04, (c_est, err / T)


def test_rotation_about_field():
    """rotate_about: unit norm, the component along h kept, the perpendicular
    part turned by 2 pi u; u = 1/2 is the over-relaxation reflection."""
    rng = np.random.default_rng(4)
    for _ in range(200):
        h = rng.normal(size=3).astype(np.float32)
        s = rng.normal(size=3)
        s = (s / np.linalg.norm(s)).astype(np.float32)
        u = np.float32(rng.random())
        out = oracle.rotate_about(h, s, u).astype(np.float64)
        n = h / np.linalg.norm(h)
        assert abs(np.linalg.norm(out) - 1) < 1e-6
        assert abs(n @ out - n @ s) < 2e-6
        p0, p1 = s - (n @ s) * n, out - (n @ out) * n
        if np.linalg.norm(p0) > 1e-3:
            ang = np.arctan2(n @ np.cross(p0, p1), p0 @ p1) % (2 * np.pi)
            d = (ang - 2 * np.pi * float(u)) % (2 * np.pi)
            assert min(d, 2 * np.pi - d) < 1e-4
        half = oracle.rotate_about(h, s, np.float32(0.5)).astype(np.float64)
        assert np.allclose(half, oracle.overrelax(h, s), atol=2e-6)
    s = np.array([0.6, 0.0, 0.8], np.float32)
    assert np.array_equal(oracle.rotate_about(np.zeros(3, np.float32), s, 0.3), s)


def test_post_rotation_keeps_the_boltzmann_law():
    """Metropolis sweeps with and without the rotation about h sample the same
    energy distribution (the rotation keeps every local energy)."""
    L, T = 2, 0.5
    res = []
    for on in (False, True):
        oracle.set_post_rotate(on)
        try:
            res.append(equilibrium_energy(L, T, 4000, 300, None, 0.3, n_or=0, seed=5 + on))
        finally:
            oracle.set_post_rotate(False)
    (e0, s0), (e1, s1) = res
    assert abs(e0 - e1) < 4 * np.hypot(s0, s1), res


def test_phase_order_independence():
    """Race check of the sublattice-parallel update: the updates of a colour
    phase visited in a shuffled order (and by several threads) give the same
    spins and acceptance counts bit for bit, so no update reads a value
    another update of the same phase writes."""
    L, temps = 4, [0.2, 0.9]
    kw = dict(n_metro=1, n_or=2, cos_cone=0.4)
    a = oracle.OracleRun(L, temps, seed=9, **kw)
    a.run(4)
    rng = np.random.default_rng(2)
    oracle.set_update_order(rng.permutation(4 * L ** 3))
    try:
        b = oracle.OracleRun(L, temps, seed=9, **kw)
        b.run(4)
    finally:
        oracle.set_update_order(None)
    assert np.array_equal(a.spins, b.spins)
    assert np.array_equal(a.accepted, b.accepted)
    assert np.allclose(a.dE, b.dE, rtol=1e-12, atol=1e-9)


def test_cone_rule_moves_toward_the_target_and_stays_bounded():
    """The oracle's cone rule, pinned by what it is for rather than by its
    formula: all rejected shrinks the cone (cos grows), all accepted widens it
    (cos falls), and cos_cone never leaves [-1, 1) however many rounds push
    it against either end."""
    c = np.array([0.2, 0.2, 0.99999994, -1.0], np.float32)
    prop = np.full(4, 1000, np.uint64)
    acc = np.array([0, 1000, 0, 1000], np.uint64)
    c0 = c.copy()
    oracle.adapt_cone_step(0.5, acc, prop, c)
    assert c[0] > c0[0] and c[1] < c0[1]
    for _ in range(200):
        oracle.adapt_cone_step(0.5, acc, prop, c)
    assert c[0] < 1.0 and c[2] < 1.0 and c[1] == -1.0 and c[3] == -1.0
    assert c[0] > 0.9999 and c[2] > 0.9999


def test_cone_rule_converges_from_both_sides():
    """From the whole sphere at low temperature (most proposals rejected) and
    from a tiny cone at moderate temperature (almost all accepted), rounds of
    the rule bring each replica's Metropolis acceptance to the 50% target;
    where even the whole sphere is accepted more than half the time (high T)
    the cone opens to the whole sphere."""
    def ratio(o, sweeps=20):
        o.accepted[:] = 0
        o.run(sweeps)
        return o.accepted / (sweeps * 16 * o.L ** 3)

    o = oracle.OracleRun(4, [0.05, 0.3], seed=5, n_metro=1, n_or=1, cos_cone=-1.0)
    o.adapt_cone(0.5, 40, 3)
    r = ratio(o)
    assert np.all(np.abs(r - 0.5) < 0.1), r
    o = oracle.OracleRun(4, [0.3, 20.0], seed=6, n_metro=1, n_or=1, cos_cone=0.9999)
    o.adapt_cone(0.5, 80, 3)
    r = ratio(o)
    assert abs(r[0] - 0.5) < 0.1, r
    assert r[1] > 0.5 and o.cosc[1] == -1.0, (r, o.cosc)


def _schedule(sets, angle):
    class _S:
        def __enter__(self):
            oracle.set_update_schedule(sets, angle)

        def __exit__(self, *a):
            oracle.set_update_schedule(False, False)
    return _S()


def test_drawn_sets_cover_all_24_uniformly():
    """The drawn-set schedule picks each of the 24 independent sets about
    equally often (chi-square over 24,000 phase draws), differently for the
    Metropolis and over-relaxation stages, and the same for every replica."""
    n = 24000
    idx = np.array([oracle.drawn_set(s // 8, (s // 4) & 1, 0, s & 3, 77) for s in range(n)])
    counts = np.bincount(idx, minlength=24)
    chi2 = ((counts - n / 24) ** 2 / (n / 24)).sum()
    assert chi2 < 60.0, (chi2, counts)  # 23 degrees of freedom: p ~ 4e-5
    met = [oracle.drawn_set(s, 0, 0, 1, 77) for s in range(200)]
    ore = [oracle.drawn_set(s, 1, 0, 1, 77) for s in range(200)]
    assert met != ore


def test_drawn_sets_are_independent_sets():
    """Every set the schedule can draw holds four mutually non-adjacent sites
    (no link of the 16-site cell, periodic images included, joins two)."""
    links = oracle.links16()
    for row in oracle.independent_sets():
        members = set(int(v) for v in row)
        assert len(members) == 4
        for site in members:
            assert not members & set(int(links[site, k, 3]) for k in range(6))


def test_random_angle_overrelaxation_conserves_energy():
    """Rotations about the local field keep every local energy, hence the
    total (to FP32 rounding), keep spins unit, and move them differently from
    the reflection by pi."""
    L, R = 4, 2
    for sets in (False, True):
        s = fresh(L, R)
        p = s.copy()
        e0 = oracle.observables(s, L, R, 1.0)[:, 0]
        with _schedule(sets, True):
            for o in range(10):
                for c in range(4):
                    oracle.overrelax_phase_ex(s, L, R, c, 3, o, 1.0, 99)
        for _ in range(10):
            for c in range(4):
                oracle.overrelax_phase(p, L, R, c, 1.0)
        e1 = oracle.observables(s, L, R, 1.0)[:, 0]
        assert np.allclose(e0, e1, rtol=0, atol=16 * L ** 3 * 1e-6), (sets, e0, e1)
        n = np.sqrt((s.reshape(R, 3, -1) ** 2).sum(axis=1))
        assert np.abs(n - 1).max() < 1e-6
        assert not np.array_equal(s, p)


def test_drawn_colour_orders_are_uniform_permutations():
    """The drawn colour order of a pass is a permutation of the four colour
    classes, each of the 24 orders about equally often (chi-square over 4,800
    passes), so every site moves exactly once a pass."""
    import collections
    n = 4800
    c = collections.Counter(tuple(oracle.drawn_colour(s, s & 1, s >> 12, p, 31) for p in range(4))
                            for s in range(n))
    assert all(sorted(k) == [0, 1, 2, 3] for k in c)
    assert len(c) == 24
    chi2 = sum((v - n / 24) ** 2 / (n / 24) for v in c.values())
    assert chi2 < 60.0, (chi2, c)


@pytest.mark.parametrize("sets,angle", [(1, False), (False, True), (1, True), (2, False), (2, True)])
def test_schedule_keeps_the_high_temperature_energy(sets, angle):
    """Drawn sets and random-angle over-relaxation sample the same law:
    E/N = -1/T at T = 40 (leading order; the next term is 2/(3T^2) = 4e-4)."""
    T = 40.0
    with _schedule(sets, angle):
        e, err = equilibrium_energy(2, T, 2000, 100, None, -1.0, n_or=1 if angle else 0)
    assert abs(e - (-1.0 / T)) < 4 * err + 1e-3, (e, err)


def test_schedule_runs_are_resumable_and_order_independent():
    """Counter-based draws: 16 sweeps equal 6 + 10 under the drawn-set,
    random-angle schedule with PT; visiting each phase's updates in a shuffled
    order gives the same spins (no update of a drawn set reads another's
    write)."""
    temps = [0.2, 0.5, 1.3]
    kw = dict(n_metro=1, n_or=2, pt_every=4, cos_cone=0.3)
    with _schedule(1, True):
        a = oracle.OracleRun(3, temps, seed=8, **kw)
        b = oracle.OracleRun(3, temps, seed=8, **kw)
        a.run(16)
        b.run(6)
        b.run(10)
        assert np.array_equal(a.spins, b.spins)
        assert np.array_equal(a.replica_at, b.replica_at)
        rng = np.random.default_rng(3)
        oracle.set_update_order(rng.permutation(4 * 27))
        try:
            c = oracle.OracleRun(3, temps, seed=8, **kw)
            c.run(16)
        finally:
            oracle.set_update_order(None)
        assert np.array_equal(a.spins, c.spins)
        assert np.array_equal(a.accepted, c.accepted)
    d = oracle.OracleRun(3, temps, seed=8, **kw)
    d.run(16)
    assert not np.array_equal(a.spins, d.spins)
