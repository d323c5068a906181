"""GPU parity: the gfx950 kernels against the CPU oracle on the same Philox
stream, through the C ABI.

Bar: spins bit-identical where the arithmetic is (single phases, init), and
within 1e-5 after 100 sweeps (they are expected to stay identical: both sides
evaluate the same float expressions in the same order); integer counts exact;
FP64 energies to 1e-12 relative (different summation order); at full size,
size-independent properties (unit spins, incremental == recomputed energy,
over-relaxation energy conservation).
"""
import sys

import numpy as np
import pytest

from oracle import oracle

pytestmark = pytest.mark.gpu

SEED = 0x5EED1234ABCD


def say(msg):
    print(f"[gpu-parity] {msg}", flush=True)
    sys.stdout.flush()


def gpu(L, temps, variant=1, track=True, **kw):
    """A ctx on device 0. Tests default to exact FP64 energy tracking in
    over-relaxation so the incremental energy can be compared at 1e-10."""
    from cudacode_amd import pyro
    say(f"create L={L} R={len(temps)} variant={variant} track={track} {kw}")
    g = pyro.Pyrochlore(L, temps, seed=kw.pop("seed", SEED), **kw)
    g.set_kernel_options(variant, track)
    return g


def orc(L, temps, **kw):
    return oracle.OracleRun(L, temps, seed=kw.pop("seed", SEED), **kw)


@pytest.mark.parametrize("L,R", [(4, 3), (5, 2), (16, 2)])
def test_init_spins_bitwise(L, R):
    with gpu(L, [1.0] * R) as g:
        s = g.get_spins()
    o = oracle.init_spins(L, R, SEED)
    assert np.array_equal(s, o)


@pytest.mark.parametrize("L", [4, 5])
def test_single_phases_bitwise(L):
    temps = [0.05, 0.4, 2.5]
    cones = np.array([0.95, 0.3, -1.0], np.float32)
    with gpu(L, temps, cos_cone=-1.0) as g:
        g.set_cos_cone(cones)
        o = orc(L, temps)
        o.cosc[:] = cones
        acc = np.zeros(3, np.uint64)
        dE = np.zeros(3)
        for step, colour in enumerate([0, 1, 2, 3, 2, 0]):
            say(f"metropolis L={L} colour={colour}")
            g.MCMultipleTemperaturesPerSystem(colour, sweep=7 + step, pass_=step % 2)
            oracle.metropolis_phase(o.spins, L, 3, colour, 7 + step, step % 2, o.beta, o.cosc,
                                    o.jeff, SEED, acc, dE)
            assert np.array_equal(g.get_spins(), o.spins), f"metropolis colour {colour}"
            say(f"overrelax L={L} colour={colour}")
            g.OverRelaxationByPiStep(colour)
            oracle.overrelax_phase(o.spins, L, 3, colour, o.jeff, dE)
            assert np.array_equal(g.get_spins(), o.spins), f"overrelax colour {colour}"
        a, p = g.acceptance()
        assert np.array_equal(a, acc)
        assert np.all(p == 6 * 4 * L ** 3)
        e_inc = g.energy_incremental()
        assert np.allclose(e_inc - (o.anchor + dE), 0, atol=1e-9 * 16 * L ** 3)


def test_c1_100_sweeps_spins_agree():
    """L=4, one replica, T=0.5, 1 Metropolis (cone) + 5 over-relaxation per sweep."""
    L, temps = 4, [0.5]
    with gpu(L, temps, n_metro=1, n_or=5, cos_cone=0.2) as g:
        o = orc(L, temps, n_metro=1, n_or=5, cos_cone=0.2)
        for chunk in range(4):
            say(f"c1 sweeps chunk {chunk}")
            g.Mc(25, use_graph=(chunk % 2 == 0))
            o.run(25)
            d = np.abs(g.get_spins() - o.spins).max()
            assert d <= 1e-5, (chunk, d)
        assert np.array_equal(g.get_spins(), o.spins)
        a, _ = g.acceptance()
        assert np.array_equal(a, o.accepted)
        assert g.sweep_counters() == (100, 0)


def test_c2_shape_100_sweeps_spins_agree_with_pt():
    """L=16 (the c2 lattice) with 4 replicas on a geometric ladder in [0.01, 1],
    1 Metropolis + 10 over-relaxation per sweep, swaps every 10 sweeps."""
    from cudacode_amd.pyro import geometric_ladder
    L = 16
    temps = geometric_ladder(0.01, 1.0, 4)
    kw = dict(n_metro=1, n_or=10, pt_every=10, cos_cone=0.5)
    with gpu(L, temps, **kw) as g:
        o = orc(L, temps, **kw)
        say("c2-shape 100 sweeps on GPU (graph)")
        g.Mc(100)
        say("c2-shape 100 sweeps on oracle")
        o.run(100)
        d = np.abs(g.get_spins() - o.spins).max()
        assert d <= 1e-5, d
        T, at = g.temperatures()
        assert np.array_equal(at, o.replica_at)
        assert np.array_equal(T, o.temps)
        assert g.sweep_counters() == (100, 10)


def test_parallel_tempering_trajectory():
    L = 3
    temps = [0.1 * 1.5 ** k for k in range(8)]
    kw = dict(n_metro=1, n_or=2, pt_every=2, cos_cone=0.0)
    with gpu(L, temps, **kw) as g:
        o = orc(L, temps, **kw)
        say("pt 40 sweeps")
        g.Mc(13, use_graph=False)
        g.Mc(27, use_graph=True)
        o.run(40)
        assert np.array_equal(g.get_spins(), o.spins)
        T, at = g.temperatures()
        assert np.array_equal(at, o.replica_at)
        assert np.array_equal(T, o.temps)
        assert np.array_equal(g.cos_cone(), o.cosc)
        assert o.swaps.value > 0


def test_explicit_swap_matches_oracle():
    L = 3
    temps = [0.3, 0.35, 0.4, 0.45, 0.5]
    with gpu(L, temps, n_metro=1, n_or=1) as g:
        o = orc(L, temps, n_metro=1, n_or=1)
        g.Mc(5, use_graph=False)
        o.run(5)
        for rnd in range(6):
            say(f"swap round {rnd}")
            n = g.TrySwapingTemperatures(rnd % 2, rnd)
            e = o.observables()[:, 0]
            m = oracle.pt_swap(o.R, rnd % 2, rnd, SEED, e, o.temps, o.beta, o.cosc, o.replica_at)
            assert n == m
        T, at = g.temperatures()
        assert np.array_equal(at, o.replica_at)


@pytest.mark.parametrize("L,R", [(4, 2), (7, 3)])
def test_observables_against_oracle(L, R):
    with gpu(L, [0.2 * (r + 1) for r in range(R)], n_metro=1, n_or=1, cos_cone=0.5) as g:
        g.Mc(10)
        s = g.get_spins()
        o = oracle.observables(s, L, R, 1.0)
        E = g.TotalEnergyv2()
        assert np.allclose(E, o[:, 0], rtol=1e-12, atol=1e-10)
        sub = g.sublattice_sums()
        assert np.allclose(sub, o[:, 1:], rtol=1e-12, atol=1e-10)
        M = g.Magnetization()
        assert np.allclose(M, o[:, 1:].reshape(R, 4, 3).sum(axis=1), atol=1e-9)
        assert np.allclose(g.energy_incremental(), E, rtol=1e-11, atol=1e-9)


def test_closed_form_states_on_gpu():
    L = 4
    N = 16 * L ** 3
    axes = np.array([[-1, -1, -1], [1, 1, -1], [-1, 1, 1], [1, -1, 1]]) / np.sqrt(3)
    s = np.zeros((1, 3, 16, L ** 3), np.float32)
    for site in range(16):
        s[0, :, site, :] = axes[site % 4][:, None]
    with gpu(L, [1.0], spin_len2=0.25) as g:
        g.set_spins(s)
        assert g.TotalEnergyv2()[0] == pytest.approx(-0.25 * N, rel=1e-6)
        m = g.DoMeasurement()[0]
        assert abs(abs(m[0]) - 1) < 1e-6 and np.allclose(m[1:], 0, atol=1e-6)
        s[:] = 0
        s[0, 0] = 1.0
        g.set_spins(s)
        assert g.TotalEnergyv2()[0] == pytest.approx(0.75 * N, rel=1e-12)
        assert np.allclose(g.DoMeasurement()[0, 6:9], [1, 0, 0])


def test_graph_and_eager_identical():
    L = 8
    temps = [0.1, 0.3, 0.9]
    kw = dict(n_metro=1, n_or=4, pt_every=3, cos_cone=0.4)
    with gpu(L, temps, **kw) as a, gpu(L, temps, **kw) as b:
        a.Mc(30, use_graph=True)
        b.Mc(30, use_graph=False)
        assert np.array_equal(a.get_spins(), b.get_spins())
        assert np.array_equal(a.temperatures()[1], b.temperatures()[1])


def test_energy_series_recording():
    L = 4
    temps = [0.2, 0.5, 1.0]
    with gpu(L, temps, n_metro=1, n_or=2, pt_every=5, cos_cone=0.3) as g:
        g.record_start(50)
        g.Mc(20)
        rows = g.record_fetch()
        assert rows.shape == (20, 3)
        e = g.energy_incremental()
        _, at = g.temperatures()
        assert np.array_equal(rows[-1], e[at])
        assert np.allclose(g.TotalEnergyv2()[at], rows[-1], rtol=1e-11)


def test_cone_adaptation_reaches_target():
    L = 4
    temps = [0.05, 0.5, 5.0]
    with gpu(L, temps, n_metro=1, n_or=1, cos_cone=-1.0) as g:
        g.CalculateSolidAngle(target=0.5, rounds=30, sweeps_per_round=5)
        g.Mc(40)
        r = g.CalculateRatio()
        assert np.all(np.abs(r[:2] - 0.5) < 0.12), r
        assert r[2] > 0.4  # at high T even the whole sphere is accepted often


def test_cone_adaptation_equals_oracle():
    """The library's adaptation loop (pyro_adapt_cone: rounds of sweeps,
    per-replica acceptance, the 0.9 / 1/0.9 rule, with exchanges carrying the
    cones between replicas) equals the oracle's restatement bit for bit:
    cones, spins, rungs after 12 rounds from the whole sphere."""
    L = 4
    temps = [0.05, 0.2, 0.5, 5.0]
    kw = dict(n_metro=1, n_or=1, pt_every=2, cos_cone=-1.0)
    with gpu(L, temps, **kw) as g:
        cone = g.CalculateSolidAngle(target=0.5, rounds=12, sweeps_per_round=4)
        s = g.get_spins()
        T, at = g.temperatures()
    o = oracle.OracleRun(L, temps, seed=SEED, **kw)
    o.adapt_cone(0.5, 12, 4)
    assert np.array_equal(cone, o.cosc), (cone, o.cosc)
    assert np.array_equal(s, o.spins)
    assert np.array_equal(T, o.temps) and np.array_equal(at, o.replica_at)
    assert len(set(cone.tolist())) > 1  # the replicas' cones really differ


def test_bad_arguments_raise():
    from cudacode_amd import pyro
    with gpu(4, [1.0]) as g:
        with pytest.raises(pyro.PyroError, match="colour"):
            g.MCMultipleTemperaturesPerSystem(4, 0)
        with pytest.raises(pyro.PyroError, match="expected"):
            g.set_spins(np.zeros(10, np.float32))
        with pytest.raises(pyro.PyroError, match="cos_cone"):
            g.set_cos_cone([1.0])


def test_full_size_c2_properties():
    """The c2 workload itself: L=16 x 64 replicas, T in [0.01, 1], 1 + 10 passes,
    swaps every 10 sweeps. After 50 sweeps every spin is a unit vector and the
    incremental FP64 energy equals the recomputed one."""
    from cudacode_amd.pyro import geometric_ladder
    temps = geometric_ladder(0.01, 1.0, 64)
    with gpu(16, temps, n_metro=1, n_or=10, pt_every=10, cos_cone=0.5) as g:
        g.Mc(50)
        s = g.get_spins().reshape(64, 3, -1).astype(np.float64)
        n = np.sqrt((s ** 2).sum(axis=1))
        assert np.abs(n - 1).max() < 1e-6
        E = g.TotalEnergyv2()
        assert np.allclose(g.energy_incremental(), E, rtol=1e-10, atol=1e-8)
        # low-T replicas are far below the high-T ones
        T, at = g.temperatures()
        e_per = E / (16 * 16 ** 3)
        assert e_per[at[-1]] < -0.9 and e_per[at[0]] > e_per[at[-1]]


def test_overrelaxation_only_conserves_energy_full_size():
    with gpu(16, [1.0] * 8, n_metro=0, n_or=10) as g:
        e0 = g.TotalEnergyv2()
        g.Mc(20)
        e1 = g.TotalEnergyv2()
        assert np.allclose(e0, e1, rtol=0, atol=65536 * 2e-6)
        assert np.allclose(g.energy_incremental(), e1, rtol=1e-11, atol=1e-8)


def test_large_lattice_properties():
    """L=64 (the c5 lattice, 4,194,304 spins per replica), 4 replicas."""
    with gpu(64, [0.02, 0.08, 0.2, 0.5], n_metro=1, n_or=3, pt_every=2, cos_cone=0.5) as g:
        g.Mc(4)
        s = g.get_spins().reshape(4, 3, -1)
        n = np.sqrt((s.astype(np.float64) ** 2).sum(axis=1))
        assert np.abs(n - 1).max() < 1e-6
        assert np.allclose(g.energy_incremental(), g.TotalEnergyv2(), rtol=1e-10)


@pytest.mark.parametrize("L", [4, 8, 12, 16, 24, 32])
def test_kernel_variants_identical(L):
    """The cell-per-thread kernels (compile-time colour and L, XCD-remapped
    blocks) and the site-per-thread kernels give the same spins bit for bit,
    and both equal the oracle, for every lattice size with compiled copies
    below the full-size L=64 tests."""
    temps = [0.05, 0.3, 1.5]
    kw = dict(n_metro=1, n_or=3, pt_every=2, cos_cone=0.4)
    with gpu(L, temps, variant=0, **kw) as a, gpu(L, temps, variant=1, **kw) as b:
        a.Mc(6)
        b.Mc(6)
        sa, sb = a.get_spins(), b.get_spins()
        assert np.array_equal(sa, sb)
        assert np.array_equal(a.acceptance()[0], b.acceptance()[0])
    o = orc(L, temps, **kw)
    o.run(6)
    assert np.array_equal(sb, o.spins)


def test_untracked_overrelaxation_drift_is_rounding_only():
    """Without FP64 bookkeeping in over-relaxation (the fast default), the
    incremental energy still follows the recomputed one to FP32 rounding:
    100 sweeps of 1 + 10 passes on 4 x 65,536 spins stay within 1e-6 relative."""
    temps = [0.02, 0.1, 0.4, 1.0]
    with gpu(16, temps, track=False, n_metro=1, n_or=10, cos_cone=0.5) as g:
        g.Mc(100)
        E = g.TotalEnergyv2()
        inc = g.energy_incremental()
        assert np.allclose(inc, E, rtol=1e-6, atol=1e-3)


def test_exchange_reanchors_incremental_energy():
    """The energies computed for each swap re-anchor the incremental energy:
    right after a swap it equals the recomputed energy exactly, so untracked
    over-relaxation rounding cannot build up past one exchange interval."""
    temps = [0.02, 0.1, 0.4, 1.0]
    with gpu(16, temps, track=False, n_metro=1, n_or=10, pt_every=10, cos_cone=0.5) as g:
        g.Mc(20)
        assert np.array_equal(g.energy_incremental(), g.TotalEnergyv2())
        g.Mc(3)
        assert np.allclose(g.energy_incremental(), g.TotalEnergyv2(), rtol=1e-6, atol=1e-3)


@pytest.mark.parametrize("variant", [0, 1])
def test_post_rotation_matches_oracle(variant):
    """The optional rotation about h after each Metropolis test: same spins
    as the oracle bit for bit, unchanged acceptance counting, exact books."""
    L, temps = 4, [0.1, 0.5, 2.0]
    kw = dict(n_metro=1, n_or=1, pt_every=4, cos_cone=0.4)
    with gpu(L, temps, variant=variant, **kw) as g:
        g.set_post_rotate(True)
        g.Mc(8)
        spins, acc = g.get_spins(), g.acceptance()[0]
        assert np.allclose(g.energy_incremental(), g.TotalEnergyv2(), rtol=1e-11, atol=1e-9)
    oracle.set_post_rotate(True)
    try:
        o = orc(L, temps, **kw)
        o.run(8)
    finally:
        oracle.set_post_rotate(False)
    assert np.array_equal(spins, o.spins)
    assert np.array_equal(acc, o.accepted)


@pytest.mark.parametrize("variant,R,graph,pt_every,track,streams", [
    (0, 5, True, 3, True, 2), (1, 4, False, 3, False, 2), (0, 2, False, 3, True, 2),
    (1, 4, True, 0, False, 2), (1, 3, False, 0, True, 2), (1, 5, True, 3, False, 2),
    (1, 4, False, 3, False, 4), (0, 5, True, 3, True, 4), (1, 7, True, 3, False, 4),
    (1, 3, False, 0, True, 3), (0, 5, False, 3, False, 3)])
def test_split_sweeps_identical(variant, R, graph, pt_every, track, streams):
    """Sweeps whose replica groups run on 2, 3 or 4 streams (joined only at
    exchanges, measurements and the end of a call) equal one-stream sweeps
    bit for bit: spins, rungs, acceptance, incremental energies, sweep
    counters, the energy series (including rows past its capacity) and the
    measurement rows; captured or eager, odd replica counts and uneven groups,
    with and without exchanges, calls that end mid-interval; and equal the
    oracle. Incremental
    energies (and so the series) agree to FP64 rounding only: each block adds
    its partial sum with one atomic, in whatever order the blocks finish.
    Untracked over-relaxation in the cell kernels writes the series from its
    first launch of a sweep; the other cases use a separate launch."""
    L = 8
    temps = list(np.geomspace(0.05, 2.0, R))
    kw = dict(n_metro=1, n_or=3, pt_every=pt_every, cos_cone=0.4)
    with gpu(L, temps, variant=variant, track=track, **kw) as a, \
            gpu(L, temps, variant=variant, track=track, **kw) as b:
        a.set_streams(streams)
        b.set_streams(1)
        assert a.streams() == streams
        for g in (a, b):
            g.record_start(37)
            g.measure_start(4, 12)
            g.Mc(5, use_graph=graph)
            g.Mc(60, use_graph=graph)
        assert np.array_equal(a.get_spins(), b.get_spins())
        assert np.array_equal(a.temperatures()[1], b.temperatures()[1])
        assert np.array_equal(a.acceptance()[0], b.acceptance()[0])
        assert np.allclose(a.energy_incremental(), b.energy_incremental(), rtol=1e-11, atol=0)
        assert a.sweep_counters() == b.sweep_counters()
        ra, rb = a.record_fetch(), b.record_fetch()
        say(f"series max |diff| {np.abs(ra - rb).max()}")
        assert ra.shape == (37, R) and np.allclose(ra, rb, rtol=1e-11, atol=0)
        ma, mb = a.measure_fetch(), b.measure_fetch()
        assert ma.shape[0] == 12 and np.array_equal(ma, mb)
        spins, at = a.get_spins(), a.temperatures()[1]
    o = orc(L, temps, **kw)
    o.run(65)
    assert np.array_equal(spins, o.spins)
    assert np.array_equal(at, o.replica_at)


def test_two_stream_sweeps_identical_matrix_model():
    """The same for the anisotropic exchange model in a field."""
    from cudacode_amd import pyro
    L, temps = 4, [0.1, 0.4, 1.6]
    C = pyro.exchange_matrices(*pyro.YB2TI2O7_J)
    B = pyro.zeeman_fields([0.0, 0.0, 0.5])
    kw = dict(n_metro=1, n_or=2, pt_every=2, cos_cone=0.4)
    with gpu(L, temps, **kw) as a, gpu(L, temps, **kw) as b:
        for g in (a, b):
            g.set_couplings(C, B)
        b.set_streams(1)
        a.Mc(8)
        b.Mc(8)
        assert np.array_equal(a.get_spins(), b.get_spins())
        assert np.allclose(a.energy_incremental(), b.energy_incremental(), rtol=1e-11, atol=0)


def test_stream_default_and_autotune_on_uniform_lattice():
    """Replica groups by default: 3 when a third of the replicas spans >= 256
    workgroups, else 2 when half of them does, else 1; more streams than
    replicas are refused. The autotuner tries 1, 2 and 3 on a uniform
    lattice, keeps the state, and the sweeps after it equal an untuned ctx's."""
    with gpu(4, [0.2, 0.5, 1.0, 2.0]) as small:
        assert small.streams() == 1
    with gpu(4, [0.2, 0.5]) as two:
        with pytest.raises(RuntimeError):
            two.set_streams(3)
        with pytest.raises(RuntimeError):
            two.set_streams(5)
    with gpu(16, list(np.geomspace(0.05, 1.0, 64))) as big:
        assert big.streams() == 3  # 21 replicas x 16 workgroups in group 0
    temps = list(np.geomspace(0.05, 1.0, 32))
    kw = dict(n_metro=1, n_or=2, pt_every=2, cos_cone=0.4)
    with gpu(16, temps, **kw) as g, gpu(16, temps, **kw) as h:
        assert g.streams() == 2  # 16 replicas x 16 workgroups per half
        g.Mc(2)
        h.Mc(2)
        res = g.autotune(2)
        say(f"autotune (mode, block, variant, streams, groups, ms/sweep): {res.tolist()}")
        assert set(res[res[:, 0] == 0, 3].tolist()) == {1.0, 2.0, 3.0}
        blocked = res[(res[:, 0] == 0) & (res[:, 4] > 0)]
        assert blocked[:, 4].tolist() == [2, 4, 8, 16, 32] + [4, 8, 16, 32] * 2
        assert blocked[:, 3].tolist() == [1] * 5 + [2] * 4 + [3] * 4
        assert np.all(res[:, 5] > 0)
        assert np.array_equal(g.get_spins(), h.get_spins())
        g.Mc(4)
        h.Mc(4)
        assert np.array_equal(g.get_spins(), h.get_spins())
        assert np.allclose(g.energy_incremental(), h.energy_incremental(), rtol=1e-11, atol=0)


def test_autotune_tries_four_streams_with_the_whole_pass():
    """Where the whole over-relaxation pass runs (L=16, untracked), the
    autotuner also tries four replica groups; the count it keeps stays, and
    the sweeps after it equal an untuned ctx's."""
    temps = list(np.geomspace(0.05, 1.0, 64))
    kw = dict(n_metro=1, n_or=3, pt_every=2, cos_cone=0.4)
    with gpu(16, temps, track=False, **kw) as g, gpu(16, temps, track=False, **kw) as h:
        assert g.streams() == 3 and h.streams() == 3
        g.Mc(3)
        h.Mc(3)
        res = g.autotune(2)
        say(f"autotune (mode, block, variant, streams, groups, ms/sweep): {res.tolist()}")
        plain = res[(res[:, 0] == 0) & (res[:, 4] == 0)]
        assert plain[:, 3].tolist() == [1, 2, 3, 1, 2, 3, 4]
        assert plain[:, 2].tolist() == [1, 1, 1, 0, 0, 0, 1]
        best = res[np.argmin(res[:, 5])]
        assert (g.streams(), g.group_blocking()) == (int(best[3]), int(best[4]))
        g.Mc(4)
        h.Mc(4)
        assert np.array_equal(g.get_spins(), h.get_spins())
        assert g.sweep_counters() == h.sweep_counters()
        g.set_streams(4)
        h.set_streams(4)
        g.Mc(3)
        h.Mc(3)
        assert np.array_equal(g.get_spins(), h.get_spins())
    with gpu(16, temps, track=True, **kw) as t:
        res = t.autotune(2)
        assert res[(res[:, 0] == 0) & (res[:, 4] == 0), 3].tolist() == [1, 2, 3, 1, 2, 3]


def test_full_size_c2_spins_agree_with_oracle():
    """The L=16 x 64 parallel-tempering workload at its full size (T 0.01..1,
    1 Metropolis + 10 over-relaxation, swaps every 10 sweeps, the fast
    untracked kernels on three streams): after 100 sweeps the spins, the rungs
    and the acceptance counts equal the oracle's, bit for bit."""
    from cudacode_amd.pyro import geometric_ladder
    temps = geometric_ladder(0.01, 1.0, 64)
    kw = dict(n_metro=1, n_or=10, pt_every=10, cos_cone=0.5)
    with gpu(16, temps, track=False, **kw) as g:
        assert g.streams() == 3
        g.Mc(100)
        spins, at, acc = g.get_spins(), g.temperatures()[1], g.acceptance()[0]
        assert g.sweep_counters() == (100, 10)
    say("full c2 100 sweeps on the oracle")
    o = orc(16, temps, **kw)
    o.run(100)
    assert np.array_equal(at, o.replica_at)
    assert np.array_equal(acc, o.accepted)
    assert np.array_equal(spins, o.spins)


def test_full_size_large_lattice_replicas_agree_with_oracle():
    """The L=64 x 32 ensemble at its full size (1.61 GB of spins): replicas 0
    and 31 after 5 sweeps of 1 Metropolis + 10 over-relaxation equal the
    oracle's replay of each replica alone (its own Philox counters), bit for
    bit."""
    from cudacode_amd.pyro import geometric_ladder
    temps = geometric_ladder(0.02, 0.5, 32)
    kw = dict(n_metro=1, n_or=10, pt_every=0, cos_cone=0.5)
    with gpu(64, temps, track=False, **kw) as g:
        g.Mc(5)
        s = g.get_spins().reshape(32, -1)
        picks = {r: s[r].copy() for r in (0, 31)}
        del s
    for r, sr in picks.items():
        say(f"L=64 replica {r}: 5 sweeps on the oracle")
        oracle.set_replica_base(r)
        try:
            o = orc(64, [temps[r]], **kw)
            o.run(5)
        finally:
            oracle.set_replica_base(0)
        assert np.array_equal(sr, o.spins.reshape(-1))


def test_large_lattice_with_exchanges_agrees_with_oracle():
    """L=64 (the c5 lattice) with parallel tempering on: 4 replicas of the c5
    ladder's span, 1 Metropolis + 10 over-relaxation passes per sweep (the
    fused default kernels), exchanges every 2 sweeps, 6 sweeps on two replica
    groups: spins and rungs equal the oracle's bit for bit, energies to
    1e-12."""
    temps = [0.02, 0.06, 0.18, 0.5]
    kw = dict(n_metro=1, n_or=10, pt_every=2, cos_cone=0.5)
    with gpu(64, temps, track=False, **kw) as g:
        g.set_streams(2)
        g.Mc(6)
        spins, E, at = g.get_spins(), g.TotalEnergyv2(), g.temperatures()[1]
    say("L=64 x 4 with exchanges: 6 sweeps on the oracle")
    o = orc(64, temps, **kw)
    o.run(6)
    assert np.array_equal(at, o.replica_at)
    assert np.array_equal(spins, o.spins)
    assert np.allclose(E, o.observables()[:, 0], rtol=1e-12)


@pytest.mark.parametrize("L,n_or,track,streams,graph", [
    (4, 3, False, 1, True), (8, 4, True, 1, False), (8, 5, False, 2, True),
    (12, 2, True, 2, True), (16, 10, False, 2, False), (24, 3, False, 3, True),
    (32, 2, True, 2, False), (64, 3, False, 2, True), (64, 2, True, 1, False)])
def test_fused_overrelaxation_identical(L, n_or, track, streams, graph):
    """Over-relaxation passes fused two colours per launch (alternating
    between the two spin buffers, an odd last pass in place) equal one colour
    per launch bit for bit: spins, rungs, acceptance, counters, measurement
    rows; incremental energies and the series to FP64 rounding; and the
    oracle. The fused timing launches leave the state as it was."""
    R = 4
    temps = list(np.geomspace(0.05, 2.0, R))
    kw = dict(n_metro=1, n_or=n_or, pt_every=3, cos_cone=0.4)
    with gpu(L, temps, track=track, **kw) as a, gpu(L, temps, track=track, **kw) as b:
        b.set_or_fusion(False)
        for g in (a, b):
            g.set_streams(streams)
            g.record_start(20)
            g.measure_start(4, 8)
            g.Mc(3, use_graph=graph)
            g.Mc(17, use_graph=graph)
        assert np.array_equal(a.get_spins(), b.get_spins())
        assert np.array_equal(a.temperatures()[1], b.temperatures()[1])
        assert np.array_equal(a.acceptance()[0], b.acceptance()[0])
        assert a.sweep_counters() == b.sweep_counters()
        assert np.array_equal(a.measure_fetch(), b.measure_fetch())
        ra, rb = a.record_fetch(), b.record_fetch()
        assert ra.shape == (20, R) and np.allclose(ra, rb, rtol=1e-11, atol=0)
        assert np.allclose(a.energy_incremental(), b.energy_incremental(), rtol=1e-11, atol=0)
        spins = a.get_spins()
        a.time_phase(3, 6)
        assert np.array_equal(a.get_spins(), spins)
    o = orc(L, temps, **kw)
    o.run(20)
    assert np.array_equal(spins, o.spins)


@pytest.mark.parametrize("n_or,streams,graph,R", [
    (10, 3, False, 6), (4, 2, True, 4), (3, 1, True, 3), (5, 3, True, 7)])
def test_whole_pass_overrelaxation_identical(n_or, streams, graph, R):
    """At L=16, over-relaxation passes of one launch each (colours 0-3 of a
    z-plane per block, plane z + 1's first two colours recomputed where plane
    z's last two read them) equal passes of two launches and of four, bit for
    bit: spins, rungs, acceptance, counters, measurement rows; the series and
    incremental energies to FP64 rounding; and the oracle. Timing launches
    (kind 5) leave the state as it was; with over-relaxation energy tracking
    on, or at another size, the whole pass is not used and kind 5 is refused."""
    from cudacode_amd import pyro
    temps = list(np.geomspace(0.02, 2.0, R))
    kw = dict(n_metro=1, n_or=n_or, pt_every=2, cos_cone=0.4)
    runs = []
    for level in (2, 1, 0):
        with gpu(16, temps, track=False, **kw) as g:
            g.set_or_fusion(level)
            g.set_streams(streams)
            g.record_start(12)
            g.measure_start(3, 4)
            g.Mc(2, use_graph=graph)
            g.Mc(10, use_graph=graph)
            runs.append((g.get_spins(), g.temperatures()[1], g.acceptance()[0],
                         g.sweep_counters(), g.measure_fetch(), g.record_fetch(),
                         g.energy_incremental()))
            if level == 2:
                spins = g.get_spins()
                avg, _ = g.time_phase(5, 4)
                assert avg > 0 and np.array_equal(g.get_spins(), spins)
            else:
                with pytest.raises(pyro.PyroError):
                    g.time_phase(5, 1)
    (s2, t2, a2, c2, m2, r2, e2) = runs[0]
    for (s, t, a, c, m, r, e) in runs[1:]:
        assert np.array_equal(s2, s)
        assert np.array_equal(t2, t)
        assert np.array_equal(a2, a)
        assert c2 == c
        assert np.array_equal(m2, m)
        assert np.allclose(r2, r, rtol=1e-11, atol=0)
        assert np.allclose(e2, e, rtol=1e-11, atol=0)
    o = orc(16, temps, **kw)
    o.run(12)
    assert np.array_equal(s2, o.spins)
    assert np.array_equal(t2, o.replica_at)
    with gpu(16, temps, track=True, **kw) as g:
        with pytest.raises(pyro.PyroError):
            g.time_phase(5, 1)
    with gpu(8, temps, track=False, **kw) as g:
        with pytest.raises(pyro.PyroError):
            g.time_phase(5, 1)
        with pytest.raises(pyro.PyroError):
            g.set_or_fusion(3)


@pytest.mark.parametrize("R,pt_every,graph,streams", [
    (2, 0, False, 1), (4, 2, True, 2), (3, 2, False, 3)])
def test_whole_pass_at_l64_identical(R, pt_every, graph, streams):
    """At L=64 the whole over-relaxation pass runs in strips of 4 rows of 64
    cells (a row per wave; the rows below and above and plane z + 1's first
    two colours recomputed where the strip reads them): spins, rungs and
    counters equal two colours per launch bit for bit, the series and the
    incremental energies to FP64 rounding; a timing launch (kind 5) leaves
    the state as it was."""
    temps = list(np.geomspace(0.05, 1.5, R))
    kw = dict(n_metro=1, n_or=3, pt_every=pt_every, cos_cone=0.4)
    runs = []
    for level in (2, 1):
        with gpu(64, temps, track=False, **kw) as g:
            g.set_or_fusion(level)
            g.set_streams(streams)
            g.record_start(6)
            g.Mc(2, use_graph=graph)
            if level == 2:
                spins = g.get_spins()
                avg, _ = g.time_phase(5, 2)
                assert avg > 0 and np.array_equal(g.get_spins(), spins)
            g.Mc(4, use_graph=graph)
            runs.append((g.get_spins(), g.temperatures()[1], g.sweep_counters(), g.record_fetch(),
                         g.energy_incremental(), g.TotalEnergyv2()))
    (s2, t2, c2, r2, e2, f2), (s1, t1, c1, r1, e1, _) = runs
    assert np.array_equal(s2, s1)
    assert np.array_equal(t2, t1) and c2 == c1
    assert np.allclose(r2, r1, rtol=1e-11, atol=0)
    assert np.allclose(e2, e1, rtol=1e-11, atol=0)
    assert np.allclose(e2, f2, rtol=1e-8, atol=0)


@pytest.mark.parametrize("streams,graph", [(1, True), (2, True), (2, False)])
def test_fused_metropolis_identical(streams, graph):
    """Metropolis phases fused two colours per launch in place (L=16, one
    z-plane per block, the neighbours' first-colour values exchanged between
    lanes) equal one colour per launch bit for bit: spins, rungs, acceptance
    and counters over sweeps of 2 Metropolis + 1 over-relaxation passes with
    exchanges; incremental energies to FP64 rounding (the block sums run over
    both colours). Timing launches of the pair (kind 4) advance the state as
    four one-colour launches do; with the post-Metropolis rotation on, the
    sweeps fall back to one colour per launch."""
    R = 6
    temps = list(np.geomspace(0.05, 2.0, R))
    kw = dict(n_metro=2, n_or=1, pt_every=2, cos_cone=0.4)
    with gpu(16, temps, **kw) as a, gpu(16, temps, **kw) as b:
        b.set_or_fusion(False)
        for g in (a, b):
            g.set_streams(streams)
            g.Mc(9, use_graph=graph)
        assert np.array_equal(a.get_spins(), b.get_spins())
        assert np.array_equal(a.temperatures()[1], b.temperatures()[1])
        assert np.array_equal(a.acceptance()[0], b.acceptance()[0])
        assert a.acceptance()[0].sum() > 0
        assert a.sweep_counters() == b.sweep_counters()
        assert np.allclose(a.energy_incremental(), b.energy_incremental(), rtol=1e-11, atol=0)
        assert np.allclose(a.energy_incremental(), a.TotalEnergyv2(), rtol=1e-10, atol=1e-8)
        a.time_phase(4, 2)
        b.time_phase(0, 4)
        assert np.array_equal(a.get_spins(), b.get_spins())
        assert np.array_equal(a.acceptance()[0], b.acceptance()[0])
        for g in (a, b):
            g.set_post_rotate(True)
            g.Mc(3, use_graph=graph)
        assert np.array_equal(a.get_spins(), b.get_spins())
        with pytest.raises(RuntimeError):
            a.time_phase(4, 2)


@pytest.mark.parametrize("L,groups,graph,track,streams", [
    (8, 2, True, True, 1), (8, 3, False, False, 1), (16, 4, True, False, 1),
    (8, 8, True, False, 1), (8, 4, True, False, 2), (16, 8, False, False, 3)])
def test_group_blocked_sweeps_identical(L, groups, graph, track, streams):
    """Cache-blocked sweeps (each exchange or measurement interval run replica
    group after group on one stream, the draws of the interval's sweeps keyed
    as in an unblocked run) equal unblocked sweeps bit for bit: spins, rungs,
    acceptance, counters and measurement rows; the series and incremental
    energies to FP64 rounding; and the oracle. One group in flight at a time,
    or one per stream. Calls that end inside an interval and graph blocks
    both cut intervals."""
    from cudacode_amd import pyro
    R = 8
    temps = list(np.geomspace(0.05, 2.0, R))
    kw = dict(n_metro=1, n_or=3, pt_every=3, cos_cone=0.4)
    out = []
    for G in (groups, 0):
        with gpu(L, temps, track=track, **kw) as g:
            g.set_streams(streams if G else 1)
            g.set_group_blocking(G)
            assert g.group_blocking() == (G if G >= 2 else 0)
            g.record_start(20)
            g.measure_start(4, 8)
            g.Mc(5, use_graph=graph)
            g.Mc(15, use_graph=graph)
            out.append((g.get_spins(), g.temperatures()[1], g.acceptance()[0], g.sweep_counters(),
                        g.measure_fetch(), g.record_fetch(), g.energy_incremental()))
    (s1, t1, a1, c1, m1, r1, e1), (s0, t0, a0, c0, m0, r0, e0) = out
    assert np.array_equal(s1, s0)
    assert np.array_equal(t1, t0)
    assert np.array_equal(a1, a0)
    assert c1 == c0 == (20, 6)
    assert np.array_equal(m1, m0)
    assert r1.shape == (20, R) and np.allclose(r1, r0, rtol=1e-11, atol=0)
    assert np.allclose(e1, e0, rtol=1e-11, atol=0)
    o = orc(L, temps, **kw)
    o.run(20)
    assert np.array_equal(s1, o.spins)
    assert np.array_equal(t1, o.replica_at)


def test_group_blocking_refusals():
    from cudacode_amd import pyro
    with gpu(4, [0.5, 1.0, 2.0]) as g:
        with pytest.raises(pyro.PyroError):
            g.set_group_blocking(4)  # more groups than replicas
        with pytest.raises(pyro.PyroError):
            g.set_group_blocking(-1)
        g.set_group_blocking(1)
        assert g.group_blocking() == 0


def test_graph_block_sizes():
    """Sweeps per captured graph: a whole number of exchange and measurement
    intervals, at least 16; a call that starts on a block and spans whole
    blocks equals the same sweeps run eagerly."""
    with gpu(4, [0.5, 1.0], n_metro=1, n_or=2, pt_every=10) as g:
        assert g.graph_block() == 20
        g.measure_start(4, 8)
        assert g.graph_block() == 20
    with gpu(4, [0.5, 1.0], n_metro=1, n_or=2, pt_every=3) as g:
        g.measure_start(4, 8)
        assert g.graph_block() == 24
    with gpu(4, [0.5, 1.0], n_metro=1, n_or=2, pt_every=0) as g, \
            gpu(4, [0.5, 1.0], n_metro=1, n_or=2, pt_every=0) as h:
        assert g.graph_block() == 16
        g.Mc(32, use_graph=True)
        h.Mc(32, use_graph=False)
        assert np.array_equal(g.get_spins(), h.get_spins())
