"""GPU parity of the reference's other update schedule, through the C ABI:
phases on independent sets drawn from the 24 (Pyrochlore.cu:848,
IndependentSubLattices[Update[0] % 24]) and over-relaxation by a random
angle about the local field (OverRelaxationStep, Pyrochlore.cu:1009-1045).

Bar: spins, rungs and acceptance bit-identical to the oracle run with the
same schedule (both sides evaluate the same float expressions on the same
Philox words); FP64 energies to 1e-12 relative; one-stream == split-stream
and captured == eager bit for bit; the options refused where they do not run.
"""
import sys

import numpy as np
import pytest

from oracle import oracle

pytestmark = pytest.mark.gpu

SEED = 0x51ED5E7


def say(msg):
    print(f"[gpu-schedule] {msg}", flush=True)
    sys.stdout.flush()


class schedule:
    """The oracle's schedule switch for the duration of a block."""

    def __init__(self, sets, angle):
        self.sets, self.angle = sets, angle

    def __enter__(self):
        oracle.set_update_schedule(self.sets, self.angle)

    def __exit__(self, *a):
        oracle.set_update_schedule(False, False)


def gpu(L, temps, sets, angle, variant=1, track=True, **kw):
    from cudacode_amd import pyro
    say(f"create L={L} R={len(temps)} sets={sets} angle={angle} variant={variant} {kw}")
    g = pyro.Pyrochlore(L, temps, seed=SEED, **kw)
    g.set_kernel_options(variant, track)
    g.set_update_schedule(sets, angle)
    return g


@pytest.mark.parametrize("L,sets,angle,variant", [
    (4, 1, False, 1), (4, 0, True, 1), (4, 1, True, 0), (6, 1, True, 1),
    (8, 1, True, 1), (16, 1, False, 1), (4, 2, False, 1), (6, 2, True, 1), (16, 2, True, 1)])
def test_schedule_matches_oracle(L, sets, angle, variant):
    """Drawn sets (1), drawn colour orders (2) and/or random-angle
    over-relaxation with exchanges: the
    oracle's spins, rungs and acceptance bit for bit, FP64 energies to 1e-12
    and the incremental energy equal to the recomputed one, at compiled and
    runtime lattice sizes and from either kernel-variant setting."""
    temps = [0.08, 0.3, 0.9, 2.5]
    kw = dict(n_metro=1, n_or=2, pt_every=3, cos_cone=0.35)
    n = 10
    with gpu(L, temps, sets, angle, variant=variant, **kw) as g:
        g.Mc(n)
        spins, E = g.get_spins(), g.TotalEnergyv2()
        inc, acc, at = g.energy_incremental(), g.acceptance()[0], g.temperatures()[1]
    with schedule(sets, angle):
        o = oracle.OracleRun(L, temps, seed=SEED, **kw)
        o.run(n)
    assert np.array_equal(spins, o.spins)
    assert np.array_equal(acc, o.accepted)
    assert np.array_equal(at, o.replica_at)
    assert np.allclose(E, o.observables()[:, 0], rtol=1e-12)
    assert np.allclose(inc, E, rtol=1e-11, atol=1e-9)


@pytest.mark.parametrize("sets", [1, 2])
def test_schedule_with_two_metropolis_passes_matches_oracle(sets):
    """Two Metropolis and three over-relaxation passes a sweep: the pass index
    keys the set or colour-order draws of both stages; spins, rungs and
    acceptance equal the oracle's bit for bit."""
    L, temps = 6, [0.1, 0.4, 1.5]
    kw = dict(n_metro=2, n_or=3, pt_every=4, cos_cone=0.4)
    with gpu(L, temps, sets, True, **kw) as g:
        g.Mc(8)
        spins, acc, at = g.get_spins(), g.acceptance()[0], g.temperatures()[1]
    with schedule(sets, True):
        o = oracle.OracleRun(L, temps, seed=SEED, **kw)
        o.run(8)
    assert np.array_equal(spins, o.spins)
    assert np.array_equal(acc, o.accepted)
    assert np.array_equal(at, o.replica_at)


def test_schedule_matrix_model_matches_oracle():
    """The same with the anisotropic exchange in a field (each drawn site
    takes its own basis couplings and Zeeman term)."""
    from cudacode_amd import pyro
    L, temps = 4, [0.05, 0.3, 1.2]
    kw = dict(n_metro=1, n_or=2, pt_every=3, cos_cone=0.3)
    C = pyro.exchange_matrices(*pyro.YB2TI2O7_J)
    B = pyro.zeeman_fields([0.4, 0.1, 0.9], 4.18, 1.77)
    with gpu(L, temps, True, True, spin_len2=0.25, **kw) as g:
        g.set_couplings(C, B)
        _, coup = g.couplings()
        g.Mc(8)
        spins, E, acc = g.get_spins(), g.TotalEnergyv2(), g.acceptance()[0]
    with schedule(True, True):
        o = oracle.OracleRun(L, temps, seed=SEED, spin_len2=0.25, **kw)
        oracle.set_couplings(coup)
        try:
            o.anchor = o.observables()[:, 0].copy()
            o.run(8)
            E_o = o.observables()[:, 0]
        finally:
            oracle.set_couplings(None)
    assert np.array_equal(spins, o.spins)
    assert np.array_equal(acc, o.accepted)
    assert np.allclose(E, E_o, rtol=1e-12)


@pytest.mark.parametrize("streams,graph,sets", [(2, True, 1), (3, False, 1), (3, True, 2)])
def test_schedule_split_streams_and_graphs_identical(streams, graph, sets):
    """Replica groups on several streams draw the same set for the same
    phase: split sweeps equal one-stream sweeps bit for bit, captured or
    eager."""
    L, temps = 8, [0.1, 0.2, 0.4, 0.8, 1.6]
    kw = dict(n_metro=1, n_or=3, pt_every=4, cos_cone=0.4)
    out = []
    for n_streams, use_graph in ((1, True), (streams, graph)):
        with gpu(L, temps, sets, True, **kw) as g:
            g.set_streams(n_streams)
            g.Mc(12, use_graph=use_graph)
            out.append((g.get_spins(), g.acceptance()[0], g.temperatures()[1],
                        g.energy_incremental()))
    (s1, a1, t1, e1), (s2, a2, t2, e2) = out
    assert np.array_equal(s1, s2)
    assert np.array_equal(a1, a2)
    assert np.array_equal(t1, t2)
    assert np.array_equal(e1, e2)


def test_random_angle_overrelaxation_full_size_properties():
    """At L=16 x 8: over-relaxation-only sweeps by a random angle keep the
    energy to FP32 rounding and the spins unit; the incremental energy
    follows the recomputed one."""
    L, temps = 16, list(np.geomspace(0.01, 1.0, 8))
    with gpu(L, temps, True, True, n_metro=0, n_or=5, pt_every=0) as g:
        e0 = g.TotalEnergyv2()
        g.Mc(6)
        e1 = g.TotalEnergyv2()
        assert np.allclose(e0, e1, rtol=0, atol=16 * L ** 3 * 2e-6)
        assert np.allclose(g.energy_incremental(), e1, rtol=1e-11, atol=1e-8)
        bad, _ = g.check_spins(1e-5)
        assert bad == 0


def test_schedule_refused_outside_execution_mode_0():
    from cudacode_amd import pyro
    say("batch ctx refuses the schedule options; bad schedule values refused")
    with pyro.Batch([4, 8], [0.5, 1.0], seed=SEED) as b:
        with pytest.raises(pyro.PyroError):
            b.set_update_schedule(True, False)
    with pyro.Pyrochlore(4, [0.5], seed=SEED) as g:
        with pytest.raises(pyro.PyroError):
            g.set_update_schedule(3, False)
