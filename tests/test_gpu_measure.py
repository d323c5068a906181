"""Measurement series, simulated annealing and the run protocol, through the
C ABI on the GPU.

- a measurement row holds, per rung, the full FP64 energy and the moduli of
  the 5 order-parameter groups of the replica at that rung: checked against
  the host fold of the sublattice sums and against the oracle's energies
  (1e-12 relative: different summation order);
- rows written inside captured sweeps equal rows of eager sweeps, bit for
  bit, with exchanges and measurements on different cadences;
- annealing follows the reference schedule T0 - (T0 - T) m / steps
  (MonteCarlo.cu:88-92) with no exchanges: the spins equal the oracle's run
  through the same beta schedule, bit for bit, and the exchange cadence
  resumes afterwards;
- the protocol writes the reference's files.
"""
import os
import sys

import numpy as np
import pytest

from oracle import oracle

pytestmark = pytest.mark.gpu

SEED = 0xA11EA1


def say(msg):
    print(f"[gpu-measure] {msg}", flush=True)
    sys.stdout.flush()


def ctx(L, temps, **kw):
    from cudacode_amd import pyro
    say(f"create L={L} R={len(temps)} {kw}")
    return pyro.Pyrochlore(L, list(temps), seed=kw.pop("seed", SEED), **kw)


def test_rows_match_recomputed_observables():
    from cudacode_amd import pyro
    L, temps = 4, [0.25, 0.6, 1.3, 2.9]
    kw = dict(n_metro=1, n_or=2, cos_cone=0.5)
    with ctx(L, temps, **kw) as g:
        g.measure_start(3, 10)
        g.Mc(9)
        rows = g.measure_fetch()
        assert rows.shape == (3, 4, pyro.MEAS_Q)
        T, at = g.temperatures()
        E = g.TotalEnergyv2()[at]
        op = g.DoMeasurement()[at]
        assert np.array_equal(rows[-1, :, 0], E)
        mods = np.stack([np.abs(op[:, 0]), np.linalg.norm(op[:, 1:3], axis=1),
                         np.linalg.norm(op[:, 3:6], axis=1), np.linalg.norm(op[:, 6:9], axis=1),
                         np.linalg.norm(op[:, 9:12], axis=1)], axis=1)
        assert np.allclose(rows[-1, :, 1:], mods, rtol=1e-12, atol=1e-15)
        M = g.Magnetization()[at] / g.N
        assert np.allclose(rows[-1, :, 4], np.linalg.norm(M, axis=1), rtol=1e-12)
        g.measure_stop()
        g.Mc(6)
        assert g.measure_fetch().shape[0] == 3  # stopped: no new rows
    o = oracle.OracleRun(L, temps, seed=SEED, **kw)
    o.run(9)
    e_orc = o.observables()[:, 0][o.replica_at]
    assert np.allclose(rows[-1, :, 0], e_orc, rtol=1e-12)


def test_captured_rows_equal_eager_rows():
    L, temps = 4, [0.2, 0.4, 0.8, 1.6, 3.2]
    kw = dict(n_metro=1, n_or=1, pt_every=5, cos_cone=0.4)
    with ctx(L, temps, **kw) as a, ctx(L, temps, **kw) as b:
        for g in (a, b):
            g.measure_start(2, 64)
        a.Mc(40)
        b.Mc(40, use_graph=False)
        ra, rb = a.measure_fetch(), b.measure_fetch()
        assert ra.shape[0] == 20
        assert np.array_equal(ra, rb)
        assert np.array_equal(a.get_spins(), b.get_spins())
        assert np.array_equal(a.temperatures()[1], b.temperatures()[1])
        assert a.sweep_counters() == b.sweep_counters() == (40, 8)


def test_anneal_follows_reference_schedule():
    L, temps = 4, [0.3, 0.7, 1.5]
    kw = dict(n_metro=1, n_or=1, pt_every=3, cos_cone=0.3)
    steps, per, T0 = 4, 2, 5.0
    with ctx(L, temps, **kw) as g:
        g.SimulatedAnnealing(steps, per, T0)
        s_gpu = g.get_spins()
        assert g.sweep_counters() == (steps * per, 0)
        assert np.array_equal(g.TotalEnergyv2(), g.energy_incremental())
        g.Mc(6)
        s_gpu2 = g.get_spins()
        at_gpu = g.temperatures()[1]
    o = oracle.OracleRun(L, temps, seed=SEED, **kw)
    Tt = np.array(temps, np.float64)
    dT = 1.0 / steps
    o.pt_every = 0
    for m in range(steps):
        o.beta = (1.0 / (T0 - (T0 - Tt) * m * dT)).astype(np.float32)
        o.run(per)
    o.beta = (1.0 / Tt).astype(np.float32)
    o.pt_every = 3
    assert np.array_equal(s_gpu, o.spins)
    o.run(6)
    assert np.array_equal(s_gpu2, o.spins)
    assert np.array_equal(at_gpu, o.replica_at)


def test_protocol_writes_reference_files(tmp_path):
    from cudacode_amd import measure
    L, temps = 4, [0.3, 0.6, 1.2, 2.4]
    with ctx(L, temps, n_metro=1, n_or=2, pt_every=4, cos_cone=0.5) as g:
        s = measure.run_protocol(g, str(tmp_path), rep=0, anneal_steps=5, anneal_sweeps=2,
                                 pt_rounds=2, thermalisation=10, n_measures=40,
                                 sweeps_per_measure=2, save_every=20, store_confs=2,
                                 confs_every=10, n_bins=8)
        N = g.N
    names = sorted(p.name for p in tmp_path.iterdir())
    for k in range(4):
        assert f"InitialConf{k}.dat" in names and f"Conf{k}.dat" in names
        assert os.path.getsize(tmp_path / f"InitialConf{k}.dat") == 3 * N * 8
        assert os.path.getsize(tmp_path / f"Conf{k}.dat") == 2 * 3 * N * 8
    e = measure.read_table(str(tmp_path / "Energies0.dat"))
    assert np.allclose(e[:, 0], sorted(temps, reverse=True))
    assert np.all(e[:, 8] == 40)
    assert np.allclose(e[:, 1], s.e[:, 0])
    # E/N rises with T; the coldest rung sits near the ground-state value -1.
    assert np.all(np.diff(e[::-1, 1]) > 0)
    assert -1.0 < e[-1, 1] < -0.6
    op = measure.read_table(str(tmp_path / "OrderParameters0.dat"))
    assert op.shape == (4, 21)


def test_batch_rows_and_protocol(tmp_path):
    """Mixed lattice sizes: measurement rows follow each rung's replica (its
    own N for the order parameters), and the protocol runs on a batch
    (configuration files are skipped: they hold one lattice size)."""
    from cudacode_amd import measure, pyro
    Ls = [4, 8, 4, 8]
    temps = [0.3, 0.3, 1.2, 1.2]
    with pyro.Batch(Ls, temps, seed=SEED, n_metro=1, n_or=1, pt_every=3, cos_cone=0.5) as g:
        g.measure_start(3, 8)
        g.Mc(6)
        rows = g.measure_fetch()
        assert rows.shape == (2, 4, pyro.MEAS_Q)
        g.measure_stop()
        g.set_passes(1, 1)
        _, at = g.temperatures()
        g.measure_start(1, 4)
        g.Mc(1)  # sweep 7: no exchange after it, so rows see the current rungs
        last = g.measure_fetch()[-1]
        E = g.TotalEnergyv2()[at]
        assert np.array_equal(last[:, 0], E)
        M = g.Magnetization()[at] / np.asarray(g.N)[at][:, None]
        assert np.allclose(last[:, 4], np.linalg.norm(M, axis=1), rtol=1e-12)
        g.measure_stop()
        s = measure.run_protocol(g, str(tmp_path), anneal_steps=2, anneal_sweeps=2, pt_rounds=1,
                                 thermalisation=3, n_measures=8, sweeps_per_measure=3,
                                 save_every=8, store_confs=1, confs_every=4, n_bins=4)
    names = sorted(p.name for p in tmp_path.iterdir())
    assert names == ["Energies0.dat", "Magnetization0.dat", "OrderParameters0.dat", "run0.jsonl"]
    assert s.n == 8 and np.all(np.isfinite(s.e[:, 0]))
