"""Measurement statistics and the result files, on synthetic rows (CPU).

The rows have the layout the device writes (n, R, 6): E, |m_A2|, |m_E|,
|m_T1|, |m_F|, |m_T2| per rung. Checks: means, the specific heat against its
definition on an exactly known series, binning errors that shrink like
1/sqrt(n) for independent data, tau_int of an AR(1) series against its closed
form, and that the three files carry the same numbers back.
"""
import numpy as np
import pytest

from cudacode_amd import measure, stats


def synthetic_rows(n=4096, R=3, seed=1):
    rng = np.random.default_rng(seed)
    rows = np.empty((n, R, 6))
    for k in range(R):
        rows[:, k, 0] = -1000.0 * (k + 1) + rng.normal(0, 5.0 * (k + 1), n)
        rows[:, k, 1:] = np.abs(rng.normal(0.1 * (k + 1), 0.01, (n, 5)))
    return rows


def test_summary_means_and_specific_heat():
    rows = synthetic_rows()
    T = np.array([2.0, 1.0, 0.5])
    N = 1024
    s = measure.summarise(rows, T, N, n_bins=32)
    for k in range(3):
        E = rows[:, k, 0]
        assert s.e[k, 0] == pytest.approx(E.mean() / N, rel=1e-14)
        assert s.e2[k, 0] == pytest.approx(((E / N) ** 2).mean(), rel=1e-14)
        c_direct = (np.mean(E * E) - E.mean() ** 2) / (N * T[k] ** 2)
        # jackknife removes only the O(1/n) bias of the plug-in variance
        assert s.c[k, 0] == pytest.approx(c_direct, rel=2e-3)
        assert s.c[k, 1] > 0
        for g in range(5):
            assert s.m[k, g, 0] == pytest.approx(rows[:, k, 1 + g].mean(), rel=1e-14)
            assert s.m2[k, g, 0] == pytest.approx((rows[:, k, 1 + g] ** 2).mean(), rel=1e-14)


def test_specific_heat_of_two_level_series_is_exact():
    """E alternating between 0 and 2 (mean 1, variance 1): C/N = 1/(N T^2)
    with zero jackknife error, since every bin holds the same multiset."""
    n, N, T = 1024, 8.0, 0.5
    E = np.tile([0.0, 2.0], n // 2)
    rows = np.zeros((n, 1, 6))
    rows[:, 0, 0] = E
    s = measure.summarise(rows, [T], N, n_bins=16)
    assert s.c[0, 0] == pytest.approx(1.0 / (N * T * T), rel=1e-12)
    assert s.c[0, 1] == pytest.approx(0.0, abs=1e-12)


def test_binning_error_scales_for_independent_data():
    rng = np.random.default_rng(3)
    x = rng.normal(0, 1, 1 << 16)
    err = stats.binning_error(x, 64)
    assert err == pytest.approx(1 / np.sqrt(x.size), rel=0.25)


def test_tau_int_of_ar1_matches_closed_form():
    """AR(1) x_t = a x_{t-1} + noise: tau_int = (1 + a) / (2 (1 - a))."""
    rng = np.random.default_rng(5)
    a, n = 0.8, 200_000
    x = np.empty(n)
    x[0] = 0.0
    eps = rng.normal(size=n)
    for t in range(1, n):
        x[t] = a * x[t - 1] + eps[t]
    tau, _, err = stats.tau_int(x)
    assert tau == pytest.approx((1 + a) / (2 * (1 - a)), abs=4 * err + 0.05)


def test_files_round_trip(tmp_path):
    rows = synthetic_rows(n=512, R=4)
    T = np.array([4.0, 2.0, 1.0, 0.5])
    s = measure.summarise(rows, T, 256)
    paths = measure.write_all(str(tmp_path), 3, s)
    assert sorted(p.name for p in tmp_path.iterdir()) == [
        "Energies3.dat", "Magnetization3.dat", "OrderParameters3.dat"]
    e = measure.read_table(paths["energies"])
    assert e.shape == (4, 9)
    assert np.allclose(e[:, 0], T)
    assert np.allclose(e[:, 1], s.e[:, 0], rtol=1e-11)
    assert np.allclose(e[:, 5], s.c[:, 0], rtol=1e-11)
    assert np.all(e[:, 8] == 512)
    op = measure.read_table(paths["order_parameters"])
    assert op.shape == (4, 21)
    assert np.allclose(op[:, 1:11:2], s.m[:, :, 0], rtol=1e-11)
    assert np.allclose(op[:, 11:21:2], s.m2[:, :, 0], rtol=1e-11)
    mag = measure.read_table(paths["magnetization"])
    assert np.allclose(mag[:, 1], s.m[:, 3, 0], rtol=1e-11)
    assert np.allclose(mag[:, 3], s.m2[:, 3, 0], rtol=1e-11)


def test_short_series_has_no_error_bars():
    rows = synthetic_rows(n=3, R=2)
    s = measure.summarise(rows, [1.0, 0.5], 64)
    assert np.isnan(s.e[:, 1]).all() and np.isnan(s.c).all()
    assert np.isfinite(s.e[:, 0]).all()
