"""Estimators for Monte Carlo time series: integrated autocorrelation time,
binning and jackknife errors.

The reference keeps running sums of E and E^2 per temperature in a Measure
class (missing Measure.hpp; used at Pyrochlore.cu:1426-1438,1643-1650) and
prints no error bars. Here the series themselves are kept, so errors and the
autocorrelation time that the throughput metric is corrected by can be
estimated.
"""
from __future__ import annotations

import numpy as np


def autocorrelation(x) -> np.ndarray:
    """Normalised autocorrelation rho(t), t = 0..n-1, of a 1-D series (FFT, zero padded)."""
    x = np.asarray(x, np.float64)
    n = x.size
    d = x - x.mean()
    f = np.fft.rfft(d, 2 * n)
    acf = np.fft.irfft(f * np.conj(f))[:n]
    if acf[0] <= 0:
        return np.zeros(n)
    return acf / acf[0]


def tau_int(x, c: float = 6.0):
    """Integrated autocorrelation time tau = 1/2 + sum_{t=1}^{W} rho(t) with
    Sokal's automatic window: the smallest W with W >= c * tau(W)
    (A. D. Sokal, "Monte Carlo methods in statistical mechanics", 1996).
    Returns (tau, W, err) with Madras-Sokal error tau * sqrt(2(2W+1)/n)."""
    x = np.asarray(x, np.float64)
    n = x.size
    if n < 4:
        return float("nan"), 0, float("nan")
    rho = autocorrelation(x)
    taus = 0.5 + np.cumsum(rho[1:])
    W = np.arange(1, n)
    ok = W >= c * taus
    if not ok.any():
        w = n - 1
    else:
        w = int(W[np.argmax(ok)])
    tau = float(max(taus[w - 1], 0.5))
    err = tau * np.sqrt(2.0 * (2 * w + 1) / n)
    return tau, w, float(err)


def binning_error(x, n_bins: int = 32) -> float:
    """Standard error of the mean from n_bins contiguous bin means."""
    x = np.asarray(x, np.float64)
    m = x.size // n_bins
    if m < 1:
        return float("nan")
    b = x[: m * n_bins].reshape(n_bins, m).mean(axis=1)
    return float(b.std(ddof=1) / np.sqrt(n_bins))


def jackknife(x, fn, n_bins: int = 32):
    """Jackknife estimate and error of fn(bin means) over n_bins contiguous bins.
    fn takes a 1-D array of samples and returns a scalar."""
    x = np.asarray(x, np.float64)
    m = x.size // n_bins
    if m < 1:
        return float("nan"), float("nan")
    xs = x[: m * n_bins].reshape(n_bins, m)
    full = fn(xs.ravel())
    loo = np.array([fn(np.delete(xs, i, axis=0).ravel()) for i in range(n_bins)])
    est = n_bins * full - (n_bins - 1) * loo.mean()
    err = np.sqrt((n_bins - 1) / n_bins * ((loo - loo.mean()) ** 2).sum())
    return float(est), float(err)


def specific_heat(e_series, T: float, n_spins: int, n_bins: int = 32):
    """C/N = (<E^2> - <E>^2) / (N T^2) with jackknife error."""
    def c(e):
        return (np.mean(e * e) - np.mean(e) ** 2) / (n_spins * T * T)
    return jackknife(e_series, c, n_bins)


def independent_samples_per_s(sweeps_per_s: float, tau: float) -> float:
    """Sweeps per second divided by 2 tau_int: the rate of statistically
    independent measurements."""
    return sweeps_per_s / (2.0 * tau)
