"""MI355X-native pyrochlore Heisenberg Monte Carlo step.

The compute lives in the gfx950 shared library ``libpyro.so`` (HIP kernels and
the C ABI of ``include/pyro.h``); :mod:`cudacode_amd.pyro` is its host-side
interface and :mod:`cudacode_amd.stats` the estimators (binning, jackknife,
integrated autocorrelation time) used to report results.
"""
__all__ = ["pyro", "stats"]
