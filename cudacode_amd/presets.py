"""Typed run configurations, one preset per benchmark workload.

The reference configures a run through its popt table (Pyrochlore.cu:1161-1216:
system size, steps, temperature range and scale, over-relaxation count,
exchange cadence, J1..J4, field, seed); here a `RunConfig` carries the same
fields and `build()` turns it into a ctx. The five presets are the workloads
the benchmark measures: a single small lattice, the L=16 x 64 parallel
tempering ensemble, 256 replicas of mixed sizes, the learned proposal on the
L=16 ensemble, and the L=64 x 32 large-lattice ensemble.
"""
from __future__ import annotations

from dataclasses import dataclass, field, replace
from typing import Optional, Sequence

import numpy as np

from . import pyro

DEFAULT_SEED = 56081376  # the reference's default --Rng (Pyrochlore.cu:1202)


@dataclass(frozen=True)
class RunConfig:
    name: str
    L: int | Sequence[int]            # one size, or one per replica (mixed batch)
    T_min: float
    T_max: float
    n_T: int                          # distinct temperatures (geometric ladder)
    replicas_per_T: int = 1
    n_metro: int = 1
    n_or: int = 10
    n_learned: int = 0                # learned passes per sweep (needs trained weights)
    pt_every: int = 10                # sweeps between exchanges (0: none)
    sweeps: int = 10_000
    cos_cone: float = 0.5
    target_acceptance: float = 0.5    # cone adaptation target
    J: float = 1.0
    spin_len2: float = 1.0
    seed: int = DEFAULT_SEED
    Lset: Optional[Sequence[int]] = field(default=None)  # sizes a batch draws from

    def temperatures(self) -> np.ndarray:
        ladder = pyro.geometric_ladder(self.T_min, self.T_max, self.n_T) if self.n_T > 1 \
            else np.array([self.T_min])
        return np.repeat(ladder, self.replicas_per_T)

    def sizes(self) -> np.ndarray:
        """Per-replica lattice sizes (drawn from the seed for a mixed batch)."""
        R = self.n_T * self.replicas_per_T
        if self.Lset is not None:
            return np.random.default_rng(self.seed).choice(list(self.Lset), size=R)
        return np.full(R, int(self.L))

    def build(self):
        """A ctx for this configuration (Pyrochlore, or Batch for mixed sizes)."""
        T = self.temperatures()
        kw = dict(seed=self.seed, n_metro=self.n_metro, n_or=self.n_or, pt_every=self.pt_every,
                  cos_cone=self.cos_cone, J=self.J, spin_len2=self.spin_len2)
        if self.Lset is not None:
            return pyro.Batch(self.sizes(), T, **kw)
        return pyro.Pyrochlore(int(self.L), list(T), **kw)


PRESETS = {
    "c1": RunConfig("c1", L=4, T_min=0.5, T_max=0.5, n_T=1, n_or=5, pt_every=0, sweeps=10_000),
    "c2": RunConfig("c2", L=16, T_min=0.01, T_max=1.0, n_T=64, sweeps=100_000),
    "c3": RunConfig("c3", L=0, T_min=0.05, T_max=1.0, n_T=32, replicas_per_T=8,
                    Lset=(8, 12, 16, 24)),
    "c4": RunConfig("c4", L=16, T_min=0.01, T_max=1.0, n_T=64, n_metro=0, n_learned=1,
                    sweeps=100_000),
    "c5": RunConfig("c5", L=64, T_min=0.02, T_max=0.5, n_T=32, n_learned=1, sweeps=10_000),
}


def preset(name: str, **overrides) -> RunConfig:
    """A preset, optionally with fields replaced (e.g. sweeps=100)."""
    return replace(PRESETS[name], **overrides)
