"""Run-configuration presets (host arithmetic only)."""
import numpy as np

from cudacode_amd import presets, pyro


def test_presets_match_the_benchmark_workloads():
    c2 = presets.preset("c2")
    T = c2.temperatures()
    assert T.size == 64 and np.isclose(T.min(), 0.01) and np.isclose(T.max(), 1.0)
    assert np.allclose(np.diff(np.log(np.sort(T))), np.log(100) / 63)  # geometric
    c3 = presets.preset("c3")
    L = c3.sizes()
    assert L.size == 256 and set(np.unique(L)) <= {8, 12, 16, 24}
    assert np.array_equal(L, np.random.default_rng(presets.DEFAULT_SEED).choice([8, 12, 16, 24], 256))
    assert np.array_equal(np.unique(c3.temperatures(), return_counts=True)[1], np.full(32, 8))
    c1 = presets.preset("c1", sweeps=100)
    assert c1.sweeps == 100 and list(c1.temperatures()) == [0.5] and c1.pt_every == 0
    assert presets.preset("c5").n_learned == 1 and presets.preset("c4").n_metro == 0
    assert presets.DEFAULT_SEED == 56081376
