"""Matrix field model (anisotropic exchange + Zeeman field) on the GPU,
through the C ABI.

- the matrix-model phases give the oracle's spins bit for bit (both kernel
  variants, with exchanges), and FP64 energies to 1e-12;
- with C = -J I and no field the energy equals the scalar model's;
- closed form: a uniform configuration u has
  E = -(N/4) sum_b [u.B_b + sum_{b' != b} u^T C_{b b'} u] (each site has two
  neighbours of each other basis site), evaluated from the FP32 values the
  kernels use;
- over-relaxation conserves the energy with the field on; the incremental
  energy follows the recomputed one;
- the model is refused where it is not supported.
"""
import sys

import numpy as np
import pytest

from oracle import oracle

pytestmark = pytest.mark.gpu

SEED = 0xC0FFEE11


def say(msg):
    print(f"[gpu-couplings] {msg}", flush=True)
    sys.stdout.flush()


def model(spin_len2=0.25):
    from cudacode_amd import pyro
    C = pyro.exchange_matrices(*pyro.YB2TI2O7_J)
    B = pyro.zeeman_fields([0.4, 0.1, 0.9], 4.18, 1.77)
    return C, B


def ctx(L, temps, variant=1, track=True, spin_len2=0.25, **kw):
    from cudacode_amd import pyro
    say(f"create L={L} R={len(temps)} variant={variant} {kw}")
    g = pyro.Pyrochlore(L, temps, seed=SEED, spin_len2=spin_len2, **kw)
    g.set_kernel_options(variant, track)
    return g


@pytest.mark.parametrize("L,variant", [(4, 0), (4, 1), (8, 1), (12, 0), (16, 1)])
def test_matrix_model_matches_oracle(L, variant):
    """The anisotropic exchange in a field (Yb2Ti2O7 couplings) follows the
    oracle bit for bit: spins, acceptance, rungs; FP64 energies to 1e-12, for
    both kernel variants and the compiled lattice sizes."""
    temps = [0.05, 0.2, 0.6, 1.5]
    kw = dict(n_metro=1, n_or=2, pt_every=3, cos_cone=0.3)
    C, B = model()
    with ctx(L, temps, variant=variant, **kw) as g:
        g.set_couplings(C, B)
        is_mat, coup = g.couplings()
        assert is_mat
        g.Mc(12)
        spins = g.get_spins()
        E = g.TotalEnergyv2()
        inc = g.energy_incremental()
        acc = g.acceptance()[0]
        at = g.temperatures()[1]
    o = oracle.OracleRun(L, temps, seed=SEED, spin_len2=0.25, **kw)
    oracle.set_couplings(coup)
    try:
        o.anchor = o.observables()[:, 0].copy()
        o.run(12)
        assert np.array_equal(spins, o.spins)
        assert np.array_equal(acc, o.accepted)
        assert np.array_equal(at, o.replica_at)
        assert np.allclose(E, o.observables()[:, 0], rtol=1e-12)
        assert np.allclose(inc, E, rtol=1e-11, atol=1e-9)
    finally:
        oracle.set_couplings(None)


def test_heisenberg_matrices_equal_scalar_energy():
    from cudacode_amd import pyro
    L, temps = 4, [0.3, 1.0]
    with ctx(L, temps, spin_len2=1.0, n_metro=1, n_or=1) as a, \
            ctx(L, temps, spin_len2=1.0, n_metro=1, n_or=1) as b:
        a.Mc(5)
        s = a.get_spins()
        b.set_spins(s)
        b.set_couplings(pyro.exchange_matrices(1.0, 1.0, 0.0, 0.0), None)
        assert np.allclose(b.TotalEnergyv2(), a.TotalEnergyv2(), rtol=1e-14)
        b.set_couplings(None)  # back to the scalar model
        assert not b.couplings()[0]
        assert np.array_equal(b.TotalEnergyv2(), a.TotalEnergyv2())


def test_uniform_state_closed_form():
    L, temps = 4, [0.5]
    N = 16 * L ** 3
    C, B = model()
    u = np.array([0.36, -0.48, 0.8], np.float32)
    with ctx(L, temps, n_metro=0, n_or=0) as g:
        g.set_couplings(C, B)
        _, coup = g.couplings()
        soa = np.concatenate([np.full(16 * L ** 3, u[c], np.float32) for c in range(3)])
        g.set_spins(soa)
        E = g.TotalEnergyv2()[0]
    Ch = coup[:144].astype(np.float64).reshape(4, 4, 3, 3)
    Bh = coup[144:].astype(np.float64).reshape(4, 3)
    ud = u.astype(np.float64)
    expect = -(N / 4) * sum(ud @ Bh[b] + sum(ud @ Ch[b, bp] @ ud for bp in range(4) if bp != b)
                            for b in range(4))
    assert E == pytest.approx(expect, rel=1e-12)


def test_overrelaxation_conserves_energy_in_field():
    L, temps = 8, [0.2, 0.9]
    C, B = model()
    with ctx(L, temps, n_metro=1, n_or=0, cos_cone=0.5) as g:
        g.set_couplings(C, B)
        g.Mc(10)
        e0 = g.TotalEnergyv2()
        g.set_passes(0, 5)
        g.Mc(10)
        e1 = g.TotalEnergyv2()
        assert not np.array_equal(e0, e1)  # spins moved
        assert np.allclose(e0, e1, rtol=1e-5)
        assert np.allclose(g.energy_incremental(), e1, rtol=1e-11, atol=1e-9)


def test_matrix_model_refusals():
    from cudacode_amd import pyro
    C, B = model()
    with pyro.Batch([4, 8], [0.5, 0.5], seed=SEED) as g:
        with pytest.raises(pyro.PyroError, match="uniform lattice"):
            g.set_couplings(C, B)
    with ctx(4, [0.5, 1.0], n_metro=1, n_or=1) as g:
        g.set_couplings(C, B)
        with pytest.raises(pyro.PyroError, match="execution mode 0"):
            g.set_exec_mode(1)
        g.mlp_set_weights(pyro.mlp_init(1))
        with pytest.raises(pyro.PyroError, match="scalar model"):
            g.set_learned_passes(1)
        bad = C.copy()
        bad[0, 1, 0, 2] += 0.1
        with pytest.raises(pyro.PyroError, match="transpose"):
            g.set_couplings(bad, B)


def test_g_magnetization_from_device_sums_equals_oracle():
    """The g-tensor-weighted magnetisation (MagnetizationGPU,
    Pyrochlore.cu:277-342) from the device's FP64 sublattice sums equals the
    oracle's fold of the oracle's sums of the same spins (reference g1..g4
    sign patterns), in the Yb2Ti2O7 model under a field; at g_xy = g_z = 1 it
    is the plain magnetisation."""
    from cudacode_amd import pyro
    gxy, gz = 4.18, 1.77
    temps = [0.3, 1.0, 3.0]
    with ctx(8, temps, n_metro=1, n_or=2, cos_cone=0.3) as g:
        g.set_couplings(pyro.exchange_matrices(*pyro.YB2TI2O7_J),
                        pyro.zeeman_fields([0.2, 0.5, 1.0], gxy, gz))
        g.Mc(30)
        M = g.MagnetizationG(gxy, gz)
        M1 = g.MagnetizationG(1.0, 1.0)
        plain = g.Magnetization()
        s = g.get_spins()
    sub = oracle.observables(s, 8, len(temps), 1.0)[:, 1:]
    ref = np.array([oracle.magnetization_g(x, gxy, gz) for x in sub])
    say(f"M_g {M.tolist()}")
    assert np.allclose(M, ref, rtol=1e-12, atol=1e-9)
    assert np.array_equal(M1, plain)
    assert np.abs(M - plain).max() > 1.0  # the g-tensor matters here
