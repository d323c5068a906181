"""The C ABI library loads, exports every entry point of include/pyro.h, and
rejects bad arguments before touching a device. No GPU compute here."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from cudacode_amd import pyro

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    names = []
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if h.endswith(".h"):
            text = open(os.path.join(ROOT, "include", h)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names += re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(pyro_\w+)\s*\(", text, flags=re.M)
    return names


def test_every_declared_symbol_is_exported():
    lib = pyro.lib()
    names = header_functions()
    assert len(names) >= 30
    for n in names:
        assert hasattr(lib, n), n
    assert set(pyro.EXPORTS) <= set(names)


def test_version_string():
    assert b"gfx950" in pyro.lib().pyro_version()


def test_create_rejects_bad_config_before_device_use():
    with pytest.raises(pyro.PyroError, match="L=1"):
        pyro.Pyrochlore(1, [1.0])
    with pytest.raises(pyro.PyroError, match="not positive"):
        pyro.Pyrochlore(4, [1.0, -2.0])
    with pytest.raises(pyro.PyroError, match="cos_cone"):
        pyro.Pyrochlore(4, [1.0], cos_cone=1.5)
    with pytest.raises(pyro.PyroError, match="negative"):
        pyro.Pyrochlore(4, [1.0], n_or=-1)


def test_null_ctx_is_an_error_not_a_crash():
    lib = pyro.lib()
    assert lib.pyro_sweep(None, 1, 0) == -1
    assert b"null ctx" in lib.pyro_last_error()
    assert lib.pyro_destroy(None) == 0


AIAO = np.array([[-1, -1, -1], [1, 1, -1], [-1, 1, 1], [1, -1, 1]], np.float64) / np.sqrt(3)


def test_order_parameter_closed_forms():
    """The twelve rows of Measurement (Pyrochlore.cu:580-624) evaluated on ideal
    q=0 states: all-in-all-out gives |A2| = 1; a uniform state gives the ferro
    vector; states of the E doublet give |m_E| = 1; each state leaves the other
    groups at zero."""
    N = 1024.0
    q = N / 4

    def fold(vecs):
        return pyro.fold_order_params(np.concatenate([q * np.asarray(v) for v in vecs]), N)[0]

    m = fold(AIAO)
    assert abs(abs(m[0]) - 1) < 1e-12
    assert np.allclose(m[1:], 0, atol=1e-12)
    m = fold([[0.6, 0, 0.8]] * 4)
    assert np.allclose(m[6:9], [0.6, 0, 0.8]) and np.allclose(np.delete(m, [6, 7, 8]), 0)
    # E doublet (psi_2 / psi_3): local x / y axes of each sublattice.
    psi2 = np.array([[-2, 1, 1], [2, -1, 1], [-2, -1, -1], [2, 1, -1]]) / np.sqrt(6)
    m = fold(psi2)
    assert abs(np.hypot(m[1], m[2]) - 1) < 1e-12
    assert np.allclose(np.delete(m, [1, 2]), 0, atol=1e-12)
    psi3 = np.array([[0, -1, 1], [0, 1, 1], [0, 1, -1], [0, -1, -1]]) / np.sqrt(2)
    m = fold(psi3)
    assert abs(np.hypot(m[1], m[2]) - 1) < 1e-12
    assert np.allclose(np.delete(m, [1, 2]), 0, atol=1e-12)


def test_order_parameter_rows_orthogonal():
    """The 12 rows span the 12-dimensional space of sublattice vectors with
    equal norms (|row|^2 = 4) and are mutually orthogonal: the decomposition
    into the A2, E, T1, ferro and T2 groups is complete and loses nothing."""
    N = 4.0
    rows = np.array([pyro.fold_order_params(np.eye(12)[i], N)[0] * N for i in range(12)]).T
    assert np.allclose(rows @ rows.T, 4 * np.eye(12), atol=1e-12)
