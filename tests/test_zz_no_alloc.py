"""The entry point that allocates nothing (pyro_observables_noalloc), through
the C ABI: on caller-owned device spins and a caller-owned device output, on
the caller's own stream, it returns what pyro_energy and pyro_sublattice_sums
return for a ctx holding the same spins, bit for bit; two calls in a row
share the library's scratch (the second on a smaller input, so stale
partials of the first lie beyond what it reads, then the larger again), and
a request above the scratch maximum is refused before any device work.

The caller's buffer and stream come from the HIP runtime the library itself
links (found in this process's map once the library is loaded): device
pointers and stream handles of another runtime copy (such as the one bundled
with a Python GPU framework) are not interchangeable with it."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[gpu-noalloc] {msg}", flush=True)


def hip_runtime():
    """The libamdhip64 that libpyro.so loaded (same handle, same runtime)."""
    from cudacode_amd import pyro
    pyro.lib()
    with open("/proc/self/maps") as f:
        paths = {ln.split()[-1] for ln in f if "libamdhip64" in ln and "/" in ln}
    rocm = sorted(p for p in paths if "/torch/" not in p)
    assert rocm, paths
    hip = C.CDLL(rocm[0])
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
    hip.hipStreamDestroy.argtypes = [C.c_void_p]
    hip.hipStreamSynchronize.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    return hip


def ok(rc):
    assert rc == 0, f"HIP error {rc}"


def reference(g):
    E = g.TotalEnergyv2()
    sub = g.sublattice_sums()
    return np.concatenate([E[:, None], sub], axis=1)


def run_noalloc(hip, stream, out, g, L, R, jeff):
    from cudacode_amd import pyro
    ptr, n = g.device_ptr()
    assert n == R * 48 * L ** 3
    ok(hip.hipMemset(out, 0xFF, R * 13 * 8))  # NaN patterns: every value must be written
    g.sync()  # the ctx's own stream is done with the spins
    say(f"pyro_observables_noalloc L={L} R={R} jeff={jeff}")
    pyro.observables_noalloc(0, ptr, L, R, jeff, out.value, stream.value)
    ok(hip.hipStreamSynchronize(stream))
    host = np.empty((R, 13), np.float64)
    ok(hip.hipMemcpy(host.ctypes.data_as(C.c_void_p), out, host.nbytes, 2))  # device to host
    return host


def test_noalloc_equals_the_allocating_entry_points():
    from cudacode_amd import pyro
    big = pyro.Pyrochlore(16, pyro.geometric_ladder(0.05, 2.0, 12), seed=3, n_metro=1, n_or=2,
                          pt_every=3, cos_cone=0.4, J=1.0, spin_len2=0.25)
    small = pyro.Pyrochlore(4, [0.3, 1.1], seed=4, n_metro=1, n_or=1, cos_cone=0.4, J=1.5)
    hip = hip_runtime()
    out, stream = C.c_void_p(), C.c_void_p()
    ok(hip.hipMalloc(C.byref(out), 12 * 13 * 8))
    ok(hip.hipStreamCreate(C.byref(stream)))
    try:
        big.Mc(9)
        small.Mc(5)
        got = run_noalloc(hip, stream, out, big, 16, 12, 0.25)
        ref = reference(big)
        assert np.array_equal(got, ref), np.abs(got - ref).max()
        got2 = run_noalloc(hip, stream, out, small, 4, 2, 1.5)
        ref2 = reference(small)
        assert np.array_equal(got2, ref2), np.abs(got2 - ref2).max()
        big.Mc(3)  # and the larger input again, after the smaller one
        got3 = run_noalloc(hip, stream, out, big, 16, 12, 0.25)
        assert np.array_equal(got3, reference(big))
        say("bitwise equal to pyro_energy / pyro_sublattice_sums on both inputs")
    finally:
        hip.hipStreamSynchronize(stream)
        hip.hipStreamDestroy(stream)
        hip.hipFree(out)
        big.close()
        small.close()


def test_noalloc_refuses_above_its_scratch():
    from cudacode_amd import pyro
    with pytest.raises(pyro.PyroError, match="scratch"):
        pyro.observables_noalloc(0, 1 << 20, 64, 33, 1.0, 1 << 20, 0)
