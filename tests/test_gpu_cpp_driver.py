"""A C++ host program (cudacode_amd/examples/pt_driver.cpp, built next to the
library) runs the parallel-tempering workload through the C ABI alone, and
its result equals the Python binding's run of the same configuration bit for
bit: the spins (an FNV-1a hash of their bytes), the sweep and exchange
counters, and the FP64 energy of every rung."""
import os
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "cudacode_amd", "pt_driver")


def fnv1a(data: bytes) -> int:
    a = np.frombuffer(data, np.uint8)
    h = 1469598103934665603
    for b in a.tolist():
        h = ((h ^ b) * 1099511628211) & 0xFFFFFFFFFFFFFFFF
    return h


def parse(out):
    ladder, rungs, r = {}, [], {}
    for line in out.splitlines():
        f = line.split()
        if f[0] == "ladder":
            ladder[int(f[1])] = float(f[2])
        elif f[0] == "rung":
            rungs.append(float(f[5]))
        elif f[0] == "counters":
            r["counters"] = (int(f[1]), int(f[2]))
        elif f[0] == "spins_fnv1a":
            r["hash"] = int(f[1], 16)
    r["ladder"] = [ladder[k] for k in sorted(ladder)]
    r["e_per_rung"] = np.array(rungs)
    return r


@pytest.mark.parametrize("L,R,n_or,pt_every,sweeps", [(4, 3, 2, 3, 20), (8, 4, 3, 5, 30)])
def test_cpp_driver_equals_python_binding(L, R, n_or, pt_every, sweeps):
    from cudacode_amd import pyro
    assert os.path.exists(DRIVER), "built by cudacode_amd.build"
    seed = 0xC0FFEE
    args = [DRIVER, str(L), str(R), "0.1", "1.0", str(seed), str(n_or), str(pt_every), str(sweeps)]
    print(f"[gpu-cpp] {' '.join(args[1:])}", flush=True)
    run = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert run.returncode == 0, (run.stdout, run.stderr)
    c = parse(run.stdout)
    assert len(c["ladder"]) == R and c["counters"][0] >= sweeps
    g = pyro.Pyrochlore(L, c["ladder"], seed=seed, J=1.0, spin_len2=1.0, n_metro=1, n_or=n_or,
                        pt_every=pt_every, cos_cone=0.5)
    try:
        g.CalculateSolidAngle(0.5, 5, 10)
        g.Mc(sweeps)
        spins = g.get_spins()
        E = g.TotalEnergyv2()
        T, at = g.temperatures()
        counters = g.sweep_counters()
    finally:
        g.close()
    assert fnv1a(spins.tobytes()) == c["hash"]
    assert counters == c["counters"]
    N = 16 * L ** 3
    assert np.allclose(E[at] / N, c["e_per_rung"], rtol=0, atol=1e-12)
