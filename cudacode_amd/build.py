"""Build the gfx950 shared library libpyro.so in-tree with hipcc.

Run as ``python -m cudacode_amd.build`` or through ``__graft_entry__.build()``.
The library is the product: kernels and the C ABI of include/pyro.h.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libpyro.so")

SOURCES = [
    "pyro_kernels.hip",
    "pyro_learned.hip",
    "pyro_batch.hip",
    "pyro_io.hip",
    "pyro_aniso.hip",
    "pyro_capi.hip",
]

# -ffp-contract=off: no fused multiply-add contraction, so float expressions
# round exactly as written (the host restatement evaluates the same sequence).
# -packed-fp32-ops: keep the compiler from forming packed FP32 instructions.
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-command-line-argument",
    "-Wno-unknown-warning-option",
]

# Per-source additions. The learned-proposal kernels keep MFMA accumulators in
# ordinary VGPRs (gfx950 MFMAs read and write either register file), which
# saves the accumulator-to-VGPR copies before every activation.
EXTRA_FLAGS = {
    "pyro_learned.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"],
}


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the gfx950 library cannot be built")


def build(verbose: bool = False) -> str:
    """Compile every source for gfx950 and link libpyro.so. Returns its path."""
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs, procs = [], []
    for src in SOURCES:  # translation units compile in parallel
        obj = os.path.join(objdir, os.path.splitext(src)[0] + ".o")
        cmd = [hipcc(), *HIP_FLAGS, *EXTRA_FLAGS.get(src, []), "-I", os.path.join(ROOT, "include"), "-c",
               os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd)))
        objs.append(obj)
    failed = [src for src, p in procs if p.wait() != 0]
    if failed:
        raise RuntimeError(f"hipcc failed on {', '.join(failed)}")
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", tmp]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    build_examples()
    return LIB


EXAMPLES = ["pt_driver"]


def build_examples() -> None:
    """Host programs in examples/ that drive the library through the C ABI
    (include/pyro.h), linked next to libpyro.so; a C++ compiler is enough."""
    cxx = shutil.which("g++") or shutil.which("c++")
    if cxx is None:
        raise RuntimeError("no C++ compiler for the examples")
    for name in EXAMPLES:
        out = os.path.join(HERE, name)
        cmd = [cxx, "-std=c++17", "-O2", "-Wall", "-I", os.path.join(ROOT, "include"),
               os.path.join(HERE, "examples", name + ".cpp"), "-L", HERE, "-lpyro",
               "-Wl,-rpath,$ORIGIN", "-o", out + ".tmp"]
        subprocess.run(cmd, check=True)
        os.replace(out + ".tmp", out)


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
