"""The C ABI from a C++ caller, the way the reference's host code would bind
it: a program that includes include/pyro.h and links cudacode_amd/libpyro.so
builds with g++, and its calls that need no GPU return the known answers:
the version string, the link table and the independent sets (equal to the
library's own Python view of them), the Philox4x32-10 known-answer vectors of
the Random123 distribution, a configuration refused before any device use
with its error text, and a null ctx refused without a crash."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from cudacode_amd import pyro

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROGRAM = r"""
#include <cstdio>
#include "pyro.h"

int main() {
  std::printf("version %s\n", pyro_version());
  int8_t links[16][6][4];
  int8_t sets[24][4];
  if (pyro_links16(links) || pyro_independent_sets(sets)) return 1;
  std::printf("links");
  for (int s = 0; s < 16; ++s)
    for (int n = 0; n < 6; ++n)
      for (int c = 0; c < 4; ++c) std::printf(" %d", links[s][n][c]);
  std::printf("\nsets");
  for (int k = 0; k < 24; ++k)
    for (int c = 0; c < 4; ++c) std::printf(" %d", sets[k][c]);
  std::printf("\n");
  const uint32_t ctrs[2][4] = {{0, 0, 0, 0}, {0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu}};
  const uint32_t keys[2][2] = {{0, 0}, {0xffffffffu, 0xffffffffu}};
  for (int v = 0; v < 2; ++v) {
    uint32_t out[4];
    if (pyro_philox4x32_10(ctrs[v], keys[v], out)) return 2;
    std::printf("philox %08x %08x %08x %08x\n", out[0], out[1], out[2], out[3]);
  }
  const double temps[1] = {1.0};
  pyro_config cfg = {};
  cfg.L = 1;  // refused: a cell has neighbours in the next cell on every side
  cfg.n_replicas = 1;
  cfg.temperatures = temps;
  cfg.J = 1.0;
  cfg.spin_len2 = 1.0;
  cfg.cos_cone = -1.0f;
  pyro_ctx *ctx = nullptr;
  const int rc = pyro_create(&cfg, &ctx);
  std::printf("create %d %s\n", rc, pyro_last_error());
  std::printf("sweep-null %d %s\n", pyro_sweep(nullptr, 1, 0), pyro_last_error());
  return ctx == nullptr ? 0 : 3;
}
"""


@pytest.fixture(scope="module")
def caller_output(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ is not installed")
    lib_dir = os.path.join(ROOT, "cudacode_amd")
    assert os.path.exists(os.path.join(lib_dir, "libpyro.so")), "build the library first"
    d = tmp_path_factory.mktemp("cpp_caller")
    src, exe = d / "caller.cpp", d / "caller"
    src.write_text(PROGRAM)
    subprocess.run([gxx, "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), str(src),
                    "-L", lib_dir, "-lpyro", f"-Wl,-rpath,{lib_dir}", "-o", str(exe)],
                   check=True, capture_output=True, text=True)
    run = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert run.returncode == 0, (run.returncode, run.stdout, run.stderr)
    return {line.split(" ", 1)[0]: line.split(" ", 1)[1] for line in run.stdout.splitlines()
            if " " in line} | {"philox": [l.split()[1:] for l in run.stdout.splitlines()
                                          if l.startswith("philox")]}


def test_version_and_tables(caller_output):
    assert "gfx950" in caller_output["version"]
    links = np.array(caller_output["links"].split(), np.int64).reshape(16, 6, 4)
    sets = np.array(caller_output["sets"].split(), np.int64).reshape(24, 4)
    assert np.array_equal(links, pyro.links16())
    assert np.array_equal(sets, pyro.independent_sets())


def test_philox_known_answers(caller_output):
    assert caller_output["philox"] == [["6627e8d5", "e169c58d", "bc57ac4c", "9b00dbd8"],
                                       ["408f276d", "41c83b0e", "a20bc7c6", "6d5451fd"]]


def test_errors_are_codes_and_text(caller_output):
    rc, text = caller_output["create"].split(" ", 1)
    assert int(rc) == -1 and "L=1" in text
    rc, text = caller_output["sweep-null"].split(" ", 1)
    assert int(rc) == -1 and "null ctx" in text
