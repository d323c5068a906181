import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

# Under pytest-xdist every worker would start an OpenMP team as wide as the
# machine for the oracle's tiny lattices; the spinning teams then starve each
# other (a 0.5 s test took minutes). One thread per worker keeps parallel runs
# as fast as serial ones. Set before the oracle library is loaded.
if os.environ.get("PYTEST_XDIST_WORKER"):
    os.environ["OMP_NUM_THREADS"] = "1"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")
    config.addinivalue_line("markers", "slow: long CPU statistics test")


def gpu_available() -> bool:
    try:
        import torch  # noqa: F401  (only to ask the runtime for a device)
        return torch.cuda.is_available()
    except Exception:
        return False
