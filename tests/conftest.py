import os
import sys

import pytest

# PyTorch ships its own HIP runtime.  Imported first, it is the runtime the
# library binds to (same soname), so torch tensors, streams and events and
# the library share one runtime; a process that loads the library first and
# torch later ends up with two runtimes, and torch then finds no GPU.
try:
    import torch  # noqa: F401
except ImportError:
    pass

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def dto():
    import dual_threshold_optimization_amd as d
    d._lib.lib()
    return d


@pytest.fixture(scope="session")
def O():
    from oracle import oracle
    oracle.lib()
    return oracle


@pytest.fixture(scope="session")
def gpu(dto):
    """Fails (does not skip) on a GPU test run without a device: the HIP path must be the one that runs."""
    dto._lib.require_device()
    return True
