import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")


@pytest.fixture(scope="session")
def hip():
    """The engine library bound to cuda:0 (GPU tests only)."""
    import torch
    from rt_amd import lib
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a HIP device")
    lib.init(0)
    return lib
