import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device)")
    config.addinivalue_line("markers", "slow: long-running CPU test")


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O

    O.build()
    return O


@pytest.fixture(scope="session")
def dev():
    # torch (which some GPU tests use for device memory) ships its own HIP runtime; it must see the
    # device before the library's runtime does, so it is initialised first
    import torch

    if torch.cuda.is_available():
        torch.cuda.init()
    import dmtz_amd

    return dmtz_amd.default_context()
