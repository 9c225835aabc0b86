import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


@pytest.fixture(scope="session")
def gpu():
    import torch

    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a visible GPU")
    from psmma_git_amd import api

    api.lib()  # raises if the HIP library is missing: no fallback
    return torch.device("cuda:0")


@pytest.fixture(scope="session")
def ora():
    from oracle import oracle

    oracle.lib()
    return oracle
