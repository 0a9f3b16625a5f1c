import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU and the built libpff_hip.so")


def gpu_available() -> bool:
    try:
        import torch

        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def ora():
    from oracle import pff_oracle

    pff_oracle.lib()
    return pff_oracle


@pytest.fixture(scope="session")
def pf():
    if not gpu_available():
        pytest.skip("no GPU")
    import pyfastflow_amd

    pyfastflow_amd._lib.load()
    return pyfastflow_amd
