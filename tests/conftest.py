import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X); runs the HIP kernels")


def _has_gpu():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    if not _has_gpu():
        pytest.skip("no ROCm GPU visible")
    import torch
    import qumasun_amd
    qumasun_amd.lib()   # fails loudly when the HIP library is missing
    return torch.device("cuda:0")


@pytest.fixture(scope="session")
def oracle():
    import oracle as O
    O.olib()
    return O
