import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through libgdsjoin.so)")


def _has_gpu():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.fixture(scope="session")
def gpu():
    if not _has_gpu():
        pytest.skip("no GPU")
    import torch
    torch.cuda.init()
    return torch.device("cuda:0")


@pytest.fixture(scope="session")
def sj(gpu):
    from gpu_self_join_amd import SelfJoin
    h = SelfJoin(0)
    yield h
    h.close()
