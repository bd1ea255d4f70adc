import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line('markers', 'gpu: needs a ROCm GPU (MI355X) and the built libmacromax_hip.so')


def gpu_available() -> bool:
    try:
        import torch
        return torch.cuda.is_available() and torch.cuda.device_count() > 0
    except Exception:
        return False


@pytest.fixture(scope='session')
def device():
    import torch
    if not gpu_available():
        pytest.fail('GPU test collected without a GPU: run it on the MI355X box (pytest -m gpu).')
    return torch.device('cuda:0')
