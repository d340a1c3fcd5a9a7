import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs through the HIP library)")


_ops_cache = {}


def oracle_operators(n=4):
    """FP64 operators from the oracle's precompute (cached per session)."""
    import oracle as O
    if n not in _ops_cache:
        _ops_cache[n] = O.operators(n)
    return _ops_cache[n]


@pytest.fixture(scope="session")
def ops4():
    return oracle_operators(4)


@pytest.fixture(scope="session")
def cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU visible")
    return torch.device("cuda:0")
