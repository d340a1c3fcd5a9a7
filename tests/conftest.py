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


def truncated_operators(U, K, n, r):
    """Rank-r operators from a set of rank >= r: U's first r columns and the
    leading r x r block of every E_k = U^T K0_k U (column-major)."""
    import numpy as np
    n3 = n ** 3
    r0 = U.size // n3
    assert r <= r0, (r, r0)
    Ur = np.ascontiguousarray(U.reshape(r0, n3)[:r]).ravel()
    Kr = np.ascontiguousarray(K.reshape(316, r0, r0)[:, :r, :r]).ravel()
    return Ur, Kr


_ops_rank = {}


def operators_upto(n, r):
    """Oracle operators of order n at rank r (from one cached set of rank
    max(r, n^3/2) per order)."""
    import oracle as O
    rmax = max(r, n ** 3 // 2)
    if n not in _ops_rank or _ops_rank[n][0].size // n ** 3 < rmax:
        _ops_rank[n] = O.operators(n, rmax)
    U, K, _ = _ops_rank[n]
    return truncated_operators(U, K, n, r)
