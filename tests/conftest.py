import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP device (MI355X)")


@pytest.fixture(scope="session", autouse=True)
def _torch_device_first(request):
    """torch carries its own HIP runtime beside the system one libsblas links.
    When both share a process, torch must open the device first: if libsblas
    initialises the GPU before torch does, torch then finds no GPU."""
    if any(item.get_closest_marker("gpu") for item in request.session.items):
        import torch
        if torch.cuda.is_available():
            torch.zeros(1, device="cuda")
    yield


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as ora
    if not os.path.exists(ora.LIB_PATH):
        ora.build()
    return ora


def random_csr(rng, m, n, mean_len, empty_frac=0.0, long_rows=0, long_len=0):
    """Random CSR with sorted unique columns; some empty rows and long rows."""
    lens = rng.poisson(mean_len, size=m).astype(np.int64)
    lens = np.minimum(lens, n)
    if empty_frac > 0:
        lens[rng.random(m) < empty_frac] = 0
    for _ in range(long_rows):
        lens[rng.integers(0, m)] = min(long_len, n)
    rp = np.zeros(m + 1, dtype=np.int64)
    np.cumsum(lens, out=rp[1:])
    col = np.empty(int(rp[-1]), dtype=np.int32)
    for r in range(m):
        k = int(lens[r])
        if k:
            col[rp[r]:rp[r + 1]] = np.sort(rng.choice(n, size=k, replace=False))
    val = rng.uniform(-1, 1, size=len(col))
    return rp.astype(np.int32), col, val
