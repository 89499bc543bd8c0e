"""Host-side invariants of the row blocks the LDS-staged CSR kernel runs on:
every block's element window fits the block's LDS (256 * items elements) and
its rows fit the y staging (1024 rows), blocks tile the elements of [0, nnz)
exactly once in order, every row of [0, m) is in some block, and a row is
shared only where a chunk edge or a block capacity cuts it."""
import ctypes

import numpy as np
import pytest

from conftest import random_csr


def _blocks(m, rp, nnz, C, mode, items):
    from hpda_sblas_pnnl_amd import _lib
    L = _lib.lib()
    f = L.sblas_internal_row_blocks
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_int,
                  ctypes.c_int, ctypes.c_void_p, ctypes.c_int]
    rp = np.ascontiguousarray(rp, np.int32)
    n = f(m, rp.ctypes.data, nnz, C, mode, items, None, 0)
    out = np.zeros((max(n, 1), 4), np.int32)
    f(m, rp.ctypes.data, nnz, C, mode, items, out.ctypes.data, n)
    return out[:n]


def _check(m, rp, C, mode, items):
    nnz = int(rp[-1])
    B = _blocks(m, rp, nnz, C, mode, items)
    cap = 256 * items
    assert len(B) >= 1
    assert np.all(B[:, 3] - B[:, 2] <= cap)
    assert np.all(B[:, 3] >= B[:, 2])
    assert np.all(B[:, 1] - B[:, 0] <= 1024)
    assert np.all(B[:, 1] > B[:, 0])
    # element windows tile [0, nnz)
    assert B[0, 2] == 0 and B[-1, 3] == nnz
    assert np.all(B[1:, 2] == B[:-1, 3])
    # rows covered, in order; consecutive blocks share at most one row
    assert B[0, 0] == 0 and B[-1, 1] == m
    assert np.all((B[1:, 0] == B[:-1, 1]) | (B[1:, 0] == B[:-1, 1] - 1))
    # each block's window lies inside its rows' elements
    for rb, re, lo, hi in B:
        assert rp[rb] <= lo or rb == re
        assert hi <= rp[re] or hi == lo
    return B


@pytest.mark.parametrize("mode", [0, 1, 2, 3])
@pytest.mark.parametrize("items", [8, 16])
def test_row_block_invariants(mode, items):
    rng = np.random.default_rng(21)
    for trial in range(15):
        m = int(rng.integers(1, 3000))
        rp, _, _ = random_csr(rng, m, 5000, float(rng.uniform(0.2, 60)),
                              empty_frac=float(rng.choice([0, 0.3, 0.9])),
                              long_rows=int(rng.integers(0, 4)), long_len=5000)
        for C in (1, 2, 3, 8, 17):
            _check(m, rp, C, mode, items)


def test_long_rows_and_empty():
    # a row of 100k elements between empty rows; no elements at all
    rp = np.array([0, 0, 100_000, 100_000, 100_001, 100_001], np.int32)
    B = _check(5, rp, 1, 0, 8)
    assert (B[:, 1] - B[:, 0] == 1).sum() >= 100_000 // 2048
    _check(5, rp, 8, 1, 16)
    rp0 = np.zeros(7, np.int32)
    B = _check(6, rp0, 4, 0, 8)
    assert len(B) == 1 and tuple(B[0]) == (0, 6, 0, 0)


def test_csr5_sigma_rule():
    """Tile depth of the CSR5-style form with sigma = 0: the reference's rule
    on integer nnz / m (include/anonymouslib_cuda.h:294-318: 4 up to 4, nnz/m
    up to 32, 32 up to 256, 6 beyond) snapped to the depths the kernel is
    built for, nearest in log2 (midpoints 8*sqrt(2) = 11.3, 16*sqrt(2) = 22.6)."""
    from hpda_sblas_pnnl_amd import _lib
    f = _lib.lib().sblas_internal_autotune_sigma
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_int64, ctypes.c_int]
    m = 1000
    cases = {1: 8, 4: 8, 5: 8, 11: 8, 12: 16, 16: 16, 22: 16, 23: 32, 32: 32, 33: 32, 140: 32,
             256: 32, 257: 8, 2000: 8}
    for per_row, want in cases.items():
        assert f(per_row * m, m) == want, per_row
        assert f(per_row * m + m - 1, m) == want, per_row  # integer division, as the reference
    assert f(0, 0) == 8
