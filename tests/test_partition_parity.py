"""Partition descriptors of libsblas (the product) against the oracle,
bit-exact, for every rule and chunk count (host code; no GPU needed)."""
import numpy as np
import pytest

from conftest import random_csr
from test_oracle_pins import CHUNK_COUNTS, FIELDS, _matrices


def _eq(a, b):
    assert len(a) == len(b)
    for x, y in zip(a, b):
        for f in FIELDS:
            assert int(x[f]) == int(y[f]), (f, x, y)


@pytest.fixture(scope="module")
def S():
    from hpda_sblas_pnnl_amd import spmv
    return spmv


MAPPINGS = [None, "split", "alt"]


def _mapping(kind, C):
    if kind is None:
        return None
    if kind == "split":
        return [0] * (C // 2) + [1] * (C - C // 2)
    return [i % 2 for i in range(C)]


def test_v1_v2_rowblock(S, oracle):
    for m, rp in _matrices():
        nnz = int(rp[-1])
        for C in CHUNK_COUNTS:
            st, ch = S.partition("v1", C=C, m=m, nnz=nnz, row_ptr=rp)
            ost, och = oracle.part_v1(m, rp, nnz, C)
            assert st == 0 and ost == 0
            _eq(ch, och)
            st, ch = S.partition("rowblock", C=C, m=m, row_ptr=rp)
            ost, och = oracle.part_rowblock(m, rp, C)
            assert st == ost == 0
            _eq(ch, och)
        for nb, q in [(1, 1), (5, 3), (64, 4), (1 << 20, 2)]:
            st, ch = S.partition("v2", C=0, m=m, nnz=nnz, row_ptr=rp, nb=nb, q=q)
            ost, och = oracle.part_v2(m, rp, nnz, nb, q)
            assert st == 0 and ost == 0
            _eq(ch, och)


def test_v1_numa_family(S, oracle):
    for m, rp in _matrices():
        nnz = int(rp[-1])
        ri = np.repeat(np.arange(m, dtype=np.int32), np.diff(rp))
        for C in CHUNK_COUNTS:
            for kind in MAPPINGS:
                mp = _mapping(kind, C)
                st, ch, nodes = S.partition("v1_numa", C=C, m=m, nnz=nnz, row_ptr=rp,
                                            numa_mapping=mp)
                ost, och, onodes = oracle.part_v1_numa(m, rp, nnz, C, mp)
                assert (st == 0) == (ost == 0)
                if st == 0:
                    _eq(ch, och)
                st, ch, _ = S.partition("v1_numa_csc", C=C, n=m, nnz=nnz, col_ptr=rp,
                                        numa_mapping=mp)
                ost, och, _ = oracle.part_v1_numa_csc(m, rp, nnz, C, mp)
                assert (st == 0) == (ost == 0)
                if st == 0:
                    _eq(ch, och)
                st, ch, _ = S.partition("v1_numa_coo", C=C, m=m, row_idx=ri, numa_mapping=mp)
                ost, och, _ = oracle.part_v1_numa_coo(m, ri, C, mp)
                assert (st == 0) == (ost == 0), (st, ost)
                if st == 0:
                    _eq(ch, och)
            st, ch = S.partition("baseline_coo", C=C, m=m, row_idx=ri)
            ost, och = oracle.part_baseline_coo(m, ri, C)
            assert st == ost == 0
            _eq(ch, och)
            st, ch = S.partition("baseline_csc", C=C, n=m, col_ptr=rp)
            ost, och = oracle.part_baseline_csc(m, rp, C)
            assert st == ost == 0
            _eq(ch, och)


def test_large_and_undefined(S, oracle):
    rng = np.random.default_rng(11)
    # banded-like row pointer with 2e6 rows and ~2.8e8 elements
    lens = np.clip(np.rint(rng.normal(140, 40, 2_000_000)), 8, 400).astype(np.int64)
    rp = np.zeros(len(lens) + 1, np.int64)
    np.cumsum(lens, out=rp[1:])
    rp = rp.astype(np.int32)
    nnz = int(rp[-1])
    m = len(lens)
    # v1 at C=8 overflows the reference's int products: both report it
    st, _ = S.partition("v1", C=8, m=m, nnz=nnz, row_ptr=rp)
    ost, _ = oracle.part_v1(m, rp, nnz, 8)
    assert ost == oracle.ORA_UNDEFINED and st == 4  # SBLAS_ERR_UNDEFINED
    for C in (1, 2, 7, 8, 16):
        st, ch, nodes = S.partition("v1_numa", C=C, m=m, nnz=nnz, row_ptr=rp)
        ost, och, _ = oracle.part_v1_numa(m, rp, nnz, C)
        assert st == ost == 0
        _eq(ch, och)


def test_searches_exported(S, oracle):
    rng = np.random.default_rng(12)
    for _ in range(200):
        n = int(rng.integers(1, 30))
        a = np.sort(rng.integers(0, 20, size=n + 1)).astype(np.int32)
        for key in range(-1, 22):
            assert S.get_row_from_index(n, a, key) == oracle.get_row_from_index(n, a, key)
            # within the searched range both agree; the reference's read of a[n]
            # is past the caller's array in its COO call sites
            assert S.findFirstInSorted(a, n, key) == oracle.find_first_in_sorted(a, n, key, alen=n)
            assert S.findLastInSorted(a, n, key) == oracle.find_last_in_sorted(a, n, key, alen=n)
