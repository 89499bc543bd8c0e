"""Matrix ingestion on the host side, no GPU: the oracle's COO sort and
conversions pinned against brute force and scipy, the Matrix Market reader
(include/mmio.h:254-367, test/dspmv_test_dgx1.cu:133-199) against
scipy.io.mmread, and the reader's and helpers' error statuses."""

import numpy as np
import pytest
import scipy.io
import scipy.sparse as sp


def _coo(rng, m, n, nnz, dup_frac=0.2):
    row = rng.integers(0, m, nnz).astype(np.int32)
    col = rng.integers(0, n, nnz).astype(np.int32)
    ndup = int(nnz * dup_frac)
    if ndup:  # repeat some (row, col) pairs with other values
        src = rng.integers(0, nnz, ndup)
        dst = rng.integers(0, nnz, ndup)
        row[dst], col[dst] = row[src], col[src]
    val = rng.uniform(-1, 1, nnz)
    return row, col, val


@pytest.mark.parametrize("by_col", [False, True])
def test_oracle_coo_sort_brute_force(oracle, by_col):
    """Stable lexicographic order, checked against Python's sorted() with the
    input position as the last key."""
    rng = np.random.default_rng(1)
    row, col, val = _coo(rng, 40, 30, 600, dup_frac=0.5)
    r, c, v = oracle.coo_sort(row, col, val, by_col=by_col)
    key = (lambda k: (col[k], row[k], k)) if by_col else (lambda k: (row[k], col[k], k))
    order = sorted(range(len(row)), key=key)
    assert np.array_equal(r, row[order]) and np.array_equal(c, col[order])
    assert np.array_equal(v, val[order])


def test_oracle_coo_to_csr_csc_scipy(oracle):
    """Without duplicates the conversions equal scipy's canonical CSR / CSC."""
    rng = np.random.default_rng(2)
    m, n = 300, 200
    dense = (rng.random((m, n)) < 0.05) * rng.uniform(-1, 1, (m, n))
    dense[7, :] = 0.0  # an empty row
    dense[:, 11] = 0.0  # an empty column
    row, col = np.nonzero(dense)
    perm = rng.permutation(len(row))
    row, col = row[perm].astype(np.int32), col[perm].astype(np.int32)
    val = dense[row, col]
    v, rp, ci = oracle.coo_to_csr(m, n, row, col, val)
    ref = sp.csr_matrix(dense)
    ref.sort_indices()
    assert np.array_equal(rp, ref.indptr) and np.array_equal(ci, ref.indices)
    assert np.array_equal(v, ref.data)
    v, cp, ri = oracle.coo_to_csc(m, n, row, col, val)
    ref = sp.csc_matrix(dense)
    ref.sort_indices()
    assert np.array_equal(cp, ref.indptr) and np.array_equal(ri, ref.indices)
    assert np.array_equal(v, ref.data)


def _write(path, text):
    with open(path, "w", newline="") as f:
        f.write(text)
    return str(path)


def test_mm_read_real_general_vs_scipy(tmp_path):
    from hpda_sblas_pnnl_amd.spmv import mm_read_coo
    rng = np.random.default_rng(3)
    A = sp.random(120, 90, density=0.05, random_state=4, format="coo")
    A.data = rng.standard_normal(A.nnz) * 1e3
    path = str(tmp_path / "a.mtx")
    scipy.io.mmwrite(path, A, symmetry="general", precision=17)
    m, n, row, col, val, sym = mm_read_coo(path)
    B = scipy.io.mmread(path).tocoo()
    assert (m, n, sym) == (120, 90, "general")
    assert np.array_equal(row, B.row) and np.array_equal(col, B.col)
    assert np.array_equal(val, B.data)  # correctly rounded parse on both sides


def test_mm_read_integer_pattern_comments(tmp_path):
    from hpda_sblas_pnnl_amd.spmv import mm_read_coo
    text = ("%%MatrixMarket MATRIX Coordinate INTEGER General\r\n"
            "% a comment\r\n%another\r\n"
            "3 4 3\r\n1 1 7\r\n3 4 -2\r\n\r\n2 2 +5\r\n")
    m, n, row, col, val, sym = mm_read_coo(_write(tmp_path / "i.mtx", text))
    assert (m, n, sym) == (3, 4, "general")
    assert row.tolist() == [0, 2, 1] and col.tolist() == [0, 3, 1]
    assert val.tolist() == [7.0, -2.0, 5.0]
    # scipy refuses the explicit '+' sign that fscanf("%d") accepts
    B = scipy.io.mmread(_write(tmp_path / "i2.mtx", text.replace("+5", "5"))).tocoo()
    assert sorted(zip(B.row, B.col, B.data)) == sorted(zip(row, col, val))
    text = "%%MatrixMarket matrix coordinate pattern general\n2 2 2\n1 2\n2 1\n"
    m, n, row, col, val, sym = mm_read_coo(_write(tmp_path / "p.mtx", text))
    assert row.tolist() == [0, 1] and col.tolist() == [1, 0] and val.tolist() == [1.0, 1.0]


def test_mm_read_symmetric_not_expanded(tmp_path):
    """The stored triangle comes back as stored (the harness does not expand
    it); expanding it by hand gives scipy's matrix."""
    from hpda_sblas_pnnl_amd.spmv import mm_read_coo
    text = ("%%MatrixMarket matrix coordinate real symmetric\n"
            "3 3 4\n1 1 2.5\n2 1 -1\n3 2 0.125\n3 3 4\n")
    path = _write(tmp_path / "s.mtx", text)
    m, n, row, col, val, sym = mm_read_coo(path)
    assert sym == "symmetric" and len(val) == 4
    full = sp.coo_matrix((val, (row, col)), shape=(m, n)).toarray()
    full = full + np.tril(full, -1).T
    assert np.array_equal(full, scipy.io.mmread(path).toarray())


@pytest.mark.parametrize("text,status", [
    ("%%MatrixMarket matrix coordinate complex general\n1 1 1\n1 1 1 0\n", 7),
    ("%%MatrixMarket matrix array real general\n1 1\n1\n", 7),
    ("%MatrixMarket matrix coordinate real general\n1 1 1\n1 1 1\n", 1),
    ("%%MatrixMarket matrix coordinate real general\n2 2 2\n1 1 1\n", 1),
    ("%%MatrixMarket matrix coordinate real general\n2 2 1\n3 1 1\n", 1),
    ("%%MatrixMarket matrix coordinate real general\n2 2 1\n1 1 abc\n", 1),
    ("%%MatrixMarket matrix coordinate real\n2 2 1\n1 1 1\n", 1),
])
def test_mm_read_errors(tmp_path, text, status):
    from hpda_sblas_pnnl_amd._lib import SblasError
    from hpda_sblas_pnnl_amd.spmv import mm_read_coo
    with pytest.raises(SblasError) as e:
        mm_read_coo(_write(tmp_path / "bad.mtx", text))
    assert e.value.status == status
    with pytest.raises(SblasError):
        mm_read_coo(str(tmp_path / "missing.mtx"))


def test_mm_read_empty(tmp_path):
    from hpda_sblas_pnnl_amd.spmv import mm_read_coo
    m, n, row, col, val, sym = mm_read_coo(
        _write(tmp_path / "e.mtx", "%%MatrixMarket matrix coordinate real general\n5 6 0\n"))
    assert (m, n, len(row), len(col), len(val)) == (5, 6, 0, 0, 0)


def test_helpers_reject_bad_arguments_without_touching_arrays():
    """Bad sizes fail with a status before any device work (so this runs
    without a GPU) and leave the arrays as they were."""
    from hpda_sblas_pnnl_amd._lib import SblasError
    from hpda_sblas_pnnl_amd.spmv import coo2csr, sortCOORow
    row = np.array([1, 0], np.int32)
    col = np.array([0, 1], np.int32)
    val = np.array([1.0, 2.0])
    with pytest.raises(SblasError) as e:
        sortCOORow(-1, 2, 2, val, row, col)
    assert e.value.status == 1
    assert row.tolist() == [1, 0] and val.tolist() == [1.0, 2.0]
    with pytest.raises(SblasError):
        coo2csr(2, -2, 2, val, row, col)
    with pytest.raises(TypeError):
        sortCOORow(2, 2, 2, val.astype(np.float32), row, col)


def test_oracle_drop_zero_x_brute_force(oracle):
    """The filter against the reference's host loop (src/dspmv_mgpu_v1.cu:
    35-46) restated as plain Python: -0.0 dropped, NaN kept."""
    rng = np.random.default_rng(6)
    m, n = 50, 40
    lens = rng.integers(0, 9, m)
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    col = rng.integers(0, n, rp[-1]).astype(np.int32)
    val = rng.uniform(-1, 1, rp[-1])
    x = rng.uniform(-1, 1, n)
    x[rng.random(n) < 0.4] = 0.0
    x[3], x[4] = -0.0, np.nan
    rp2, c2, v2 = [0], [], []
    for i in range(m):
        for j in range(rp[i], rp[i + 1]):
            if x[col[j]] != 0.0:
                v2.append(val[j])
                c2.append(col[j])
        rp2.append(len(v2))
    r, c, v = oracle.csr_drop_zero_x(m, rp, col, val, x)
    assert r.tolist() == rp2 and c.tolist() == c2 and v.tolist() == v2
