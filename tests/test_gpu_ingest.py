"""GPU parity of matrix ingestion (src/spmv_helper.cu:277-642): the device
COO sorts and conversions against the oracle's stable order, bit for bit,
and a Matrix Market file through the reference's harness path
(test/dspmv_test_dgx1.cu:133-199) into an SpMV."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _coo(rng, m, n, nnz, dup_frac=0.2):
    row = rng.integers(0, m, nnz).astype(np.int32)
    col = rng.integers(0, n, nnz).astype(np.int32)
    ndup = int(nnz * dup_frac)
    if ndup:
        src = rng.integers(0, nnz, ndup)
        dst = rng.integers(0, nnz, ndup)
        row[dst], col[dst] = row[src], col[src]
    return row, col, rng.uniform(-1, 1, nnz)


@pytest.mark.parametrize("m,n,nnz", [(5000, 7000, 200_000), (1, 3, 5), (70_000, 3, 50_000),
                                     (3, 100_000, 40_000)])
def test_sort_helpers(oracle, m, n, nnz):
    """sortCOORow / sortCOOCol in place, duplicates kept in input order."""
    from hpda_sblas_pnnl_amd.spmv import sortCOOCol, sortCOORow
    rng = np.random.default_rng(m + n)
    row, col, val = _coo(rng, m, n, nnz)
    for fn, by_col in ((sortCOORow, False), (sortCOOCol, True)):
        r, c, v = row.copy(), col.copy(), val.copy()
        fn(m, n, nnz, v, r, c)
        er, ec, ev = oracle.coo_sort(row, col, val, by_col=by_col)
        assert np.array_equal(r, er) and np.array_equal(c, ec) and np.array_equal(v, ev)


def test_convert_helpers(oracle):
    """coo2csr / coo2csc: pointers, indices and values bit for bit; the COO
    input untouched; empty rows and columns at both ends and inside."""
    from hpda_sblas_pnnl_amd.spmv import coo2csc, coo2csr
    rng = np.random.default_rng(7)
    m, n, nnz = 9000, 6000, 150_000
    row, col, val = _coo(rng, m, n, nnz)
    keep = (row % 97 != 3) & (col % 89 != 5) & (row > 2) & (row < m - 3) & (col > 1) & (col < n - 4)
    row, col, val = row[keep], col[keep], val[keep]
    nnz = len(row)
    r0, c0, v0 = row.copy(), col.copy(), val.copy()
    got = coo2csr(m, n, nnz, val, row, col)
    for g, e in zip(got, oracle.coo_to_csr(m, n, row, col, val)):
        assert np.array_equal(g, e)
    got = coo2csc(m, n, nnz, val, row, col)
    for g, e in zip(got, oracle.coo_to_csc(m, n, row, col, val)):
        assert np.array_equal(g, e)
    assert np.array_equal(row, r0) and np.array_equal(col, c0) and np.array_equal(val, v0)


def test_convert_empty_and_bad_indices():
    from hpda_sblas_pnnl_amd._lib import SblasError
    from hpda_sblas_pnnl_amd.spmv import coo2csr, sortCOORow
    v, rp, ci = coo2csr(4, 4, 0, np.zeros(0), np.zeros(0, np.int32), np.zeros(0, np.int32))
    assert rp.tolist() == [0] * 5 and len(v) == 0 and len(ci) == 0
    row = np.array([0, 4, 1], np.int32)  # 4 is out of range for m = 4
    col = np.array([1, 0, 2], np.int32)
    val = np.array([1.0, 2.0, 3.0])
    with pytest.raises(SblasError) as e:
        sortCOORow(4, 4, 3, val, row, col)
    assert e.value.status == 1 and row.tolist() == [0, 4, 1]
    with pytest.raises(SblasError):
        coo2csr(4, 3, 3, val, np.array([0, 1, 1], np.int32), np.array([0, 3, 1], np.int32))


@pytest.mark.parametrize("dtype", ["f64", "f32", "bf16"])
def test_device_sort_and_pointers(oracle, dtype):
    """Device tensors, every value type: values ride along bit for bit."""
    import torch
    from hpda_sblas_pnnl_amd.device import coo_ptr, coo_sort
    rng = np.random.default_rng(11)
    m, n, nnz = 20_000, 30_000, 300_000
    row, col, val = _coo(rng, m, n, nnz)
    if dtype == "f32":
        val = val.astype(np.float32)
    elif dtype == "bf16":
        val = (val.astype(np.float32).view(np.uint32) >> 16).astype(np.uint16).view(np.int16)
    for by_col in (False, True):
        dr, dc, dv = (torch.from_numpy(a.copy()).cuda() for a in (row, col, val))
        coo_sort(dr, dc, dv, m, n, by_col=by_col)
        er, ec, ev = oracle.coo_sort(row, col, val, by_col=by_col)
        assert np.array_equal(dr.cpu().numpy(), er) and np.array_equal(dc.cpu().numpy(), ec)
        assert np.array_equal(dv.cpu().numpy(), ev)
        major, cnt = (dc, n) if by_col else (dr, m)
        ptr = coo_ptr(major, cnt)
        assert np.array_equal(ptr.cpu().numpy(), oracle.coo_ptr(cnt, ec if by_col else er))


def test_rmat_roundtrip_full_size():
    """R-MAT scale 22 (about 6.9e7 elements) as COO in a random order: the
    device row sort must give back the generator's CSR exactly (its pairs are
    unique, so the sorted order is unique)."""
    import torch
    from hpda_sblas_pnnl_amd import gen as G
    from hpda_sblas_pnnl_amd.device import coo_ptr, coo_sort
    A = G.rmat_csr(22, 17, seed=0)
    rp = torch.from_numpy(A.row_ptr).cuda()
    col = torch.from_numpy(A.col).cuda()
    val = torch.from_numpy(A.val).cuda()
    row = torch.repeat_interleave(torch.arange(A.m, dtype=torch.int32, device="cuda"),
                                  torch.diff(rp).long())
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    perm = torch.randperm(A.nnz, device="cuda", generator=g)
    r, c, v = row[perm].contiguous(), col[perm].contiguous(), val[perm].contiguous()
    coo_sort(r, c, v, A.m, A.n)
    assert torch.equal(r, row) and torch.equal(c, col) and torch.equal(v, val)
    assert torch.equal(coo_ptr(r, A.m), rp)


def test_matrix_market_to_spmv(oracle, tmp_path):
    """The harness's file path: .mtx -> COO -> coo2csr -> spMV_mgpu_v1, within
    1e-12 of the oracle on the same COO."""
    import scipy.io
    import scipy.sparse as sp
    from hpda_sblas_pnnl_amd.spmv import coo2csr, mm_read_coo, spMV_mgpu_v1
    rng = np.random.default_rng(13)
    A = sp.random(3000, 2500, density=0.01, random_state=14, format="coo")
    A.data = rng.standard_normal(A.nnz)
    path = str(tmp_path / "m.mtx")
    scipy.io.mmwrite(path, A, precision=17)
    m, n, row, col, val, _ = mm_read_coo(path)
    v, rp, ci = coo2csr(m, n, len(val), val, row, col)
    x = rng.uniform(-1, 1, n)
    y = rng.uniform(-1, 1, m)
    er, ec, ev = oracle.coo_sort(row, col, val)
    ref = oracle.coo_spmv(m, n, ev, er, ec, x, 1.5, -0.5, y)
    spMV_mgpu_v1(m, n, len(v), 1.5, v, rp, ci, x, -0.5, y, 1, 1)
    assert oracle.err_normwise(y, ref) <= 1e-12


def test_device_drop_zero_x(oracle):
    """spMspV's filter on the device: bit for bit the host loop's arrays,
    -0.0 dropped and NaN kept, empty rows and rows emptied by the filter."""
    import torch
    from conftest import random_csr
    from hpda_sblas_pnnl_amd.device import csr_drop_zero_x
    rng = np.random.default_rng(21)
    m, n = 30_000, 20_000
    rp, col, val = random_csr(rng, m, n, 12, empty_frac=0.2)
    x = rng.uniform(-1, 1, n)
    x[rng.random(n) < 0.6] = 0.0
    x[rng.random(n) < 0.01] = -0.0
    x[5] = np.nan
    d = [torch.from_numpy(a).cuda() for a in (rp, col, val, x)]
    r, c, v = csr_drop_zero_x(*d, m, n)
    er, ec, ev = oracle.csr_drop_zero_x(m, rp, col, val, x)
    assert np.array_equal(r.cpu().numpy(), er) and np.array_equal(c.cpu().numpy(), ec)
    assert np.array_equal(v.cpu().numpy(), ev)


def test_spmspv_skips_zero_x_products(oracle):
    """spMspV drops the products with x[col] == 0 before multiplying: an
    infinite value facing a zero x leaves its row finite (spMV would give
    NaN), and the result equals the oracle on the filtered matrix.  v2 with
    many tasks after the filter runs fused."""
    from conftest import random_csr
    from hpda_sblas_pnnl_amd import spmv as S
    rng = np.random.default_rng(22)
    m, n = 5000, 4000
    rp, col, val = random_csr(rng, m, n, 25)
    x = rng.uniform(-1, 1, n)
    x[rng.random(n) < 0.5] = 0.0
    j = int(np.flatnonzero(x[col] == 0.0)[0])
    val[j] = np.inf
    y0 = rng.uniform(-1, 1, m)
    fr, fc, fv = oracle.csr_drop_zero_x(m, rp, col, val, x)
    ref = oracle.csr_spmv(m, n, fv, fr, fc, x, 1.25, 0.5, y0)
    assert np.all(np.isfinite(ref))
    for call in (lambda y: S.spMspV_mgpu_v1(m, n, len(val), 1.25, val, rp, col, x, 0.5, y, 1, 1),
                 lambda y: S.spMspV_mgpu_v1(m, n, len(val), 1.25, val, rp, col, x, 0.5, y, 1, 3),
                 lambda y: S.spMspV_mgpu_v2(m, n, len(val), 1.25, val, rp, col, x, 0.5, y, 1, 2,
                                            700, 4),
                 lambda y: S.spMspV_mgpu_v2(m, n, len(val), 1.25, val, rp, col, x, 0.5, y, 1, 1,
                                            100, 2)):
        y = y0.copy()
        call(y)
        assert oracle.err_normwise(y, ref) <= 1e-12
    y = y0.copy()
    S.spMV_mgpu_v1(m, n, len(val), 1.25, val, rp, col, x, 0.5, y, 1, 1)
    assert not np.all(np.isfinite(y))
