"""GPU parity of the CSR path against the oracle (host CSR loop of
test/dspmv_test_smt.cu:39-52), through the C ABI.

FP64 bar: normwise relative error <= 1e-12 (summation order is the only
difference); integer-valued inputs are exact in any order, so those compare
bitwise.  Reduced-precision values compare against the oracle run on the same
rounded values (FP64 accumulation on both sides), same 1e-12 bar.
"""
import os

import numpy as np
import pytest

from conftest import random_csr

pytestmark = pytest.mark.gpu

TOL = 1e-12


@pytest.fixture(scope="module")
def S():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    from hpda_sblas_pnnl_amd import spmv
    return spmv


@pytest.fixture
def chunks(monkeypatch):
    def set_chunks(c):
        monkeypatch.setenv("SBLAS_CHUNKS", str(c))
    return set_chunks


def _ref(oracle, m, n, val, rp, col, x, alpha, beta, y0):
    return oracle.csr_spmv(m, n, val, rp, col, x, alpha, beta, y0)


def _csc_of(m, n, rp, col, val):
    from hpda_sblas_pnnl_amd import gen as G
    A = G.CSR(m, n, np.asarray(rp, np.int32), np.asarray(col, np.int32), np.asarray(val))
    return G.to_csc(A)


ENTRY_POINTS = ["v1_k1", "v1_k2", "v1_k3", "baseline", "v1_numa", "v2", "baseline_coo", "v1_numa_coo",
                "baseline_csc", "v1_numa_csc"]


def _call(S, ep, m, n, rp, col, val, x, alpha, beta, y, kernel=1):
    nnz = len(val)
    if ep.startswith("v1_k"):
        return S.spMV_mgpu_v1(m, n, nnz, alpha, val, rp, col, x, beta, y, 1, int(ep[-1]))
    if ep == "baseline":
        return S.spMV_mgpu_baseline(m, n, nnz, alpha, val, rp, col, x, beta, y, 1)
    if ep == "v1_numa":
        return S.spMV_mgpu_v1_numa(m, n, nnz, alpha, val, rp, col, x, beta, y, 1, kernel, [0], 0, 1)
    if ep == "v2":
        return S.spMV_mgpu_v2(m, n, nnz, alpha, val, rp, col, x, beta, y, 1, kernel,
                              max(1, nnz // 5), 2)
    ri = np.repeat(np.arange(m, dtype=np.int32), np.diff(rp))
    if ep == "baseline_coo":
        return S.spMV_mgpu_baseline_coo(m, n, nnz, alpha, val, ri, col, x, beta, y, 1)
    if ep == "v1_numa_coo":
        return S.spMV_mgpu_v1_numa_coo(m, n, nnz, alpha, val, ri, col, x, beta, y, 1, kernel,
                                       [0], 1, 1)
    cp, rows, cv = _csc_of(m, n, rp, col, val)
    if ep == "baseline_csc":
        return S.spMV_mgpu_baseline_csc(m, n, nnz, alpha, cv, cp, rows, x, beta, y, 1)
    if ep == "v1_numa_csc":
        return S.spMV_mgpu_v1_numa_csc(m, n, nnz, alpha, cv, cp, rows, x, beta, y, 1, kernel,
                                       [0], 0, 0)
    raise AssertionError(ep)


@pytest.mark.parametrize("ep", ENTRY_POINTS)
def test_g_matrix_known_answer(S, oracle, ep):
    """'g' matrix, x = 1, alpha = 1, beta = 0: y_i equals the row length
    exactly (test/dspmv_test_smt.cu:176-232, :359)."""
    n = 800
    r, c, v = oracle.gen_g(n)
    rp = np.zeros(n + 1, np.int32)
    np.add.at(rp, r + 1, 1)
    rp = np.cumsum(rp).astype(np.int32)
    y = np.zeros(n)
    ret = _call(S, ep, n, n, rp, c, v, np.ones(n), 1.0, 0.0, y)
    want = np.where(np.arange(n) < n // 8, 720.0, 8.0)
    assert np.array_equal(y, want), ep
    assert ret.comp_time >= 0.0


@pytest.mark.parametrize("ep", ENTRY_POINTS)
@pytest.mark.parametrize("C", [1, 2, 3, 7, 8, 16])
def test_entry_points_beta_nonzero(S, oracle, chunks, ep, C):
    """Random matrix with empty first/last rows and empty rows at chunk edges,
    alpha, beta != 0 and y != 0, at several chunk counts."""
    chunks(C)
    rng = np.random.default_rng(100 + C)
    m, n = 3000, 2500
    rp, col, val = random_csr(rng, m, n, 12, empty_frac=0.3, long_rows=2, long_len=2000)
    lens = np.diff(rp)
    # force empty leading / trailing rows
    keep = np.ones(len(val), bool)
    keep[rp[0]:rp[2]] = False
    keep[rp[m - 2]:rp[m]] = False
    lens[:2] = 0
    lens[-2:] = 0
    col, val = col[keep], val[keep]
    rp = np.concatenate([[0], np.cumsum(lens)]).astype(np.int32)
    x = rng.uniform(-1, 1, n)
    y0 = rng.uniform(-1, 1, m)
    y = y0.copy()
    _call(S, ep, m, n, rp, col, val, x, 1.25, -0.5, y)
    ref = _ref(oracle, m, n, val, rp, col, x, 1.25, -0.5, y0)
    assert oracle.err_normwise(y, ref) <= TOL, ep
    assert np.array_equal(y[:2], -0.5 * y0[:2]) and np.array_equal(y[-2:], -0.5 * y0[-2:])


@pytest.mark.parametrize("C", [1, 2, 5, 8, 13])
def test_integer_inputs_bitwise(S, oracle, chunks, C):
    chunks(C)
    rng = np.random.default_rng(7)
    m, n = 5000, 5000
    rp, col, _ = random_csr(rng, m, n, 40, empty_frac=0.1, long_rows=3, long_len=4000)
    val = rng.integers(-8, 9, size=len(col)).astype(np.float64)
    x = rng.integers(-16, 17, size=n).astype(np.float64)
    y0 = rng.integers(-4, 5, size=m).astype(np.float64)
    for ep in ("v1_k1", "v1_numa", "v2"):
        y = y0.copy()
        _call(S, ep, m, n, rp, col, val, x, 2.0, 3.0, y)
        assert np.array_equal(y, _ref(oracle, m, n, val, rp, col, x, 2.0, 3.0, y0)), ep


def test_powerlaw_config(S, oracle):
    """n = 20,000 power-law matrix, alpha = 1, beta = 0, seed 0."""
    from hpda_sblas_pnnl_amd import gen as G
    A = G.powerlaw_csr(20000, seed=0)
    x = G.values(A.n, 0, 2, "uniform")
    for ep in ("v1_k1", "baseline", "v1_numa"):
        y = np.zeros(A.m)
        _call(S, ep, A.m, A.n, A.row_ptr, A.col, A.val, x, 1.0, 0.0, y)
        ref = _ref(oracle, A.m, A.n, A.val, A.row_ptr, A.col, x, 1.0, 0.0, np.zeros(A.m))
        assert oracle.err_normwise(y, ref) <= TOL, ep


@pytest.mark.parametrize("C", [1, 4, 64])
def test_degenerate_shapes(S, oracle, chunks, C):
    chunks(C)
    # one long row, then nothing
    m, n = 4, 10000
    rp = np.array([0, 0, 9000, 9000, 9000], np.int32)
    col = np.arange(9000, dtype=np.int32)
    rng = np.random.default_rng(3)
    val = rng.uniform(-1, 1, 9000)
    x = rng.uniform(-1, 1, n)
    y0 = rng.uniform(-1, 1, m)
    for ep in ("v1_k1", "v1_numa", "v2", "baseline"):
        y = y0.copy()
        _call(S, ep, m, n, rp, col, val, x, 0.5, 2.0, y)
        assert oracle.err_normwise(y, _ref(oracle, m, n, val, rp, col, x, 0.5, 2.0, y0)) <= TOL
    # no nonzeros at all: y = beta * y
    rp0 = np.zeros(m + 1, np.int32)
    y = y0.copy()
    S.spMV_mgpu_v1(m, n, 0, 1.0, np.zeros(0), rp0, np.zeros(0, np.int32), x, 0.25, y, 1, 1)
    assert np.array_equal(y, 0.25 * y0)
    # fewer nonzeros than chunks
    rp1 = np.array([0, 1, 1, 2, 3], np.int32)
    c1 = np.array([5, 7, 9], np.int32)
    v1 = np.array([1.5, -2.0, 4.0])
    y = y0.copy()
    S.spMV_mgpu_v1(m, n, 3, 1.0, v1, rp1, c1, x, -1.0, y, 1, 1)
    assert oracle.err_normwise(y, _ref(oracle, m, n, v1, rp1, c1, x, 1.0, -1.0, y0)) <= TOL


def test_spmspv_drops_zero_x(S, oracle):
    rng = np.random.default_rng(5)
    m, n = 2000, 2000
    rp, col, val = random_csr(rng, m, n, 20)
    x = rng.uniform(-1, 1, n)
    x[rng.random(n) < 0.5] = 0.0
    y0 = rng.uniform(-1, 1, m)
    ref = _ref(oracle, m, n, val, rp, col, x, 1.0, 0.5, y0)
    y = y0.copy()
    S.spMspV_mgpu_v1(m, n, len(val), 1.0, val, rp, col, x, 0.5, y, 1, 1)
    assert oracle.err_normwise(y, ref) <= TOL
    y = y0.copy()
    S.spMspV_mgpu_v2(m, n, len(val), 1.0, val, rp, col, x, 0.5, y, 1, 1, 5000, 2)
    assert oracle.err_normwise(y, ref) <= TOL


# ---------------------------------------------------------------------------
# plan API on device tensors
# ---------------------------------------------------------------------------

def _dev(*arrs):
    import torch
    return [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in arrs]


@pytest.mark.parametrize("kernel", ["vector", "auto"])
@pytest.mark.parametrize("chunks_", [1, 8])
def test_plan_fp64_rmat(oracle, kernel, chunks_):
    import torch
    from hpda_sblas_pnnl_amd import gen as G
    from hpda_sblas_pnnl_amd.device import SpmvPlan
    A = G.rmat_csr(scale=14, edge_factor=17, seed=0)
    x = G.values(A.n, 0, 2, "uniform")
    y0 = G.values(A.m, 0, 3, "uniform")
    rp, col, val, dx, dy = _dev(A.row_ptr, A.col, A.val, x, y0)
    plan = SpmvPlan("csr", A.m, A.n, rp, col, val, kernel=kernel, chunks=chunks_)
    plan.spmv(dx, dy, 1.0, 0.5)
    torch.cuda.synchronize()
    ref = oracle.csr_spmv(A.m, A.n, A.val, A.row_ptr, A.col, x, 1.0, 0.5, y0)
    assert oracle.err_normwise(dy.cpu().numpy(), ref) <= TOL
    # binary values, integer-valued x: exact
    xi = np.round(x * 8)
    dxi = _dev(xi)[0]
    dy2 = torch.zeros(A.m, dtype=torch.float64, device="cuda")
    plan.spmv(dxi, dy2, 1.0, 0.0)
    torch.cuda.synchronize()
    assert np.array_equal(dy2.cpu().numpy(),
                          oracle.csr_spmv(A.m, A.n, A.val, A.row_ptr, A.col, xi, 1.0, 0.0,
                                          np.zeros(A.m)))


@pytest.mark.parametrize("nx", [1, 2, 0])
@pytest.mark.parametrize("vt", ["f32", "bf16"])
def test_plan_reduced_precision_values(oracle, vt, nx):
    """Narrow values in row blocks over 3 chunks: with x gathered in FP64
    (narrow_x 1, and auto on this matrix of fewer than 2^20 elements) within
    1e-12 of the oracle on the same rounded values; with x gathered from an
    FP32 copy (narrow_x 2) x's rounding is added: every row within 2^-22 of
    sum |a_ij x_j| (max-row error)."""
    import torch
    from hpda_sblas_pnnl_amd import gen as G
    from hpda_sblas_pnnl_amd.device import SpmvPlan
    A = G.powerlaw_csr(20000, seed=0)
    x = G.values(A.n, 0, 2, "uniform")
    y0 = G.values(A.m, 0, 3, "uniform")
    if vt == "f32":
        v = A.val.astype(np.float32)
        vr = v.astype(np.float64)
        ref = oracle.csr_spmv_f32val(A.m, v, A.row_ptr, A.col, x, 1.0, 0.5, y0)
        dv = torch.from_numpy(v).cuda()
    else:
        bits = G.to_bf16_bits(A.val)
        vr = G.from_bf16_bits(bits)
        ref = oracle.csr_spmv_bf16val(A.m, bits, A.row_ptr, A.col, x, 1.0, 0.5, y0)
        dv = torch.from_numpy(bits.view(np.int16)).cuda()
    rp, col, dx, dy = _dev(A.row_ptr, A.col, x, y0)
    plan = SpmvPlan("csr", A.m, A.n, rp, col, dv, kernel="vector", chunks=3, narrow_x=nx)
    assert plan.info()["narrow_x"] == (1 if nx == 2 else 0)
    plan.spmv(dx, dy, 1.0, 0.5)
    torch.cuda.synchronize()
    got = dy.cpu().numpy()
    if nx != 2:
        assert oracle.err_normwise(got, ref) <= TOL
    else:
        assert oracle.err_maxrow(A.m, vr, A.row_ptr, A.col, x, 1.0, 0.5, y0, got, ref) <= 2.0**-22
        assert oracle.err_normwise(got, ref) <= 1e-6


@pytest.mark.parametrize("k", [2, 8])
def test_plan_spmm(oracle, k):
    import torch
    from hpda_sblas_pnnl_amd import gen as G
    from hpda_sblas_pnnl_amd.device import SpmvPlan
    A = G.powerlaw_csr(20000, seed=1)
    X = G.values(A.n * k, 0, 4, "uniform").reshape(A.n, k)
    Y0 = G.values(A.m * k, 0, 5, "uniform").reshape(A.m, k)
    rp, col, val, dX, dY = _dev(A.row_ptr, A.col, A.val, X, Y0)
    plan = SpmvPlan("csr", A.m, A.n, rp, col, val, kernel="vector", chunks=4)
    plan.spmm(dX, dY, 0.75, -0.25)
    torch.cuda.synchronize()
    ref = oracle.csr_spmm(A.m, k, A.val, A.row_ptr, A.col, X, 0.75, -0.25, Y0)
    assert oracle.err_normwise(dY.cpu().numpy().ravel(), ref.ravel()) <= TOL


@pytest.mark.parametrize("k", [2, 4, 8, 16])
@pytest.mark.parametrize("variant,chunks_", [(2, 1), (2, 3), (1, 3)])
def test_plan_spmm_long_rows(oracle, k, variant, chunks_):
    """Multi-RHS blocks on rows far longer than one row block (pieces merged in
    order), empty rows and rows cut by chunk edges, beta != 0."""
    import torch
    from hpda_sblas_pnnl_amd.device import SpmvPlan
    rng = np.random.default_rng(60 + k)
    m, n = 3000, 30000
    rp, col, val = random_csr(rng, m, n, 12, empty_frac=0.3, long_rows=3, long_len=5 * 2048 + 7)
    X = rng.uniform(-1, 1, (n, k))
    Y0 = rng.uniform(-1, 1, (m, k))
    drp, dcol, dval, dX, dY = _dev(rp, col, val, X, Y0)
    plan = SpmvPlan("csr", m, n, drp, dcol, dval, kernel="vector", chunks=chunks_, variant=variant)
    plan.spmm(dX, dY, -1.25, 0.5)
    torch.cuda.synchronize()
    ref = oracle.csr_spmm(m, k, val, rp, col, X, -1.25, 0.5, Y0)
    assert oracle.err_normwise(dY.cpu().numpy().ravel(), ref.ravel()) <= TOL


def test_plan_streams_mode_matches_fused(oracle):
    import torch
    from hpda_sblas_pnnl_amd import gen as G
    from hpda_sblas_pnnl_amd.device import SpmvPlan
    A = G.banded_csr(100000, seed=0, halfwidth=5000)
    x = G.values(A.n, 0, 2, "uniform")
    rp, col, val, dx = _dev(A.row_ptr, A.col, A.val, x)
    outs = []
    # same sub-wave kernel launched fused (one launch) and per chunk on 4 streams
    for streams in (1, 4):
        plan = SpmvPlan("csr", A.m, A.n, rp, col, val, kernel="vector", chunks=8,
                        streams=streams, variant=1)
        dy = torch.zeros(A.m, dtype=torch.float64, device="cuda")
        plan.spmv(dx, dy)
        torch.cuda.synchronize()
        outs.append(dy.cpu().numpy())
        del plan
    assert np.array_equal(outs[0], outs[1])
    ref = oracle.csr_spmv(A.m, A.n, A.val, A.row_ptr, A.col, x, 1.0, 0.0, np.zeros(A.m))
    assert oracle.err_normwise(outs[0], ref) <= TOL
    # the row-block form over the same chunks
    plan = SpmvPlan("csr", A.m, A.n, rp, col, val, kernel="vector", chunks=8, variant=2)
    assert plan.info()["variant"] == 2
    dy = torch.zeros(A.m, dtype=torch.float64, device="cuda")
    plan.spmv(dx, dy)
    torch.cuda.synchronize()
    assert oracle.err_normwise(dy.cpu().numpy(), ref) <= TOL
