"""GPU parity at the full shapes of the benchmark workloads, through the C ABI
(plans on device arrays and the reference's drop-in entry points), against
the oracle's host loops (test/dspmv_test_smt.cu:39-52 and its FP32 / BF16 /
multi-vector forms):

* R-MAT scale 22 (edge factor 17, binary values, alpha 1, beta 0.5, y != 0):
  kernels 1/2/3 and the chooser's pick, the COO and CSC paths, FP32 and BF16
  values (exact for binary values) and a k = 8 block of right-hand sides;
* the power-law n = 20,000 matrix with FP32 / BF16 values;
* the banded n = 2e6 matrix (2.8e8 elements) cut into 8 chunks: CSR with the
  v1_numa rule, COO and CSC drop-ins, and the v2 form streamed from host
  memory;
* the 64-matrix batch with FP32 values in one replayed hipGraph.

Bars: FP64 paths within 1e-12 normwise of the FP64 oracle (summation order is
the only difference).  Narrow-value paths within 1e-12 of the oracle run on
the same rounded values; FP32 values within 1e-6 of the FP64 oracle (the
stated error of the FP32-value form).
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-12
TOL_FP32_VS_FP64 = 1e-6


def _threads():
    return max(1, min(16, len(os.sched_getaffinity(0))))


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    return torch


@pytest.fixture(scope="module")
def rmat(torch_dev):
    from hpda_sblas_pnnl_amd import gen as G
    A = G.rmat_csr(scale=22, edge_factor=17, seed=0)
    x = G.values(A.n, 0, 2, "uniform")
    y0 = G.values(A.m, 0, 3, "uniform")
    return A, x, y0


@pytest.fixture(scope="module")
def rmat_ref(oracle, rmat):
    A, x, y0 = rmat
    return oracle.csr_spmv(A.m, A.n, A.val, A.row_ptr, A.col, x, 1.0, 0.5, y0,
                           threads=_threads())


@pytest.fixture(scope="module")
def banded(torch_dev):
    from hpda_sblas_pnnl_amd import gen as G
    A = G.banded_csr(2_000_000, seed=0)
    x = G.values(A.n, 0, 2, "uniform")
    y0 = G.values(A.m, 0, 3, "uniform")
    return A, x, y0


@pytest.fixture(scope="module")
def banded_ref(oracle, banded):
    A, x, y0 = banded
    return oracle.csr_spmv(A.m, A.n, A.val, A.row_ptr, A.col, x, 1.0, 0.5, y0,
                           threads=_threads())


def _dev(torch, *arrays):
    return [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in arrays]


def _plan_run(torch, plan, x, y0, alpha, beta):
    dx, dy = _dev(torch, x, y0)
    plan.spmv(dx, dy, alpha, beta)
    torch.cuda.synchronize()
    return dy.cpu().numpy()


# ---------------------------------------------------------------------------
# R-MAT scale 22
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("kernel", ["auto", "vector", "merge", "csr5"])
def test_rmat22_csr_kernels(torch_dev, oracle, rmat, rmat_ref, kernel):
    from hpda_sblas_pnnl_amd.device import SpmvPlan
    torch = torch_dev
    A, x, y0 = rmat
    rp, col, val = _dev(torch, A.row_ptr, A.col, A.val)
    plan = SpmvPlan("csr", A.m, A.n, rp, col, val, kernel=kernel)
    got = _plan_run(torch, plan, x, y0, 1.0, 0.5)
    assert oracle.err_normwise(got, rmat_ref) <= TOL, plan.info()
    # the 51 % empty rows hold exactly beta * y0
    empty = np.diff(A.row_ptr) == 0
    assert empty.mean() > 0.4
    assert np.array_equal(got[empty], 0.5 * y0[empty])


def test_rmat22_coo_csc_plans(torch_dev, oracle, rmat, rmat_ref):
    from hpda_sblas_pnnl_amd import gen as G
    from hpda_sblas_pnnl_amd.device import SpmvPlan
    torch = torch_dev
    A, x, y0 = rmat
    ri, col, val = _dev(torch, G.coo_rows(A), A.col, A.val)
    got = _plan_run(torch, SpmvPlan("coo", A.m, A.n, ri, col, val), x, y0, 1.0, 0.5)
    assert oracle.err_normwise(got, rmat_ref) <= TOL
    del ri, col, val
    cp, rows, cv = G.to_csc(A)
    dcp, drows, dcv = _dev(torch, cp, rows, cv)
    got = _plan_run(torch, SpmvPlan("csc", A.m, A.n, dcp, drows, dcv), x, y0, 1.0, 0.5)
    assert oracle.err_normwise(got, rmat_ref) <= TOL


@pytest.mark.parametrize("entry", ["v1_k1", "v1_k2", "v1_k3", "baseline_coo", "v1_numa_coo",
                                   "v1_numa_csc"])
def test_rmat22_dropins(torch_dev, oracle, rmat, rmat_ref, entry, monkeypatch):
    """The reference's entry points on host arrays (each call uploads the
    matrix), 4 chunks so split rows go through the device merge."""
    from hpda_sblas_pnnl_amd import gen as G
    from hpda_sblas_pnnl_amd import spmv as S
    monkeypatch.setenv("SBLAS_CHUNKS", "4")
    A, x, y0 = rmat
    y = y0.copy()
    if entry.startswith("v1_k"):
        r = S.spMV_mgpu_v1(A.m, A.n, A.nnz, 1.0, A.val, A.row_ptr, A.col, x, 0.5, y, 1,
                           int(entry[-1]))
    elif entry == "baseline_coo":
        r = S.spMV_mgpu_baseline_coo(A.m, A.n, A.nnz, 1.0, A.val, G.coo_rows(A), A.col, x, 0.5,
                                     y, 1)
    elif entry == "v1_numa_coo":
        r = S.spMV_mgpu_v1_numa_coo(A.m, A.n, A.nnz, 1.0, A.val, G.coo_rows(A), A.col, x, 0.5,
                                    y, 1, 1, [0], 1, 1)
    else:
        cp, rows, cv = G.to_csc(A)
        r = S.spMV_mgpu_v1_numa_csc(A.m, A.n, A.nnz, 1.0, cv, cp, rows, x, 0.5, y, 1, 1, [0], 0,
                                    0)
    assert np.isfinite(r.comp_time), entry
    assert oracle.err_normwise(y, rmat_ref) <= TOL, entry


def _narrow_check(oracle, plan, A, vr, x, y0, got, ref):
    """x gathered or staged from an FP32 copy adds x's rounding: every row
    within 2^-22 of sum |a_ij x_j|; else within 1e-12 of the rounded-value
    oracle."""
    if plan.info()["narrow_x"]:
        assert oracle.err_maxrow(A.m, vr, A.row_ptr, A.col, x, 1.0, 0.5, y0, got, ref) <= 2.0**-22
        assert oracle.err_normwise(got, ref) <= TOL_FP32_VS_FP64
    else:
        assert oracle.err_normwise(got, ref) <= TOL


@pytest.mark.parametrize("nx", [0, 1])
@pytest.mark.parametrize("vt", ["f32", "bf16"])
def test_rmat22_narrow_values(torch_dev, oracle, rmat, rmat_ref, vt, nx):
    """Binary values are exact in FP32 and BF16, so the rounded-value oracles
    equal the FP64 one; with x in FP64 (narrow_x 1) the narrow forms meet it
    at the FP64 bar, with x from an FP32 copy (the default) within x's
    rounding."""
    from hpda_sblas_pnnl_amd import gen as G
    from hpda_sblas_pnnl_amd.device import SpmvPlan
    torch = torch_dev
    A, x, y0 = rmat
    if vt == "f32":
        vals = A.val.astype(np.float32)
        vr = vals.astype(np.float64)
        ref = oracle.csr_spmv_f32val(A.m, vals, A.row_ptr, A.col, x, 1.0, 0.5, y0)
    else:
        bits = G.to_bf16_bits(A.val)
        vals = bits.view(np.int16)
        vr = G.from_bf16_bits(bits)
        ref = oracle.csr_spmv_bf16val(A.m, bits, A.row_ptr, A.col, x, 1.0, 0.5, y0)
    assert np.array_equal(ref, rmat_ref)
    rp, col, dv = _dev(torch, A.row_ptr, A.col, vals)
    plan = SpmvPlan("csr", A.m, A.n, rp, col, dv, narrow_x=nx)
    got = _plan_run(torch, plan, x, y0, 1.0, 0.5)
    if nx == 1:
        assert plan.info()["narrow_x"] == 0
    _narrow_check(oracle, plan, A, vr, x, y0, got, ref)


def test_rmat22_spmm_k8(torch_dev, oracle, rmat):
    from hpda_sblas_pnnl_amd import gen as G
    from hpda_sblas_pnnl_amd.device import SpmvPlan
    torch = torch_dev
    A, _, _ = rmat
    K = 8
    X = G.values(A.n * K, 0, 4, "uniform").reshape(A.n, K)
    Y0 = G.values(A.m * K, 0, 5, "uniform").reshape(A.m, K)
    ref = oracle.csr_spmm(A.m, K, A.val, A.row_ptr, A.col, X, 1.0, 0.5, Y0).reshape(A.m, K)
    rp, col, val, dX, dY = _dev(torch, A.row_ptr, A.col, A.val, X, Y0)
    plan = SpmvPlan("csr", A.m, A.n, rp, col, val)
    plan.spmm(dX, dY, 1.0, 0.5)
    torch.cuda.synchronize()
    got = dY.cpu().numpy()
    assert oracle.err_normwise(got.ravel(), ref.ravel()) <= TOL
    # column j of the block equals one SpMV with x = X[:, j]
    y1 = oracle.csr_spmv(A.m, A.n, A.val, A.row_ptr, A.col, X[:, 3].copy(), 1.0, 0.5,
                         Y0[:, 3].copy(), threads=_threads())
    assert oracle.err_normwise(got[:, 3], y1) <= TOL


# ---------------------------------------------------------------------------
# power-law n = 20,000 with narrow values
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("vt", ["f32", "bf16"])
def test_powerlaw_narrow_values(torch_dev, oracle, vt):
    from hpda_sblas_pnnl_amd import gen as G
    from hpda_sblas_pnnl_amd.device import SpmvPlan
    torch = torch_dev
    A = G.powerlaw_csr(20000, seed=0)
    x = G.values(A.n, 0, 2, "uniform")
    y0 = G.values(A.m, 0, 3, "uniform")
    ref64 = oracle.csr_spmv(A.m, A.n, A.val, A.row_ptr, A.col, x, 1.0, 0.5, y0)
    if vt == "f32":
        vals = A.val.astype(np.float32)
        vr = vals.astype(np.float64)
        ref = oracle.csr_spmv_f32val(A.m, vals, A.row_ptr, A.col, x, 1.0, 0.5, y0)
    else:
        bits = G.to_bf16_bits(A.val)
        vals = bits.view(np.int16)
        vr = G.from_bf16_bits(bits)
        ref = oracle.csr_spmv_bf16val(A.m, bits, A.row_ptr, A.col, x, 1.0, 0.5, y0)
    rp, col, dv = _dev(torch, A.row_ptr, A.col, vals)
    plan = SpmvPlan("csr", A.m, A.n, rp, col, dv)
    got = _plan_run(torch, plan, x, y0, 1.0, 0.5)
    _narrow_check(oracle, plan, A, vr, x, y0, got, ref)
    if vt == "f32":
        assert oracle.err_normwise(got, ref64) <= TOL_FP32_VS_FP64


# ---------------------------------------------------------------------------
# banded n = 2e6, 8 chunks
# ---------------------------------------------------------------------------

def test_banded2m_csr_v1_numa_8_chunks(torch_dev, oracle, banded, banded_ref):
    """The benchmark's plan: 8 chunks by the v1_numa rule, split rows merged
    on the device."""
    from hpda_sblas_pnnl_amd.device import SpmvPlan
    torch = torch_dev
    A, x, y0 = banded
    rp, col, val = _dev(torch, A.row_ptr, A.col, A.val)
    plan = SpmvPlan("csr", A.m, A.n, rp, col, val, chunks=8, part_mode="v1_numa")
    info = plan.info()
    assert info["chunks"] == 8 and info["split_rows"] > 0
    got = _plan_run(torch, plan, x, y0, 1.0, 0.5)
    assert oracle.err_normwise(got, banded_ref) <= TOL


@pytest.mark.parametrize("entry", ["v1_numa_coo", "v1_numa_csc", "baseline_coo", "baseline_csc"])
def test_banded2m_coo_csc_dropins_8_chunks(torch_dev, oracle, banded, banded_ref, entry,
                                           monkeypatch):
    from hpda_sblas_pnnl_amd import gen as G
    from hpda_sblas_pnnl_amd import spmv as S
    monkeypatch.setenv("SBLAS_CHUNKS", "8")
    A, x, y0 = banded
    y = y0.copy()
    if entry.endswith("coo"):
        ri = G.coo_rows(A)
        fn = S.spMV_mgpu_v1_numa_coo if entry.startswith("v1") else S.spMV_mgpu_baseline_coo
        if entry.startswith("v1"):
            r = fn(A.m, A.n, A.nnz, 1.0, A.val, ri, A.col, x, 0.5, y, 1, 1, [0], 1, 1)
        else:
            r = fn(A.m, A.n, A.nnz, 1.0, A.val, ri, A.col, x, 0.5, y, 1)
    else:
        cp, rows, cv = G.to_csc(A)
        if entry.startswith("v1"):
            r = S.spMV_mgpu_v1_numa_csc(A.m, A.n, A.nnz, 1.0, cv, cp, rows, x, 0.5, y, 1, 1, [0],
                                        0, 0)
        else:
            r = S.spMV_mgpu_baseline_csc(A.m, A.n, A.nnz, 1.0, cv, cp, rows, x, 0.5, y, 1)
    assert np.isfinite(r.comp_time), entry
    assert oracle.err_normwise(y, banded_ref) <= TOL, entry


def test_banded2m_coo_csc_plans(torch_dev, oracle, banded, banded_ref):
    from hpda_sblas_pnnl_amd import gen as G
    from hpda_sblas_pnnl_amd.device import SpmvPlan
    torch = torch_dev
    A, x, y0 = banded
    ri, col, val = _dev(torch, G.coo_rows(A), A.col, A.val)
    got = _plan_run(torch, SpmvPlan("coo", A.m, A.n, ri, col, val), x, y0, 1.0, 0.5)
    assert oracle.err_normwise(got, banded_ref) <= TOL
    del ri, col, val
    cp, rows, cv = G.to_csc(A)
    dcp, drows, dcv = _dev(torch, cp, rows, cv)
    got = _plan_run(torch, SpmvPlan("csc", A.m, A.n, dcp, drows, dcv, chunks=8), x, y0, 1.0, 0.5)
    assert oracle.err_normwise(got, banded_ref) <= TOL


# ---------------------------------------------------------------------------
# 64-matrix batch, FP32 values, one hipGraph
# ---------------------------------------------------------------------------

@pytest.mark.timeout(600)
def test_batch64_fp32_graph(torch_dev, oracle):
    from hpda_sblas_pnnl_amd import gen as G
    from hpda_sblas_pnnl_amd.device import SpmvBatch
    torch = torch_dev
    mats = G.batched_workload(64, seed=0)
    items, refs32, refs64, y0s = [], [], [], []
    for i, (kind, A) in enumerate(mats):
        x = G.values(A.n, i, 2, "uniform")
        y0 = G.values(A.m, i, 3, "uniform")
        beta = 0.5 if i % 3 == 0 else 0.0
        v32 = A.val.astype(np.float32)
        refs32.append(oracle.csr_spmv_f32val(A.m, v32, A.row_ptr, A.col, x, 1.0, beta, y0))
        refs64.append(oracle.csr_spmv(A.m, A.n, A.val, A.row_ptr, A.col, x, 1.0, beta, y0,
                                      threads=_threads()))
        rp, col, dv, dx, dy = _dev(torch, A.row_ptr, A.col, v32, x, y0)
        items.append(dict(m=A.m, n=A.n, row_ptr=rp, col=col, val=dv, x=dx, y=dy, alpha=1.0,
                          beta=beta))
        y0s.append(torch.from_numpy(y0).cuda())
    batch = SpmvBatch(items)
    assert sum(A.nnz for _, A in mats) > 1_000_000_000
    kinds = {batch.choice(i)["variant"] for i in range(64)}
    assert kinds == {2, 3}  # both row blocks and column panels in the launch
    outs = []
    for rep in range(3):
        for it, y0 in zip(items, y0s):
            it["y"].copy_(y0)
        torch.cuda.synchronize()
        batch.graph_launch()
        torch.cuda.synchronize()
        outs.append([it["y"].cpu().numpy() for it in items])
    for i, (kind, A) in enumerate(mats):
        if batch.choice(i)["variant"] == 3 or A.nnz >= 2**20:
            # column panels stage x in FP32, row blocks of 2^20 elements or
            # more gather it from an FP32 copy: each row within two FP32
            # roundings per product of the FP32-value oracle
            vr = A.val.astype(np.float32).astype(np.float64)
            x = G.values(A.n, i, 2, "uniform")
            beta = 0.5 if i % 3 == 0 else 0.0
            assert oracle.err_maxrow(A.m, vr, A.row_ptr, A.col, x, 1.0, beta,
                                     y0s[i].cpu().numpy(), outs[0][i], refs32[i]) <= 2.0**-22, i
        else:
            assert oracle.err_normwise(outs[0][i], refs32[i]) <= TOL, i
        assert oracle.err_normwise(outs[0][i], refs64[i]) <= TOL_FP32_VS_FP64, i
        assert np.array_equal(outs[0][i], outs[1][i]) and np.array_equal(outs[0][i], outs[2][i])
