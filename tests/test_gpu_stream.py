"""Out-of-core v2 streaming (src/dspmv_mgpu_v2.cu:103-544 task structure):
the matrix stays in host memory and streams through the device in nb-element
tasks over q streams.  Checked against the oracle, and bit for bit against
the in-core v2 drop-in at the same tasks (same kernel, same per-row order,
same ordered merge)."""
import os

import numpy as np
import pytest

from conftest import random_csr

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _dev(*arrs):
    import torch
    return [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in arrs]


@pytest.mark.parametrize("nb,q", [(997, 1), (997, 3), (5000, 2), (200_000, 4), (37, 8)])
def test_stream_matches_oracle(oracle, nb, q):
    """Tasks far shorter than rows (a row over many tasks), rows ending on
    task edges, empty rows, beta != 0."""
    import torch
    from hpda_sblas_pnnl_amd.device import stream_spmv_csr
    rng = np.random.default_rng(nb + q)
    m, n = 6000, 9000
    rp, col, val = random_csr(rng, m, n, 15, empty_frac=0.3, long_rows=3, long_len=4000)
    x = rng.uniform(-1, 1, n)
    y0 = rng.uniform(-1, 1, m)
    dx, dy = _dev(x, y0)
    stream_spmv_csr(m, n, rp, col, val, 0.75, dx, -1.25, dy, nb, q)
    torch.cuda.synchronize()
    ref = oracle.csr_spmv(m, n, val, rp, col, x, 0.75, -1.25, y0)
    assert oracle.err_normwise(dy.cpu().numpy(), ref) <= TOL


def test_stream_dropin_vs_in_core(oracle, monkeypatch):
    """spMV_mgpu_v2 streamed (SBLAS_V2_STREAM=1) against the in-core v2 drop-in
    (SBLAS_V2_STREAM=0).  With at most 64 tasks the in-core form also runs one
    sub-wave launch per task, so the two agree bit for bit; with more it runs
    fused in the row-block form (another summation order): both within 1e-12
    of the oracle."""
    from hpda_sblas_pnnl_amd import spmv as S
    from hpda_sblas_pnnl_amd._lib import lib
    rng = np.random.default_rng(3)
    m, n = 20_000, 15_000
    rp, col, val = random_csr(rng, m, n, 30, empty_frac=0.1, long_rows=2, long_len=9000)
    x = rng.uniform(-1, 1, n)
    y0 = rng.uniform(-1, 1, m)
    ref = oracle.csr_spmv(m, n, val, rp, col, x, 1.5, 0.5, y0)
    for nb, q in ((40_000, 2), (3_000, 4)):
        out = []
        for mode in ("1", "0"):
            monkeypatch.setenv("SBLAS_V2_STREAM", mode)
            y = y0.copy()
            S.spMV_mgpu_v2(m, n, len(val), 1.5, val, rp, col, x, 0.5, y, 1, 1, nb, q)
            out.append(y)
            assert oracle.err_normwise(y, ref) <= TOL, (nb, q, mode)
        if lib().sblas_v2_num_tasks(len(val), nb, q) <= 64:
            assert np.array_equal(out[0], out[1]), (nb, q)


def test_stream_edge_cases(oracle):
    import torch
    from hpda_sblas_pnnl_amd._lib import SblasError
    from hpda_sblas_pnnl_amd.device import stream_spmv_csr
    # no elements: y = beta * y
    dx, dy = _dev(np.ones(3), np.array([1.0, -2.0, 4.0]))
    stream_spmv_csr(3, 3, np.zeros(4, np.int32), np.zeros(0, np.int32), np.zeros(0), 2.0, dx,
                    0.5, dy, 10, 2)
    assert dy.cpu().numpy().tolist() == [0.5, -1.0, 2.0]
    # an out-of-range column: an error, and no read outside x
    rp = np.array([0, 2, 3], np.int32)
    col = np.array([0, 5, 1], np.int32)
    dx, dy = _dev(np.ones(2), np.zeros(2))
    with pytest.raises(SblasError) as e:
        stream_spmv_csr(2, 2, rp, col, np.ones(3), 1.0, dx, 0.0, dy, 2, 2)
    assert e.value.status == 1
    # decreasing row pointers are refused before any device work
    with pytest.raises(SblasError):
        stream_spmv_csr(2, 2, np.array([0, 3, 2], np.int32), np.zeros(3, np.int32), np.ones(3),
                        1.0, dx, 0.0, dy, 2, 2)
    torch.cuda.synchronize()


def test_stream_banded_full_size(oracle):
    """The 2M-row, 2.8e8-element banded matrix streamed from host memory in
    nine tasks over four streams matches the oracle (1e-12 normwise) and equals
    the in-core v2 plan at the same tasks bit for bit."""
    import torch
    from hpda_sblas_pnnl_amd import gen as G
    from hpda_sblas_pnnl_amd.device import SpmvPlan, stream_spmv_csr
    A = G.banded_csr(2_000_000, seed=0)
    x = G.values(A.n, 0, 2, "uniform")
    nb = (A.nnz + 8) // 9
    dx = torch.from_numpy(x).cuda()
    dy = torch.zeros(A.m, dtype=torch.float64, device="cuda")
    stream_spmv_csr(A.m, A.n, A.row_ptr, A.col, A.val, 1.0, dx, 0.0, dy, nb, 4)
    got = dy.cpu().numpy()
    ref = oracle.csr_spmv(A.m, A.n, A.val, A.row_ptr, A.col, x, 1.0, 0.0, np.zeros(A.m),
                          threads=max(1, min(16, len(os.sched_getaffinity(0)))))
    assert oracle.err_normwise(got, ref) <= 1e-12
    rp, col, val = _dev(A.row_ptr, A.col, A.val)
    plan = SpmvPlan("csr", A.m, A.n, rp, col, val, kernel="vector", variant=1, part_mode="v2",
                    nb=nb, streams=4)
    dy.zero_()
    plan.spmv(dx, dy, 1.0, 0.0)
    torch.cuda.synchronize()
    assert np.array_equal(got, dy.cpu().numpy())
