"""Plans under hipGraph capture: every single-stream form of a plan
(row blocks with their merge, narrow-x row blocks with the FP32 copy of x,
column panels in both layouts with chunk merges, column-sorted row blocks,
row blocks over column segments, merge path, CSR5-style tiles, the COO
kernel, and a multi-vector product) launches only kernels on the caller's
stream, so a torch CUDA graph captures it.  Each graph is replayed three
times with eager work of the program between replays, y reset before each,
and every replay equals the eager product bit for bit."""
import numpy as np
import pytest

from conftest import random_csr

pytestmark = pytest.mark.gpu


def _dev(*arrays):
    import torch
    return [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in arrays]


def _banded():
    from hpda_sblas_pnnl_amd import gen as G
    return G.banded_csr(120_000, seed=21, halfwidth=12000)


def _random(m, n, mean, seed):
    rng = np.random.default_rng(seed)
    rp, col, val = random_csr(rng, m, n, mean, empty_frac=0.1, long_rows=2, long_len=min(n, 6000))
    return rp, col, val


def _cases():
    """(name, builder) pairs; a builder returns (plan, x, y0, k, alpha, beta);
    the plan keeps its device arrays alive in plan._keep."""
    from hpda_sblas_pnnl_amd.device import SpmvPlan

    def csr(rp, col, vals, m, n, **opts):
        drp, dcol, dval = _dev(rp, col, vals)
        plan = SpmvPlan("csr", m, n, drp, dcol, dval, **opts)
        plan._keep = (drp, dcol, dval)
        return plan

    def vecs(m, n, seed, k=1):
        rng = np.random.default_rng(seed)
        shape_x, shape_y = ((n,), (m,)) if k == 1 else ((n, k), (m, k))
        return rng.uniform(-1, 1, shape_x), rng.uniform(-1, 1, shape_y)

    def row_blocks():
        m, n = 20000, 15000
        rp, col, val = _random(m, n, 12, 1)
        return (csr(rp, col, val, m, n, kernel="vector", variant=2),) + vecs(m, n, 1) + (1, 1.5, -0.5)

    def row_blocks_narrow_x():
        m = n = 65536
        rp, col, val = _random(m, n, 20, 2)
        return (csr(rp, col, val.astype(np.float32), m, n, kernel="vector", variant=2),) + \
            vecs(m, n, 2) + (1, 1.0, 0.5)

    def panels_row_order():
        A = _banded()
        return (csr(A.row_ptr, A.col, A.val, A.m, A.n, variant=3, chunks=8, part_mode="v1_numa",
                    panel_rows=640, panel_layout=1),) + vecs(A.m, A.n, 3) + (1, -0.75, 1.25)

    def panels_sorted_fp32():
        A = _banded()
        return (csr(A.row_ptr, A.col, A.val.astype(np.float32), A.m, A.n, variant=3, chunks=3,
                    panel_layout=2),) + vecs(A.m, A.n, 4) + (1, 1.0, 0.5)

    def column_sorted_blocks():
        m, n = 20000, 15000
        rp, col, val = _random(m, n, 12, 5)
        return (csr(rp, col, val, m, n, kernel="vector", variant=4),) + vecs(m, n, 5) + (1, 1.0, 0.5)

    def segments():
        m, n = 20000, 15001
        rp, col, val = _random(m, n, 12, 6)
        return (csr(rp, col, val, m, n, kernel="vector", variant=6),) + vecs(m, n, 6) + (1, 1.25, -0.5)

    def merge_path():
        m, n = 20000, 15000
        rp, col, val = _random(m, n, 12, 7)
        return (csr(rp, col, val, m, n, kernel="merge"),) + vecs(m, n, 7) + (1, 1.0, 0.5)

    def csr5():
        m, n = 20000, 15000
        rp, col, val = _random(m, n, 12, 8)
        return (csr(rp, col, val, m, n, kernel="csr5", sigma=16),) + vecs(m, n, 8) + (1, 1.0, 0.5)

    def coo():
        m, n = 20000, 15000
        rp, col, val = _random(m, n, 12, 9)
        rows = np.repeat(np.arange(m, dtype=np.int32), np.diff(rp))
        dri, dcol, dval = _dev(rows, col, val)
        plan = SpmvPlan("coo", m, n, dri, dcol, dval, kernel="coo")
        plan._keep = (dri, dcol, dval)
        return (plan,) + vecs(m, n, 9) + (1, 1.0, 0.5)

    def multi_vector():
        m, n = 20000, 15000
        rp, col, val = _random(m, n, 12, 10)
        return (csr(rp, col, val, m, n, kernel="vector"),) + vecs(m, n, 10, k=4) + (4, 1.0, 0.5)

    return [("row_blocks", row_blocks), ("row_blocks_narrow_x", row_blocks_narrow_x),
            ("panels_row_order", panels_row_order), ("panels_sorted_fp32", panels_sorted_fp32),
            ("column_sorted_blocks", column_sorted_blocks), ("segments", segments),
            ("merge_path", merge_path), ("csr5", csr5), ("coo", coo),
            ("multi_vector", multi_vector)]


_NAMES = [name for name, _ in _cases()]


@pytest.mark.parametrize("name", _NAMES)
def test_plan_graph_replays_equal_eager(name):
    import torch
    build = dict(_cases())[name]
    plan, x, y0, k, alpha, beta = build()
    print(f"[graph] {name} {plan.info()['kernel']}/{plan.info()['variant']}", flush=True)
    dx, dy0 = _dev(x, y0)

    def product(dy):
        if k == 1:
            plan.spmv(dx, dy, alpha, beta)
        else:
            plan.spmm(dx, dy, alpha, beta)

    eager = dy0.clone()
    product(eager)
    torch.cuda.synchronize()
    eager = eager.cpu().numpy()
    assert np.all(np.isfinite(eager))

    static_y = dy0.clone()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):  # one eager product on the capture's kind of stream first
        product(static_y)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        product(static_y)
    other = dy0.clone()
    for rep in range(3):
        static_y.copy_(dy0)
        g.replay()
        product(other)  # eager work of the program between replays
        torch.cuda.synchronize()
        assert np.array_equal(static_y.cpu().numpy(), eager), (name, rep)
