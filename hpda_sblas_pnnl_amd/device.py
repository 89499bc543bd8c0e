"""Plans over device-resident matrices (torch tensors on a HIP device).

``SpmvPlan`` wraps the C plan API (include/sblas.h, section 3): the analogue
of the reference's CSR5 handle (include/anonymouslib_cuda.h:11-53) -- analyse
once, then ``spmv``/``spmm`` on the current torch stream as often as needed.
torch supplies device memory and streams only; every kernel is the engine's.
"""
from __future__ import annotations

import ctypes

from . import _lib
from ._lib import Opts, PlanInfo, check, lib

_VAL_TYPES = {"f64": _lib.VAL_F64, "f32": _lib.VAL_F32, "bf16": _lib.VAL_BF16}
_KERNELS = {"auto": _lib.KERNEL_AUTO, "vector": _lib.KERNEL_VECTOR, "merge": _lib.KERNEL_MERGE,
            "csr5": _lib.KERNEL_CSR5, "coo": 4, 0: 0, 1: 1, 2: 2, 3: 3, 4: 4}
_PARTS = {"v1": _lib.PART_V1, "v1_numa": _lib.PART_V1_NUMA, "v2": _lib.PART_V2,
          "rowblock": _lib.PART_ROWBLOCK}


def _torch():
    import torch  # noqa: PLC0415  (torch is plumbing; import lazily)
    return torch


def val_type_of(t) -> int:
    torch = _torch()
    if t.dtype == torch.float64:
        return _lib.VAL_F64
    if t.dtype == torch.float32:
        return _lib.VAL_F32
    if t.dtype in (torch.bfloat16, torch.int16, torch.uint16):
        return _lib.VAL_BF16
    raise TypeError(f"unsupported value dtype {t.dtype}")


def make_opts(kernel="auto", chunks=0, part_mode="v1", nb=0, streams=1, subwave=0, sigma=0,
              xcd_remap=-1, variant=0, items=0, panel_rows=0, narrow_x=0,
              panel_layout=0) -> Opts:
    o = Opts()
    lib().sblas_opts_default(ctypes.byref(o))
    o.kernel = _KERNELS[kernel]
    o.chunks = chunks
    o.part_mode = _PARTS[part_mode] if isinstance(part_mode, str) else part_mode
    o.nb = nb
    o.streams = streams
    o.subwave = subwave
    o.sigma = sigma
    o.xcd_remap = xcd_remap
    o.variant = variant
    o.items = items
    o.panel_rows = panel_rows
    o.narrow_x = narrow_x
    o.panel_layout = panel_layout
    return o


def _stream_ptr(stream):
    torch = _torch()
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


class SpmvBatch:
    """Many CSR matrices run as one persistent launch (optionally a replayed
    hipGraph).  items: list of dicts with keys m, n, row_ptr, col, val, x, y,
    alpha, beta (tensors on one HIP device) and optionally k: x (n x k) and y
    (m x k) row-major hold k right-hand sides (k in 1, 2, 4, 8, 16).  variant: 0 lets the autotuner
    pick row blocks or column panels per matrix, 2 forces row blocks, 3 asks
    for column panels wherever the format allows them."""

    def __init__(self, items, subwave=0, chunks=0, item_size=0, variant=0, panel_rows=0,
                 narrow_x=0):
        self._keep = []
        arr = (_lib.BatchItem * max(len(items), 1))()
        for i, it in enumerate(items):
            self._keep.append(it)
            arr[i].m, arr[i].n = int(it["m"]), int(it["n"])
            arr[i].nnz = int(it["val"].numel())
            arr[i].d_row_ptr = it["row_ptr"].data_ptr()
            arr[i].d_col = it["col"].data_ptr()
            arr[i].d_val = it["val"].data_ptr()
            arr[i].val_type = val_type_of(it["val"])
            arr[i].k = int(it.get("k", 1))
            arr[i].d_x = it["x"].data_ptr()
            arr[i].d_y = it["y"].data_ptr()
            arr[i].alpha = float(it.get("alpha", 1.0))
            arr[i].beta = float(it.get("beta", 0.0))
        o = make_opts(subwave=subwave, chunks=chunks, items=item_size, variant=variant,
                      panel_rows=panel_rows, narrow_x=narrow_x)
        h = ctypes.c_void_p()
        check(lib().sblas_batch_create(len(items), arr, ctypes.byref(o), ctypes.byref(h)),
              "sblas_batch_create")
        self._h = h
        self.count = len(items)

    def run(self, stream=None):
        check(lib().sblas_batch_run(self._h, _stream_ptr(stream)), "sblas_batch_run")

    def graph_launch(self, stream=None):
        check(lib().sblas_batch_graph_launch(self._h, _stream_ptr(stream)),
              "sblas_batch_graph_launch")

    def choice(self, i) -> dict:
        c = _lib.BatchChoice()
        check(lib().sblas_batch_get_choice(self._h, i, ctypes.byref(c)), "sblas_batch_get_choice")
        return c.as_dict()

    def shape(self) -> dict:
        w, g, n = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(lib().sblas_batch_get_shape(self._h, ctypes.byref(w), ctypes.byref(g),
                                          ctypes.byref(n)), "sblas_batch_get_shape")
        return {"work_items": w.value, "grid": g.value, "launches": n.value}

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().sblas_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001  (interpreter shutdown)
            pass


class SpmvPlan:
    """A matrix analysed for repeated y = alpha*A*x + beta*y.

    fmt 'csr': (row_ptr, col, val); 'coo': (row_idx, col_idx, val) sorted by
    row; 'csc': (col_ptr, row_idx, val).  Index tensors int32, values float64 /
    float32 / bfloat16, all on the same HIP device.  The tensors must stay
    alive and unchanged while the plan exists (the plan keeps references).
    """

    def __init__(self, fmt, m, n, a0, a1, val, **opts):
        self.fmt, self.m, self.n = fmt, int(m), int(n)
        self._keep = (a0, a1, val)
        self.nnz = int(val.numel())
        o = make_opts(**opts)
        h = ctypes.c_void_p()
        create = {"csr": lib().sblas_plan_create_csr, "coo": lib().sblas_plan_create_coo,
                  "csc": lib().sblas_plan_create_csc}[fmt]
        st = create(self.m, self.n, self.nnz, a0.data_ptr(), a1.data_ptr(), val.data_ptr(),
                    val_type_of(val), ctypes.byref(o), ctypes.byref(h))
        check(st, f"sblas_plan_create_{fmt}")
        self._h = h

    def spmv(self, x, y, alpha=1.0, beta=0.0, stream=None):
        st = lib().sblas_plan_spmv(self._h, float(alpha), x.data_ptr(), float(beta),
                                   y.data_ptr(), _stream_ptr(stream))
        check(st, "sblas_plan_spmv")
        return y

    def spmm(self, X, Y, alpha=1.0, beta=0.0, stream=None):
        k = 1 if X.dim() == 1 else int(X.shape[1])
        st = lib().sblas_plan_spmm(self._h, k, float(alpha), X.data_ptr(), float(beta),
                                   Y.data_ptr(), _stream_ptr(stream))
        check(st, "sblas_plan_spmm")
        return Y

    def info(self) -> dict:
        inf = PlanInfo()
        check(lib().sblas_plan_get_info(self._h, ctypes.byref(inf)), "sblas_plan_get_info")
        return inf.as_dict()

    def chunks(self):
        import numpy as np  # noqa: PLC0415
        n = lib().sblas_plan_get_chunks(self._h, None, 0)
        if n < 0:
            return -n, np.zeros(0, dtype=_lib.CHUNK_DTYPE)
        out = np.zeros(max(n, 1), dtype=_lib.CHUNK_DTYPE)
        lib().sblas_plan_get_chunks(self._h, out.ctypes.data, n)
        return 0, out[:n]

    def set_timing(self, on: bool = True):
        check(lib().sblas_plan_set_timing(self._h, int(bool(on))), "sblas_plan_set_timing")

    def kernel_times_ms(self, cap: int = 4096):
        import numpy as np  # noqa: PLC0415
        out = np.zeros(cap, dtype=np.float32)
        n = lib().sblas_plan_kernel_times(self._h, out.ctypes.data, cap)
        if n < 0:
            raise _lib.SblasError(-n, "sblas_plan_kernel_times")
        return out[:n].astype(np.float64)

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().sblas_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # noqa: BLE001  (interpreter shutdown)
            pass


def _int32_dev(t, name):
    torch = _torch()
    if not (t.is_cuda and t.dtype == torch.int32 and t.is_contiguous()):
        raise TypeError(f"{name} must be a contiguous int32 device tensor")
    return t


def coo_sort(row, col, val, m, n, by_col=False, stream=None):
    """Sorts device COO tensors in place by (row, column), or by (column, row)
    with ``by_col`` -- cusparseXcoosortByRow / ByColumn with the value gather
    (src/spmv_helper.cu:406-476).  Stable: equal pairs keep their order.
    Returns when the work on ``stream`` is done."""
    _int32_dev(row, "row"), _int32_dev(col, "col")
    if not (val.is_cuda and val.is_contiguous()):
        raise TypeError("val must be a contiguous device tensor")
    nnz = row.numel()
    if col.numel() != nnz or val.numel() != nnz:
        raise ValueError("row, col and val must have the same length")
    check(lib().sblas_coo_sort(m, n, nnz, row.data_ptr(), col.data_ptr(), val.data_ptr(),
                               val_type_of(val), 1 if by_col else 0, _stream_ptr(stream)),
          "sblas_coo_sort")


def coo_ptr(sorted_idx, count, stream=None):
    """Compressed pointers of a sorted index tensor (cusparseXcoo2csr,
    src/spmv_helper.cu:478-491): an int32 tensor of count + 1 entries."""
    torch = _torch()
    _int32_dev(sorted_idx, "sorted_idx")
    ptr = torch.empty(count + 1, dtype=torch.int32, device=sorted_idx.device)
    check(lib().sblas_coo_ptr(count, sorted_idx.numel(), sorted_idx.data_ptr(), ptr.data_ptr(),
                              _stream_ptr(stream)), "sblas_coo_ptr")
    return ptr


def csr_drop_zero_x(row_ptr, col, val, x, m, n, stream=None):
    """spMspV's filter on device tensors (src/dspmv_mgpu_v1.cu:19-74): returns
    (row_ptr, col, val) of the CSR matrix without the elements whose
    x[col] == 0, rows in order."""
    torch = _torch()
    _int32_dev(row_ptr, "row_ptr"), _int32_dev(col, "col")
    if val.dtype != torch.float64 or x.dtype != torch.float64:
        raise TypeError("val and x must be float64 device tensors")
    nnz = col.numel()
    rp2 = torch.empty(m + 1, dtype=torch.int32, device=col.device)
    c2 = torch.empty(max(nnz, 1), dtype=torch.int32, device=col.device)
    v2 = torch.empty(max(nnz, 1), dtype=torch.float64, device=col.device)
    kept = ctypes.c_int64()
    check(lib().sblas_csr_drop_zero_x(m, n, nnz, row_ptr.data_ptr(), col.data_ptr(),
                                      val.data_ptr(), x.data_ptr(), rp2.data_ptr(),
                                      c2.data_ptr(), v2.data_ptr(), ctypes.byref(kept),
                                      _stream_ptr(stream)), "sblas_csr_drop_zero_x")
    k = kept.value
    return rp2, c2[:k], v2[:k]


def stream_spmv_csr(m, n, row_ptr, col, val, alpha, x, beta, y, nb, q=2, stream=None):
    """Out-of-core CSR SpMV in the reference's v2 task structure
    (src/dspmv_mgpu_v2.cu:103-544): the matrix (numpy arrays in host memory)
    streams through the device in tasks of nb elements over q streams;
    x and y are float64 device tensors, y updated in place."""
    import numpy as np
    rp = np.ascontiguousarray(row_ptr, dtype=np.int32)
    ci = np.ascontiguousarray(col, dtype=np.int32)
    va = np.ascontiguousarray(val, dtype=np.float64)
    torch = _torch()
    if x.dtype != torch.float64 or y.dtype != torch.float64 or not (x.is_cuda and y.is_cuda):
        raise TypeError("x and y must be float64 device tensors")
    check(lib().sblas_stream_spmv_csr(m, n, len(va), rp.ctypes.data, ci.ctypes.data,
                                      va.ctypes.data, float(alpha), x.data_ptr(), float(beta),
                                      y.data_ptr(), int(nb), int(q), _stream_ptr(stream)),
          "sblas_stream_spmv_csr")
