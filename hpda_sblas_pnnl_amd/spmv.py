"""The reference's C entry points (include/spmv_kernel.h:211-296), callable from
Python on host (numpy) arrays.

Same names and argument order as the reference.  ``y`` is updated in place
with y = alpha*A*x + beta*y, ``alpha``/``beta`` are Python floats (the
reference passes pointers), and the return value is the reference's
``spmv_ret`` timing record.  Where the reference exits the process on an error
(``checkCudaErrors``, include/detail/cuda/helper_cuda.h:583-595) these raise
``SblasError``; the C ABI itself returns NaN timings and a status.
"""
from __future__ import annotations

import ctypes
import math
import os

import numpy as np

from . import _lib
from ._lib import SblasError, lib

__all__ = [
    "spMV_mgpu_baseline", "spMV_mgpu_baseline_csc", "spMV_mgpu_baseline_coo", "spMV_mgpu_v1",
    "spMV_mgpu_v1_numa", "spMV_mgpu_v1_numa_csc", "spMV_mgpu_v1_numa_coo", "spMV_mgpu_v2",
    "spMspV_mgpu_v1", "spMspV_mgpu_v1_numa", "spMspV_mgpu_v2", "get_row_from_index",
    "findFirstInSorted", "findLastInSorted", "sortCOORow", "sortCOOCol", "coo2csr", "coo2csc",
    "mm_read_coo",
]


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _scal(v):
    return ctypes.byref(ctypes.c_double(float(v)))


def _dropin(name, m, n, nnz, alpha, val, a0, a1, x, beta, y, *tail, mapping=None):
    """Calls entry point `name` with the reference's argument order: sizes,
    &alpha, the three matrix arrays, x, &beta, y, then the entry's own
    trailing integers (ngpu, kernel, ...); a numa_mapping array, when given,
    goes right after ngpu and kernel.  The converted arrays stay referenced
    here until the call returns (ctypes sees only their addresses)."""
    if not (isinstance(y, np.ndarray) and y.dtype == np.float64 and y.flags.c_contiguous):
        raise TypeError("y must be a C-contiguous float64 numpy array (updated in place)")
    val, a0, a1, x = _f64(val), _i32(a0), _i32(a1), _f64(x)
    tail = list(tail)
    if mapping is not None:
        mp = _i32(mapping)
        tail.insert(2, mp.ctypes.data)
    ret = getattr(lib(), name)(m, n, nnz, _scal(alpha), val.ctypes.data, a0.ctypes.data,
                               a1.ctypes.data, x.ctypes.data, _scal(beta), y.ctypes.data, *tail)
    if any(math.isnan(getattr(ret, f)) for f, _ in _lib.SpmvRet._fields_):
        raise SblasError(lib().sblas_last_error(), name)
    return ret


def spMV_mgpu_baseline(m, n, nnz, alpha, csrVal, csrRowPtr, csrColIndex, x, beta, y, ngpu=1):
    """Row-block CSR SpMV (src/dspmv_mgpu_baseline.cu:14-202)."""
    return _dropin("spMV_mgpu_baseline", m, n, nnz, alpha, csrVal, csrRowPtr, csrColIndex, x,
                   beta, y, ngpu)


def spMV_mgpu_baseline_csc(m, n, nnz, alpha, cscVal, cscColPtr, cscRowIdx, x, beta, y, ngpu=1):
    """Column-block CSC SpMV (src/dspmv_mgpu_baseline_csc.cu:14-239)."""
    return _dropin("spMV_mgpu_baseline_csc", m, n, nnz, alpha, cscVal, cscColPtr, cscRowIdx, x,
                   beta, y, ngpu)


def spMV_mgpu_baseline_coo(m, n, nnz, alpha, cooVal, cooRowIdx, cooColIdx, x, beta, y, ngpu=1):
    """Row-block COO SpMV (src/dspmv_mgpu_baseline_coo.cu:16-264)."""
    return _dropin("spMV_mgpu_baseline_coo", m, n, nnz, alpha, cooVal, cooRowIdx, cooColIdx, x,
                   beta, y, ngpu)


def spMV_mgpu_v1(m, n, nnz, alpha, csrVal, csrRowPtr, csrColIndex, x, beta, y, ngpu=1,
                 kernel=1):
    """nnz-chunked CSR SpMV with kernel switch 1/2/3 (src/dspmv_mgpu_v1.cu:76-360)."""
    return _dropin("spMV_mgpu_v1", m, n, nnz, alpha, csrVal, csrRowPtr, csrColIndex, x, beta, y,
                   ngpu, kernel)


def spMV_mgpu_v1_numa(m, n, nnz, alpha, csrVal, csrRowPtr, csrColIndex, x, beta, y, ngpu=1,
                      kernel=1, numa_mapping=None, part_opt=0, merg_opt=0):
    """Two-level nnz-chunked CSR SpMV (src/dspmv_mgpu_v1_numa.cu:98-837)."""
    return _dropin("spMV_mgpu_v1_numa", m, n, nnz, alpha, csrVal, csrRowPtr, csrColIndex, x,
                   beta, y, ngpu, kernel, part_opt, merg_opt,
                   mapping=numa_mapping if numa_mapping is not None else [0] * max(ngpu, 1))


def spMV_mgpu_v1_numa_csc(m, n, nnz, alpha, cscVal, cscColPtr, cscRowIndex, x, beta, y, ngpu=1,
                          kernel=1, numa_mapping=None, part_opt=0, merg_opt=0):
    """Two-level nnz-chunked CSC SpMV (src/dspmv_mgpu_v1_numa_csc.cu:43-483)."""
    return _dropin("spMV_mgpu_v1_numa_csc", m, n, nnz, alpha, cscVal, cscColPtr, cscRowIndex, x,
                   beta, y, ngpu, kernel, part_opt, merg_opt,
                   mapping=numa_mapping if numa_mapping is not None else [0] * max(ngpu, 1))


def spMV_mgpu_v1_numa_coo(m, n, nnz, alpha, cooVal, cooRowIdx, cooColIdx, x, beta, y, ngpu=1,
                          kernel=1, numa_mapping=None, part_opt=0, merg_opt=0):
    """Two-level nnz-chunked COO SpMV (src/dspmv_mgpu_v1_numa_coo.cu:35-633)."""
    return _dropin("spMV_mgpu_v1_numa_coo", m, n, nnz, alpha, cooVal, cooRowIdx, cooColIdx, x,
                   beta, y, ngpu, kernel, part_opt, merg_opt,
                   mapping=numa_mapping if numa_mapping is not None else [0] * max(ngpu, 1))


def spMV_mgpu_v2(m, n, nnz, alpha, csrVal, csrRowPtr, csrColIndex, x, beta, y, ngpu=1,
                 kernel=1, nb=1 << 20, copy_of_workspace=1):
    """Task-queue CSR SpMV, nb nonzeros per task (src/dspmv_mgpu_v2.cu:103-300)."""
    return _dropin("spMV_mgpu_v2", m, n, nnz, alpha, csrVal, csrRowPtr, csrColIndex, x, beta, y,
                   ngpu, kernel, nb, copy_of_workspace)


def spMspV_mgpu_v1(m, n, nnz, alpha, csrVal, csrRowPtr, csrColIndex, x, beta, y, ngpu=1,
                   kernel=1):
    """v1 after dropping entries with x[col] == 0 (src/dspmv_mgpu_v1.cu:19-74)."""
    return _dropin("spMspV_mgpu_v1", m, n, nnz, alpha, csrVal, csrRowPtr, csrColIndex, x, beta,
                   y, ngpu, kernel)


def spMspV_mgpu_v1_numa(m, n, nnz, alpha, csrVal, csrRowPtr, csrColIndex, x, beta, y, ngpu=1,
                        kernel=1):
    """Declared by the reference (include/spmv_kernel.h:282) but never defined:
    always raises."""
    return _dropin("spMspV_mgpu_v1_numa", m, n, nnz, alpha, csrVal, csrRowPtr, csrColIndex, x,
                   beta, y, ngpu, kernel)


def spMspV_mgpu_v2(m, n, nnz, alpha, csrVal, csrRowPtr, csrColIndex, x, beta, y, ngpu=1,
                   kernel=1, nb=1 << 20, copy_of_workspace=1):
    """v2 after dropping entries with x[col] == 0 (src/dspmv_mgpu_v2.cu:35-101)."""
    return _dropin("spMspV_mgpu_v2", m, n, nnz, alpha, csrVal, csrRowPtr, csrColIndex, x, beta,
                   y, ngpu, kernel, nb, copy_of_workspace)


def get_row_from_index(n, a, idx):
    a = _i32(a)
    return lib().get_row_from_index(n, a.ctypes.data, idx)


def findFirstInSorted(a, n, key):
    a = _i32(a)
    return lib().findFirstInSorted(a.ctypes.data, n, key)


def findLastInSorted(a, n, key):
    a = _i32(a)
    return lib().findLastInSorted(a.ctypes.data, n, key)


# --------------------------------------------------------------------------
# partition API
# --------------------------------------------------------------------------

def _desc(n):
    return np.zeros(max(n, 1), dtype=_lib.CHUNK_DTYPE)


def partition(mode: str, *, C: int, m: int = 0, n: int = 0, nnz: int = 0, row_ptr=None,
              row_idx=None, col_ptr=None, numa_mapping=None, nb: int = 0, q: int = 1):
    """Chunk descriptors of one of the reference's rules.

    mode: 'rowblock' | 'baseline_coo' | 'baseline_csc' | 'v1' | 'v1_numa' |
    'v1_numa_coo' | 'v1_numa_csc' | 'v2'.  Returns (status, chunks[, nodes]).
    """
    L = lib()
    mp = None if numa_mapping is None else _i32(numa_mapping)
    mpp = None if mp is None else mp.ctypes.data
    nodes = 1 if mp is None else int(mp.max()) + 1
    if mode == "rowblock":
        rp = _i32(row_ptr); out = _desc(C)
        return L.sblas_partition_rowblock(m, rp.ctypes.data, C, out.ctypes.data), out[:C]
    if mode == "baseline_coo":
        ri = _i32(row_idx); out = _desc(C)
        return L.sblas_partition_baseline_coo(m, len(ri), ri.ctypes.data, C, out.ctypes.data), out[:C]
    if mode == "baseline_csc":
        cp = _i32(col_ptr); out = _desc(C)
        return L.sblas_partition_baseline_csc(n, cp.ctypes.data, C, out.ctypes.data), out[:C]
    if mode == "v1":
        rp = _i32(row_ptr); out = _desc(C)
        return L.sblas_partition_v1(m, rp.ctypes.data, nnz, C, out.ctypes.data), out[:C]
    if mode == "v1_numa":
        rp = _i32(row_ptr); out, nout = _desc(C), _desc(nodes)
        st = L.sblas_partition_v1_numa(m, rp.ctypes.data, nnz, C, mpp, out.ctypes.data,
                                       nout.ctypes.data)
        return st, out[:C], nout[:nodes]
    if mode == "v1_numa_coo":
        ri = _i32(row_idx); out, nout = _desc(C), _desc(nodes)
        st = L.sblas_partition_v1_numa_coo(m, len(ri), ri.ctypes.data, C, mpp, out.ctypes.data,
                                           nout.ctypes.data)
        return st, out[:C], nout[:nodes]
    if mode == "v1_numa_csc":
        cp = _i32(col_ptr); out, nout = _desc(C), _desc(nodes)
        st = L.sblas_partition_v1_numa_csc(n, cp.ctypes.data, nnz, C, mpp, out.ctypes.data,
                                           nout.ctypes.data)
        return st, out[:C], nout[:nodes]
    if mode == "v2":
        rp = _i32(row_ptr)
        T = L.sblas_v2_num_tasks(nnz, nb, q)
        if T < 0:
            return -T, _desc(0)[:0]
        out = _desc(T)
        got = L.sblas_partition_v2(m, rp.ctypes.data, nnz, nb, q, out.ctypes.data, T)
        return (0 if got >= 0 else -got), out[:T]
    raise ValueError(f"unknown partition mode {mode!r}")


# --------------------------------------------------------------------------
# matrix ingestion (src/spmv_helper.cu:277-642, include/mmio.h)
# --------------------------------------------------------------------------

def _inplace(a, dtype, name):
    if not (isinstance(a, np.ndarray) and a.dtype == dtype and a.flags.c_contiguous):
        raise TypeError(f"{name} must be a C-contiguous {np.dtype(dtype).name} array (sorted in place)")
    return a


def _void_call(name, *args):
    getattr(lib(), name)(*args)
    st = lib().sblas_last_error()
    if st != 0:
        raise SblasError(st, name)


def _sort_coo(name, m, n, nnz, cooVal, cooRowIdx, cooColIdx):
    v = _inplace(cooVal, np.float64, "cooVal")
    r = _inplace(cooRowIdx, np.int32, "cooRowIdx")
    c = _inplace(cooColIdx, np.int32, "cooColIdx")
    if min(len(v), len(r), len(c)) < nnz:
        raise ValueError("arrays shorter than nnz")
    _void_call(name, m, n, nnz, v.ctypes.data, r.ctypes.data, c.ctypes.data)


def sortCOORow(m, n, nnz, cooVal, cooRowIdx, cooColIdx):
    """Sorts the COO arrays in place by row, then column (on the device)."""
    _sort_coo("sortCOORow", m, n, nnz, cooVal, cooRowIdx, cooColIdx)


def sortCOOCol(m, n, nnz, cooVal, cooRowIdx, cooColIdx):
    """Sorts the COO arrays in place by column, then row (on the device)."""
    _sort_coo("sortCOOCol", m, n, nnz, cooVal, cooRowIdx, cooColIdx)


def _convert(name, m, n, nnz, cooVal, cooRowIdx, cooColIdx, major):
    v, r, c = _f64(cooVal), _i32(cooRowIdx), _i32(cooColIdx)
    if min(len(v), len(r), len(c)) < nnz:
        raise ValueError("arrays shorter than nnz")
    out_v = np.empty(max(nnz, 0), np.float64)
    out_p = np.empty(max(major, 0) + 1, np.int32)
    out_i = np.empty(max(nnz, 0), np.int32)
    _void_call(name, m, n, nnz, v.ctypes.data, r.ctypes.data, c.ctypes.data, out_v.ctypes.data,
               out_p.ctypes.data, out_i.ctypes.data)
    return out_v, out_p, out_i


def coo2csr(m, n, nnz, cooVal, cooRowIdx, cooColIdx):
    """COO (any order) -> (csrVal, csrRowPtr, csrColIdx); the input is not modified."""
    return _convert("coo2csr", m, n, nnz, cooVal, cooRowIdx, cooColIdx, m)


def coo2csc(m, n, nnz, cooVal, cooRowIdx, cooColIdx):
    """COO (any order) -> (cscVal, cscColPtr, cscRowIdx); the input is not modified."""
    return _convert("coo2csc", m, n, nnz, cooVal, cooRowIdx, cooColIdx, n)


SYMMETRY = ("general", "symmetric", "skew-symmetric", "hermitian")


def mm_read_coo(path):
    """A Matrix Market coordinate file as COO: (m, n, row_idx, col_idx, val,
    symmetry) with 0-based int32 indices in file order, float64 values and
    the symmetry name (stored entries only, not expanded)."""
    m, n, nnz, sym = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    pr, pc, pv = ctypes.c_void_p(), ctypes.c_void_p(), ctypes.c_void_p()
    st = lib().sblas_mm_read_coo(os.fsencode(path), ctypes.byref(m), ctypes.byref(n),
                                 ctypes.byref(nnz), ctypes.byref(pr), ctypes.byref(pc),
                                 ctypes.byref(pv), ctypes.byref(sym))
    if st != 0:
        raise SblasError(st, f"mm_read_coo({path})")
    try:
        k = nnz.value
        row = np.ctypeslib.as_array(ctypes.cast(pr, ctypes.POINTER(ctypes.c_int)), (max(k, 1),))[:k].copy()
        col = np.ctypeslib.as_array(ctypes.cast(pc, ctypes.POINTER(ctypes.c_int)), (max(k, 1),))[:k].copy()
        val = np.ctypeslib.as_array(ctypes.cast(pv, ctypes.POINTER(ctypes.c_double)), (max(k, 1),))[:k].copy()
    finally:
        for p in (pr, pc, pv):
            lib().sblas_free(p)
    return m.value, n.value, row, col, val, SYMMETRY[sym.value]
