"""Python access to the oracle (oracle/liboracle.so) and to the reference's own
host functions compiled by oracle/build_ref.sh (oracle/_ref/libref_host.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the CPU
baseline leg of bench.py, as the checker.  The product package never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libref_host.so")

ORA_OK = 0
ORA_BAD_ARGS = -1
ORA_UNDEFINED = -2

CHUNK_DTYPE = np.dtype(
    [
        ("start_idx", "<i8"),
        ("end_idx", "<i8"),
        ("start_row", "<i4"),
        ("end_row", "<i4"),
        ("start_flag", "<i4"),
        ("end_flag", "<i4"),
        ("m_c", "<i4"),
        ("pad_", "<i4"),
        ("nnz_c", "<i8"),
    ]
)
assert CHUNK_DTYPE.itemsize == 48

_P = ctypes.c_void_p
_I = ctypes.c_int
_L = ctypes.c_int64


def build() -> None:
    """Compile liboracle.so (and the reference host functions when the
    reference tree is present)."""
    subprocess.check_call(["make", "-s", "-B", "-C", HERE, "liboracle.so"])
    if os.path.isdir(os.environ.get("SBLAS_REFERENCE_DIR", "/root/reference")):
        subprocess.check_call(["sh", os.path.join(HERE, "build_ref.sh")],
                              stdout=subprocess.DEVNULL)


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            subprocess.check_call(["make", "-s", "-C", HERE, "liboracle.so"])
        L = ctypes.CDLL(LIB_PATH)
        sig = {
            "ora_get_row_from_index": (_I, [_I, _P, _L, _I]),
            "ora_find_first_in_sorted": (_I, [_P, _I, _L, _I]),
            "ora_find_last_in_sorted": (_I, [_P, _I, _L, _I]),
            "ora_part_rowblock": (_I, [_I, _P, _I, _P]),
            "ora_part_baseline_coo": (_I, [_I, _I, _P, _I, _P]),
            "ora_part_baseline_csc": (_I, [_I, _P, _I, _P]),
            "ora_part_v1": (_I, [_I, _P, _I, _I, _P]),
            "ora_part_v1_numa": (_I, [_I, _P, _I, _I, _P, _P, _P]),
            "ora_part_v1_numa_coo": (_I, [_I, _I, _P, _I, _P, _P, _P]),
            "ora_part_v1_numa_csc": (_I, [_I, _P, ctypes.c_longlong, _I, _P, _P, _P]),
            "ora_v2_num_tasks": (_I, [_I, _I, _I]),
            "ora_part_v2": (_I, [_I, _P, _I, _I, _I, _P, _I]),
            "ora_csr_spmv": (None, [_I, _I, _I, _P, _P, _P, _P, _P, _P, _P]),
            "ora_csr_spmv_omp": (None, [_I, _I, _I, _P, _P, _P, _P, _P, _P, _P, _I]),
            "ora_coo_spmv": (None, [_I, _I, _I, _P, _P, _P, _P, _P, _P, _P]),
            "ora_csr_spmv_f32val": (None, [_I, _P, _P, _P, _P, ctypes.c_double,
                                           ctypes.c_double, _P]),
            "ora_csr_spmv_bf16val": (None, [_I, _P, _P, _P, _P, ctypes.c_double,
                                            ctypes.c_double, _P]),
            "ora_csr_spmm": (None, [_I, _I, _P, _P, _P, _P, ctypes.c_double,
                                    ctypes.c_double, _P]),
            "ora_gen_g_nnz": (_L, [_I]),
            "ora_gen_g": (_L, [_I, _P, _P, _P, _L]),
            "ora_err_normwise": (ctypes.c_double, [_L, _P, _P]),
            "ora_err_maxrow": (ctypes.c_double, [_I, _P, _P, _P, _P, ctypes.c_double,
                                                 ctypes.c_double, _P, _P, _P]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a: np.ndarray | None):
    return None if a is None else a.ctypes.data


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(a, dtype=np.float64)


# --------------------------------------------------------------------------
# searches
# --------------------------------------------------------------------------

def get_row_from_index(n: int, a, idx: int, alen: int | None = None) -> int:
    a = _i32(a)
    return lib().ora_get_row_from_index(n, _p(a), len(a) if alen is None else alen, idx)


def find_first_in_sorted(a, n: int, key: int, alen: int | None = None) -> int:
    a = _i32(a)
    return lib().ora_find_first_in_sorted(_p(a), n, len(a) if alen is None else alen, key)


def find_last_in_sorted(a, n: int, key: int, alen: int | None = None) -> int:
    a = _i32(a)
    return lib().ora_find_last_in_sorted(_p(a), n, len(a) if alen is None else alen, key)


# --------------------------------------------------------------------------
# partitions: each returns (status, chunks[, nodes])
# --------------------------------------------------------------------------

def _chunks(n: int) -> np.ndarray:
    return np.zeros(max(n, 1), dtype=CHUNK_DTYPE)


def _mapping(C: int, mapping):
    if mapping is None:
        return None, 1
    mp = _i32(mapping)
    assert len(mp) == C
    return mp, int(mp.max()) + 1


def part_rowblock(m, rp, C):
    rp = _i32(rp)
    out = _chunks(C)
    st = lib().ora_part_rowblock(m, _p(rp), C, _p(out))
    return st, out[:C]


def part_baseline_coo(m, row_idx, C):
    ri = _i32(row_idx)
    out = _chunks(C)
    st = lib().ora_part_baseline_coo(m, len(ri), _p(ri), C, _p(out))
    return st, out[:C]


def part_baseline_csc(n, cp, C):
    cp = _i32(cp)
    out = _chunks(C)
    st = lib().ora_part_baseline_csc(n, _p(cp), C, _p(out))
    return st, out[:C]


def part_v1(m, rp, nnz, C):
    rp = _i32(rp)
    out = _chunks(C)
    st = lib().ora_part_v1(m, _p(rp), nnz, C, _p(out))
    return st, out[:C]


def part_v1_numa(m, rp, nnz, C, mapping=None):
    rp = _i32(rp)
    mp, nodes = _mapping(C, mapping)
    out, nout = _chunks(C), _chunks(nodes)
    st = lib().ora_part_v1_numa(m, _p(rp), nnz, C, _p(mp), _p(out), _p(nout))
    return st, out[:C], nout[:nodes]


def part_v1_numa_coo(m, row_idx, C, mapping=None):
    ri = _i32(row_idx)
    mp, nodes = _mapping(C, mapping)
    out, nout = _chunks(C), _chunks(nodes)
    st = lib().ora_part_v1_numa_coo(m, len(ri), _p(ri), C, _p(mp), _p(out), _p(nout))
    return st, out[:C], nout[:nodes]


def part_v1_numa_csc(n, cp, nnz, C, mapping=None):
    cp = _i32(cp)
    mp, nodes = _mapping(C, mapping)
    out, nout = _chunks(C), _chunks(nodes)
    st = lib().ora_part_v1_numa_csc(n, _p(cp), nnz, C, _p(mp), _p(out), _p(nout))
    return st, out[:C], nout[:nodes]


def v2_num_tasks(nnz, nb, q):
    return lib().ora_v2_num_tasks(nnz, nb, q)


def part_v2(m, rp, nnz, nb, q):
    rp = _i32(rp)
    T = v2_num_tasks(nnz, nb, q)
    if T < 0:
        return T, _chunks(0)[:0]
    out = _chunks(T)
    st = lib().ora_part_v2(m, _p(rp), nnz, nb, q, _p(out), T)
    return (ORA_OK if st >= 0 else st), out[:T]


# --------------------------------------------------------------------------
# SpMV oracles
# --------------------------------------------------------------------------

def csr_spmv(m, n, val, rp, col, x, alpha, beta, y, threads: int = 0):
    """y <- alpha*A*x + beta*y with the reference's host loop; returns a new array."""
    val, rp, col, x = _f64(val), _i32(rp), _i32(col), _f64(x)
    y = np.array(y, dtype=np.float64, copy=True)
    a = np.array([alpha], dtype=np.float64)
    b = np.array([beta], dtype=np.float64)
    if threads:
        lib().ora_csr_spmv_omp(m, n, len(val), _p(a), _p(val), _p(rp), _p(col), _p(x),
                               _p(b), _p(y), threads)
    else:
        lib().ora_csr_spmv(m, n, len(val), _p(a), _p(val), _p(rp), _p(col), _p(x),
                           _p(b), _p(y))
    return y


def coo_spmv(m, n, val, row_idx, col_idx, x, alpha, beta, y):
    val, ri, ci, x = _f64(val), _i32(row_idx), _i32(col_idx), _f64(x)
    y = np.array(y, dtype=np.float64, copy=True)
    a = np.array([alpha], dtype=np.float64)
    b = np.array([beta], dtype=np.float64)
    lib().ora_coo_spmv(m, n, len(val), _p(a), _p(val), _p(ri), _p(ci), _p(x), _p(b), _p(y))
    return y


def csr_spmv_f32val(m, val32, rp, col, x, alpha, beta, y):
    val32 = np.ascontiguousarray(val32, dtype=np.float32)
    rp, col, x = _i32(rp), _i32(col), _f64(x)
    y = np.array(y, dtype=np.float64, copy=True)
    lib().ora_csr_spmv_f32val(m, _p(val32), _p(rp), _p(col), _p(x), alpha, beta, _p(y))
    return y


def csr_spmv_bf16val(m, val16, rp, col, x, alpha, beta, y):
    val16 = np.ascontiguousarray(val16, dtype=np.uint16)
    rp, col, x = _i32(rp), _i32(col), _f64(x)
    y = np.array(y, dtype=np.float64, copy=True)
    lib().ora_csr_spmv_bf16val(m, _p(val16), _p(rp), _p(col), _p(x), alpha, beta, _p(y))
    return y


def csr_spmm(m, k, val, rp, col, X, alpha, beta, Y):
    val, rp, col = _f64(val), _i32(rp), _i32(col)
    X = _f64(X)
    Y = np.array(Y, dtype=np.float64, copy=True)
    lib().ora_csr_spmm(m, k, _p(val), _p(rp), _p(col), _p(X), alpha, beta, _p(Y))
    return Y


def gen_g(n):
    """The reference harness's 'g' matrix as row-sorted COO (rows, cols, vals)."""
    nnz = lib().ora_gen_g_nnz(n)
    if nnz < 0:
        raise ValueError(f"'g' generator undefined for n={n}")
    r = np.empty(nnz, np.int32)
    c = np.empty(nnz, np.int32)
    v = np.empty(nnz, np.float64)
    got = lib().ora_gen_g(n, _p(r), _p(c), _p(v), nnz)
    assert got == nnz
    return r, c, v


def coo_sort(row, col, val, by_col=False):
    """COO sorted by (row, column), or by (column, row) with by_col, equal
    pairs in input order: the order cusparseXcoosortByRow / ByColumn plus the
    value gather leave (src/spmv_helper.cu:277-387; the tie order of equal
    keys is the stable one, cuSPARSE's own is not published).  Returns new
    (row, col, val) arrays."""
    row, col, val = np.asarray(row), np.asarray(col), np.asarray(val)
    order = np.lexsort((row, col)) if by_col else np.lexsort((col, row))  # stable
    return row[order], col[order], val[order]


def coo_ptr(count, sorted_idx):
    """Compressed pointer array of a sorted index array (cusparseXcoo2csr,
    src/spmv_helper.cu:478-491): ptr[r] = first position of an index >= r."""
    return np.searchsorted(np.asarray(sorted_idx), np.arange(count + 1), side="left").astype(
        np.int32)


def coo_to_csr(m, n, row, col, val):
    """coo2csr (src/spmv_helper.cu:493-573): (csrVal, csrRowPtr, csrColIdx)."""
    r, c, v = coo_sort(row, col, val)
    return v, coo_ptr(m, r), c.astype(np.int32)


def coo_to_csc(m, n, row, col, val):
    """coo2csc (src/spmv_helper.cu:575-642): (cscVal, cscColPtr, cscRowIdx)."""
    r, c, v = coo_sort(row, col, val, by_col=True)
    return v, coo_ptr(n, c), r.astype(np.int32)


def csr_drop_zero_x(m, rp, col, val, x):
    """spMspV's filter (src/dspmv_mgpu_v1.cu:19-74): the CSR matrix without the
    elements whose x[col] == 0 (-0.0 included, NaN kept), rows in order.
    Returns (row_ptr, col, val)."""
    rp, col, val, x = np.asarray(rp), np.asarray(col), np.asarray(val), np.asarray(x)
    keep = x[col] != 0.0
    kept = np.concatenate([[0], np.cumsum(keep)]).astype(np.int64)
    return kept[rp].astype(np.int32), col[keep].astype(np.int32), val[keep]


def err_normwise(y, r) -> float:
    y, r = _f64(y), _f64(r)
    return lib().ora_err_normwise(len(y), _p(y), _p(r))


def err_maxrow(m, val, rp, col, x, alpha, beta, y0, y, r) -> float:
    val, rp, col, x = _f64(val), _i32(rp), _i32(col), _f64(x)
    y0 = None if y0 is None else _f64(y0)
    y, r = _f64(y), _f64(r)
    return lib().ora_err_maxrow(m, _p(val), _p(rp), _p(col), _p(x), alpha, beta,
                                _p(y0), _p(y), _p(r))


# --------------------------------------------------------------------------
# the reference's own compiled host functions (oracle/_ref), when present
# --------------------------------------------------------------------------

_ref = None


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref():
    """ctypes handle on the reference's host functions (C++ linkage names)."""
    global _ref
    if _ref is None:
        R = ctypes.CDLL(REF_PATH)
        names = {
            "get_row_from_index": ("_Z18get_row_from_indexiPii", _I, [_I, _P, _I]),
            "findFirstInSorted": ("_Z17findFirstInSortedPiii", _I, [_P, _I, _I]),
            "findLastInSorted": ("_Z16findLastInSortedPiii", _I, [_P, _I, _I]),
            "csr_spmv_cpu": ("_Z12csr_spmv_cpuiiiPdS_PiS0_S_S_S_", None,
                             [_I, _I, _I, _P, _P, _P, _P, _P, _P, _P]),
            "coo_spmv_cpu": ("_Z12coo_spmv_cpuiiiPdS_PiS0_S_S_S_", None,
                             [_I, _I, _I, _P, _P, _P, _P, _P, _P, _P]),
        }

        class _Ref:
            pass

        h = _Ref()
        for py, (sym, res, args) in names.items():
            f = getattr(R, sym)
            f.restype = res
            f.argtypes = args
            setattr(h, py, f)
        _ref = h
    return _ref
