"""ctypes bindings of libsblas.so (include/sblas.h).

Loading fails loudly when the library has not been built: there is no
fallback path.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libsblas.so")
GEN_PATH = os.path.join(PKG, "libsblas_gen.so")

SBLAS_OK = 0
SBLAS_ERR_ARGS = 1
SBLAS_ERR_NGPU = 2
SBLAS_ERR_KERNEL = 3
SBLAS_ERR_UNDEFINED = 4
SBLAS_ERR_HIP = 5
SBLAS_ERR_ALLOC = 6
SBLAS_ERR_UNSUPPORTED = 7
SBLAS_ERR_NO_DEVICE = 8

VAL_F64, VAL_F32, VAL_BF16 = 0, 1, 2
FMT_CSR, FMT_COO, FMT_CSC = 0, 1, 2
KERNEL_AUTO, KERNEL_VECTOR, KERNEL_MERGE, KERNEL_CSR5 = 0, 1, 2, 3
PART_V1, PART_V1_NUMA, PART_V2, PART_ROWBLOCK = 0, 1, 2, 3

_P = ctypes.c_void_p
_I = ctypes.c_int
_I64 = ctypes.c_int64
_D = ctypes.c_double
_DP = ctypes.POINTER(ctypes.c_double)
_IP = ctypes.POINTER(ctypes.c_int)


class SpmvRet(ctypes.Structure):
    """spmv_ret (include/spmv_kernel.h:16-57 of the reference): seconds."""

    _fields_ = [("numa_part_time", _D), ("part_time", _D), ("comp_time", _D),
                ("comm_time", _D), ("merg_time", _D)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class ChunkDesc(ctypes.Structure):
    _fields_ = [("start_idx", _I64), ("end_idx", _I64), ("start_row", ctypes.c_int32),
                ("end_row", ctypes.c_int32), ("start_flag", ctypes.c_int32),
                ("end_flag", ctypes.c_int32), ("m_c", ctypes.c_int32),
                ("pad_", ctypes.c_int32), ("nnz_c", _I64)]


CHUNK_DTYPE = np.dtype([("start_idx", "<i8"), ("end_idx", "<i8"), ("start_row", "<i4"),
                        ("end_row", "<i4"), ("start_flag", "<i4"), ("end_flag", "<i4"),
                        ("m_c", "<i4"), ("pad_", "<i4"), ("nnz_c", "<i8")])
assert CHUNK_DTYPE.itemsize == ctypes.sizeof(ChunkDesc) == 48


class Opts(ctypes.Structure):
    _fields_ = [("kernel", _I), ("chunks", _I), ("part_mode", _I), ("nb", _I),
                ("streams", _I), ("subwave", _I), ("sigma", _I), ("xcd_remap", _I),
                ("variant", _I), ("items", _I), ("panel_rows", _I), ("narrow_x", _I),
                ("panel_layout", _I), ("reserved", _I * 3)]


class PlanInfo(ctypes.Structure):
    _fields_ = [("format", _I), ("val_type", _I), ("kernel", _I), ("chunks", _I),
                ("subwave", _I), ("sigma", _I), ("xcd_remap", _I), ("split_rows", _I),
                ("m", _I64), ("n", _I64), ("nnz", _I64), ("workspace_bytes", _I64),
                ("mean_row", _D), ("cv_row", _D), ("max_row", _I64), ("launches", _I),
                ("variant", _I), ("row_blocks", _I), ("items", _I),
                ("panel_rows", _I), ("panel_slices", _I), ("tuned", _I),
                ("trial_ms", ctypes.c_float * 3), ("panel_layout", _I), ("narrow_x", _I),
                ("panel_elements", _I64), ("segments", _I), ("virtual_rows", _I),
                ("trial_seg_ms", ctypes.c_float), ("reserved_info", _I)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_ if k != "trial_ms"}
        d["trial_ms"] = [round(float(v), 4) for v in self.trial_ms]
        d["trial_seg_ms"] = round(float(self.trial_seg_ms), 4)
        return d


# Every function include/sblas.h declares, with its ctypes signature.
_DROPIN_CSR = [_I, _I, _I, _DP, _P, _P, _P, _P, _DP, _P, _I]
SIGNATURES = {
    "sblas_last_error": (_I, []),
    "sblas_error_string": (ctypes.c_char_p, [_I]),
    "sblas_version": (ctypes.c_char_p, []),
    "spMV_mgpu_baseline": (SpmvRet, _DROPIN_CSR),
    "spMV_mgpu_baseline_csc": (SpmvRet, _DROPIN_CSR),
    "spMV_mgpu_baseline_coo": (SpmvRet, _DROPIN_CSR),
    "spMV_mgpu_v1": (SpmvRet, _DROPIN_CSR + [_I]),
    "spMV_mgpu_v1_numa": (SpmvRet, _DROPIN_CSR + [_I, _P, _I, _I]),
    "spMV_mgpu_v1_numa_csc": (SpmvRet, [_I, _I, ctypes.c_longlong, _DP, _P, _P, _P, _P, _DP,
                                        _P, _I, _I, _P, _I, _I]),
    "spMV_mgpu_v1_numa_coo": (SpmvRet, _DROPIN_CSR + [_I, _P, _I, _I]),
    "spMV_mgpu_v2": (SpmvRet, _DROPIN_CSR + [_I, _I, _I]),
    "spMspV_mgpu_v1": (SpmvRet, _DROPIN_CSR + [_I]),
    "spMspV_mgpu_v1_numa": (SpmvRet, _DROPIN_CSR + [_I]),
    "spMspV_mgpu_v2": (SpmvRet, _DROPIN_CSR + [_I, _I, _I]),
    "get_row_from_index": (_I, [_I, _P, _I]),
    "findFirstInSorted": (_I, [_P, _I, _I]),
    "findLastInSorted": (_I, [_P, _I, _I]),
    "sblas_partition_rowblock": (_I, [_I, _P, _I, _P]),
    "sblas_partition_baseline_coo": (_I, [_I, _I, _P, _I, _P]),
    "sblas_partition_baseline_csc": (_I, [_I, _P, _I, _P]),
    "sblas_partition_v1": (_I, [_I, _P, _I, _I, _P]),
    "sblas_partition_v1_numa": (_I, [_I, _P, _I, _I, _P, _P, _P]),
    "sblas_partition_v1_numa_coo": (_I, [_I, _I, _P, _I, _P, _P, _P]),
    "sblas_partition_v1_numa_csc": (_I, [_I, _P, ctypes.c_longlong, _I, _P, _P, _P]),
    "sblas_v2_num_tasks": (_I, [_I, _I, _I]),
    "sblas_partition_v2": (_I, [_I, _P, _I, _I, _I, _P, _I]),
    "sblas_opts_default": (None, [ctypes.POINTER(Opts)]),
    "sblas_plan_create_csr": (_I, [_I, _I, _I64, _P, _P, _P, _I, ctypes.POINTER(Opts),
                                   ctypes.POINTER(_P)]),
    "sblas_plan_create_coo": (_I, [_I, _I, _I64, _P, _P, _P, _I, ctypes.POINTER(Opts),
                                   ctypes.POINTER(_P)]),
    "sblas_plan_create_csc": (_I, [_I, _I, _I64, _P, _P, _P, _I, ctypes.POINTER(Opts),
                                   ctypes.POINTER(_P)]),
    "sblas_plan_spmv": (_I, [_P, _D, _P, _D, _P, _P]),
    "sblas_spmv_csr_noalloc": (_I, [_I, _I, _I64, _P, _P, _P, _I, _D, _P, _D, _P, _P]),
    "sblas_plan_spmm": (_I, [_P, _I, _D, _P, _D, _P, _P]),
    "sblas_plan_get_info": (_I, [_P, ctypes.POINTER(PlanInfo)]),
    "sblas_plan_get_chunks": (_I, [_P, _P, _I]),
    "sblas_plan_destroy": (_I, [_P]),
    "sblas_plan_set_timing": (_I, [_P, _I]),
    "sblas_plan_kernel_times": (_I, [_P, _P, _I]),
    "sblas_batch_create": (_I, [_I, _P, ctypes.POINTER(Opts), ctypes.POINTER(_P)]),
    "sblas_batch_run": (_I, [_P, _P]),
    "sblas_batch_graph_launch": (_I, [_P, _P]),
    "sblas_batch_get_choice": (_I, [_P, _I, _P]),
    "sblas_batch_get_shape": (_I, [_P, _P, _P, _P]),
    "sblas_batch_destroy": (_I, [_P]),
    "sortCOORow": (None, [_I, _I, _I, _P, _P, _P]),
    "sortCOOCol": (None, [_I, _I, _I, _P, _P, _P]),
    "coo2csr": (None, [_I, _I, _I, _P, _P, _P, _P, _P, _P]),
    "coo2csc": (None, [_I, _I, _I, _P, _P, _P, _P, _P, _P]),
    "sblas_coo_sort": (_I, [_I, _I, _I64, _P, _P, _P, _I, _I, _P]),
    "sblas_coo_ptr": (_I, [_I, _I64, _P, _P, _P]),
    "sblas_stream_spmv_csr": (_I, [_I, _I, _I64, _P, _P, _P, _D, _P, _D, _P, _I, _I, _P]),
    "sblas_csr_drop_zero_x": (_I, [_I, _I, _I64, _P, _P, _P, _P, _P, _P, _P,
                                   ctypes.POINTER(_I64), _P]),
    "sblas_mm_read_coo": (_I, [ctypes.c_char_p, ctypes.POINTER(_I), ctypes.POINTER(_I),
                               ctypes.POINTER(_I), ctypes.POINTER(_P), ctypes.POINTER(_P),
                               ctypes.POINTER(_P), ctypes.POINTER(_I)]),
    "sblas_free": (None, [_P]),
}


class BatchItem(ctypes.Structure):
    _fields_ = [("m", _I), ("n", _I), ("nnz", _I64), ("d_row_ptr", _P), ("d_col", _P),
                ("d_val", _P), ("val_type", _I), ("k", _I), ("d_x", _P), ("d_y", _P),
                ("alpha", _D), ("beta", _D)]


class BatchChoice(ctypes.Structure):
    _fields_ = [("kernel", _I), ("subwave", _I), ("items", _I), ("chunks", _I),
                ("row_blocks", _I), ("variant", _I), ("mean_row", _D), ("cv_row", _D),
                ("max_row", _I64), ("panel_rows", _I), ("reserved_", _I)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved_"}

_lib = None
_gen = None


class SblasError(RuntimeError):
    def __init__(self, status: int, where: str = ""):
        self.status = status
        msg = lib().sblas_error_string(status).decode() if _lib is not None else str(status)
        super().__init__(f"{where}: sblas status {status} ({msg})")


def lib() -> ctypes.CDLL:
    """The engine library.  Raises when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m hpda_sblas_pnnl_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(status: int, where: str) -> None:
    if status != SBLAS_OK:
        raise SblasError(status, where)


def gen() -> ctypes.CDLL:
    global _gen
    if _gen is None:
        if not os.path.exists(GEN_PATH):
            raise RuntimeError(f"{GEN_PATH} is missing: build with `python -m hpda_sblas_pnnl_amd.build`")
        G = ctypes.CDLL(GEN_PATH)
        G.sblas_gen_rowlen_powerlaw.argtypes = [_I, ctypes.c_uint64, _I, _I, _D, _D, _P]
        G.sblas_gen_rowlen_normal.argtypes = [_I, ctypes.c_uint64, _D, _D, _I, _I, _P]
        G.sblas_gen_cols.argtypes = [_I, _I, _I, ctypes.c_uint64, _P, _P]
        G.sblas_gen_values.argtypes = [_I64, ctypes.c_uint64, ctypes.c_uint64, _I, _P]
        G.sblas_gen_rmat.restype = _I64
        G.sblas_gen_rmat.argtypes = [_I, _I64, _D, _D, _D, ctypes.c_uint64,
                                     ctypes.POINTER(_IP), ctypes.POINTER(_IP)]
        G.sblas_gen_rmat_n.restype = _I64
        G.sblas_gen_rmat_n.argtypes = [_I, _I64, _D, _D, _D, ctypes.c_uint64, _I64,
                                       ctypes.POINTER(_IP), ctypes.POINTER(_IP)]
        G.sblas_gen_free.argtypes = [_P]
        G.sblas_gen_free.restype = None
        G.sblas_gen_csr_to_csc.argtypes = [_I, _I, _P, _P, _P, _P, _P, _P]
        G.sblas_gen_csr_to_coo_rows.argtypes = [_I, _P, _P]
        for f in ("sblas_gen_rowlen_powerlaw", "sblas_gen_rowlen_normal", "sblas_gen_cols",
                  "sblas_gen_values", "sblas_gen_csr_to_csc", "sblas_gen_csr_to_coo_rows"):
            getattr(G, f).restype = _I
        _gen = G
    return _gen
