"""Seeded synthetic workloads (libsblas_gen.so) and host format helpers.

Workloads, all with int32 indices and sorted unique columns per row:

* ``powerlaw_csr``: n x n, row lengths p(L) ~ (L + 2.23)^-2 on [1, 2000]
  (exponent 2 shifted so the mean is ~16), columns uniform, values U(-1, 1).
* ``rmat_csr``: R-MAT (a, b, c, d) = (.57, .19, .19, .05), 2^scale vertices,
  edge_factor * 2^scale draws, duplicates removed, self loops kept, no vertex
  permutation, values 1.0.
* ``banded_csr``: row lengths round(N(140, 40)) clipped to [8, 400], columns
  uniform within +-halfwidth of the diagonal, values N(0, 1).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from ._lib import gen


@dataclass
class CSR:
    m: int
    n: int
    row_ptr: np.ndarray   # int32[m+1]
    col: np.ndarray       # int32[nnz]
    val: np.ndarray       # float64[nnz]

    @property
    def nnz(self) -> int:
        return int(self.row_ptr[-1])


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed ({rc})")


def row_ptr_from_lengths(rowlen: np.ndarray) -> np.ndarray:
    rp = np.zeros(len(rowlen) + 1, dtype=np.int64)
    np.cumsum(rowlen, out=rp[1:])
    if rp[-1] > 2**31 - 1:
        raise ValueError("nnz exceeds int32")
    return rp.astype(np.int32)


def values(count: int, seed: int, tag: int, kind: str = "uniform") -> np.ndarray:
    out = np.empty(count, dtype=np.float64)
    k = {"uniform": 0, "normal": 1, "ones": 2}[kind]
    _check(gen().sblas_gen_values(count, seed, tag, k, out.ctypes.data), "gen values")
    return out


def _fill_cols(m, n, halfwidth, seed, rp):
    col = np.empty(int(rp[-1]), dtype=np.int32)
    _check(gen().sblas_gen_cols(m, n, halfwidth, seed, rp.ctypes.data, col.ctypes.data),
           "gen cols")
    return col


def powerlaw_csr(n: int = 20000, seed: int = 0, lmin: int = 1, lmax: int = 2000,
                 shift: float = 2.23, expo: float = 2.0) -> CSR:
    rowlen = np.empty(n, dtype=np.int32)
    _check(gen().sblas_gen_rowlen_powerlaw(n, seed, lmin, min(lmax, n), shift, expo,
                                           rowlen.ctypes.data), "gen rowlen")
    rp = row_ptr_from_lengths(rowlen)
    col = _fill_cols(n, n, -1, seed, rp)
    return CSR(n, n, rp, col, values(len(col), seed, 1, "uniform"))


def banded_csr(n: int = 2_000_000, seed: int = 0, mu: float = 140.0, sigma: float = 40.0,
               lmin: int = 8, lmax: int = 400, halfwidth: int = 50_000) -> CSR:
    rowlen = np.empty(n, dtype=np.int32)
    _check(gen().sblas_gen_rowlen_normal(n, seed, mu, sigma, lmin, lmax, rowlen.ctypes.data),
           "gen rowlen")
    # a row cannot hold more distinct columns than its window
    r = np.arange(n, dtype=np.int64)
    window = np.minimum(n - 1, r + halfwidth) - np.maximum(0, r - halfwidth) + 1
    rowlen = np.minimum(rowlen, window).astype(np.int32)
    rp = row_ptr_from_lengths(rowlen)
    col = _fill_cols(n, n, halfwidth, seed, rp)
    return CSR(n, n, rp, col, values(len(col), seed, 1, "normal"))


def rmat_csr(scale: int = 22, edge_factor: int = 17, a: float = 0.57, b: float = 0.19,
             c: float = 0.19, seed: int = 0, n: int | None = None) -> CSR:
    """R-MAT with 2^scale vertices; with n, the graph drawn at the next power of
    two >= n and cut to its first n vertices (edges with an endpoint >= n are
    dropped)."""
    if n is not None:
        scale = max(1, int(n - 1).bit_length())
    m = (1 << scale) if n is None else int(n)
    rpp = ctypes.POINTER(ctypes.c_int)()
    cpp = ctypes.POINTER(ctypes.c_int)()
    nnz = gen().sblas_gen_rmat_n(scale, edge_factor * (1 << scale), a, b, c, seed, m,
                                 ctypes.byref(rpp), ctypes.byref(cpp))
    if nnz < 0:
        raise RuntimeError("R-MAT generation failed")
    try:
        rp = np.ctypeslib.as_array(rpp, shape=(m + 1,)).copy()
        col = np.ctypeslib.as_array(cpp, shape=(max(nnz, 1),))[:nnz].copy()
    finally:
        gen().sblas_gen_free(ctypes.cast(rpp, ctypes.c_void_p))
        gen().sblas_gen_free(ctypes.cast(cpp, ctypes.c_void_p))
    return CSR(m, m, rp, col, np.ones(nnz, dtype=np.float64))


def batched_workload(count: int = 64, seed: int = 0, n_lo: int = 10_000, n_hi: int = 2_000_000):
    """count matrices with n log-uniform in [n_lo, n_hi] (seeded): even index a
    banded matrix (rows N(140, 40) clipped to [8, 400], half-width
    min(50,000, n/40)), odd index R-MAT (edge factor 16) drawn at the next power
    of two and cut to n.  Returns [(kind, CSR)]."""
    u = values(count, seed, 77, "uniform") * 0.5 + 0.5
    ns = np.rint(np.exp(np.log(n_lo) + u * (np.log(n_hi) - np.log(n_lo)))).astype(np.int64)
    out = []
    for i, n in enumerate(ns):
        n = int(n)
        if i % 2 == 0:
            out.append(("banded", banded_csr(n, seed=seed + i, halfwidth=min(50_000, n // 40))))
        else:
            out.append(("rmat", rmat_csr(edge_factor=16, seed=seed + i, n=n)))
    return out


def to_csc(A: CSR):
    """(col_ptr, row_idx, val) of A; row indices ascending within a column."""
    cp = np.empty(A.n + 1, dtype=np.int32)
    ri = np.empty(A.nnz, dtype=np.int32)
    cv = np.empty(A.nnz, dtype=np.float64)
    _check(gen().sblas_gen_csr_to_csc(A.m, A.n, A.row_ptr.ctypes.data, A.col.ctypes.data,
                                      A.val.ctypes.data, cp.ctypes.data, ri.ctypes.data,
                                      cv.ctypes.data), "csr->csc")
    return cp, ri, cv


def coo_rows(A: CSR) -> np.ndarray:
    ri = np.empty(A.nnz, dtype=np.int32)
    _check(gen().sblas_gen_csr_to_coo_rows(A.m, A.row_ptr.ctypes.data, ri.ctypes.data),
           "csr->coo")
    return ri


def to_bf16_bits(v: np.ndarray) -> np.ndarray:
    """float64 -> bfloat16 bit patterns, rounded to nearest (ties to even)
    straight from the double: the value is cut to 8 significant bits in
    binary64 (np.rint rounds half to even), which is exact in float32, then
    the float32's upper 16 bits are taken.  Values below the smallest normal
    bfloat16 (2^-126) take the float32 path (round-to-nearest-even of the
    float32 bits), whose double rounding is confined to subnormals."""
    d = np.ascontiguousarray(v, dtype=np.float64)
    frac, ex = np.frexp(d)                        # d = frac * 2^ex, |frac| in [0.5, 1)
    q = np.ldexp(np.rint(np.ldexp(frac, 8)), ex - 8)
    small = np.abs(d) < np.ldexp(1.0, -126)
    with np.errstate(over="ignore"):
        f = q.astype(np.float32).view(np.uint32) >> 16
    fs = d.astype(np.float32).view(np.uint32)
    f_small = (fs + ((fs >> 16) & 1) + np.uint32(0x7FFF)) >> 16
    return np.where(small, f_small, f).astype(np.uint16)


def from_bf16_bits(h: np.ndarray) -> np.ndarray:
    return (h.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
