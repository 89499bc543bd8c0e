"""MI355X-native SpMV engine behind the C entry points of PNNL's Sparse-BLAS
(spMV_mgpu_baseline / _v1 / _v1_numa / _v2 with CSR, CSC and COO variants).

* ``hpda_sblas_pnnl_amd.spmv``   -- the reference's entry points on host arrays
* ``hpda_sblas_pnnl_amd.device`` -- plans over device-resident tensors
* ``hpda_sblas_pnnl_amd.gen``    -- seeded synthetic workloads
* ``include/sblas.h``            -- the C ABI of libsblas.so
"""
from ._lib import SblasError, lib  # noqa: F401

__all__ = ["SblasError", "lib"]
