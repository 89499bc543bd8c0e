"""MI355X-native SpMV engine behind the C entry points of PNNL's Sparse-BLAS
(spMV_mgpu_baseline / _v1 / _v1_numa / _v2 with CSR, CSC and COO variants).

The reference's entry points on host arrays live in ``spmv``, plans and
batches over device-resident tensors in ``device``, seeded workloads in
``gen``; ``include/sblas.h`` is the C ABI of libsblas.so underneath.
"""
from ._lib import SblasError, lib
from .device import SpmvBatch, SpmvPlan
from .spmv import (spMV_mgpu_baseline, spMV_mgpu_baseline_coo, spMV_mgpu_baseline_csc,
                   spMV_mgpu_v1, spMV_mgpu_v1_numa, spMV_mgpu_v1_numa_coo, spMV_mgpu_v1_numa_csc,
                   spMV_mgpu_v2, spMspV_mgpu_v1, spMspV_mgpu_v2)

__all__ = ["SblasError", "lib", "SpmvPlan", "SpmvBatch", "spMV_mgpu_baseline",
           "spMV_mgpu_baseline_coo", "spMV_mgpu_baseline_csc", "spMV_mgpu_v1", "spMV_mgpu_v1_numa",
           "spMV_mgpu_v1_numa_coo", "spMV_mgpu_v1_numa_csc", "spMV_mgpu_v2", "spMspV_mgpu_v1",
           "spMspV_mgpu_v2"]
