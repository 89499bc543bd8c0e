// plan.hpp -- the engine's plan object behind the opaque sblas_plan handle.
#pragma once

#include <vector>

#include "internal.hpp"

namespace sblas {

struct RowStats {
  double mean = 0, cv = 0, empty_frac = 0;
  int64_t max_row = 0;
};

RowStats row_stats(int m, const int *rp);
int autotune_kernel(const RowStats &s);
int autotune_subwave(const RowStats &s);
int autotune_sigma(const RowStats &s);
const char *last_message();
void clear_error();

struct Plan {
  int format = SBLAS_FMT_CSR;
  int val_type = SBLAS_VAL_F64;
  int m = 0, n = 0;
  int64_t nnz = 0;
  // caller arrays (device): CSR rp/col/val, COO row/col/val, CSC cp/ri/val
  const int *a0 = nullptr;
  const int *a1 = nullptr;
  const void *val = nullptr;
  sblas_opts opts{};
  int device = 0;
  int kernel = SBLAS_KERNEL_VECTOR;
  int chunks = 1;
  RowStats stats;

  // reference descriptors of the chunking rule and their status
  std::vector<sblas_chunk_desc> descs;
  int desc_status = SBLAS_OK;

  // execution
  std::vector<ExecChunk> exec;
  std::vector<DevChunk> dev_chunks;
  DevChunk *d_chunks = nullptr;       // fused launch table
  DevChunk *d_chunks_solo = nullptr;  // per-chunk tables (blk_begin = 0)
  VectorLaunch vl{};
  std::vector<VectorLaunch> solo_vl;
  int *d_mrows = nullptr, *d_mptr = nullptr, *d_mslots = nullptr;
  int ngroups = 0;
  int split_rows = 0;
  double *d_partials = nullptr;       // 2 slots per chunk, x max_k
  int max_k = 16;

  // segmented tiles (COO / merge path / CSR5)
  int tiles = 0;
  int sigma = 0;
  TileCarry *d_carry = nullptr;
  int *d_tile_row = nullptr;
  uint32_t *d_tile_flags = nullptr;

  // CSC: row-major copy built once on the device (stable by-row transpose)
  int *d_csr_rp = nullptr, *d_csr_col = nullptr;
  void *d_csr_val = nullptr;

  std::vector<hipStream_t> streams;
  std::vector<hipEvent_t> events;
  std::vector<void *> device_buffers;
  int64_t workspace_bytes = 0;
  int launches = 0;

  // kernel timing (measurement only)
  bool timing = false;
  std::vector<hipEvent_t> tev;   // pairs: [2i] before, [2i+1] after
  int tcount = 0;

  ~Plan();
  void *dev_alloc(size_t bytes);
  int build_csr_chunks(const int *h_rp);
  int setup_vector(const int *h_rp);
  hipEvent_t *timing_slot();
  int run_spmv(double alpha, const double *x, double beta, double *y, hipStream_t s);
  int run_spmm(int k, double alpha, const double *X, double beta, double *Y, hipStream_t s);
};

}  // namespace sblas

struct sblas_plan_s {
  sblas::Plan impl;
};
