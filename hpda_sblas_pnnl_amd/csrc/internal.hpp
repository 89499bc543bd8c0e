// internal.hpp -- shared declarations of libsblas (not part of the public ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "partition.hpp"
#include "sblas.h"

namespace sblas {

// ---- status -----------------------------------------------------------------
void set_error(int status, const std::string &what);
int last_error();

#define SBLAS_HIP_TRY(expr)                                                    \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess) {                                                    \
      ::sblas::set_error(SBLAS_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
      return SBLAS_ERR_HIP;                                                    \
    }                                                                          \
  } while (0)

// ---- device-side chunk table ---------------------------------------------------
// One execution chunk: element range [lo, hi), row range [row_begin, row_end),
// first launch block of the chunk.  Split rows write partial sums to
// partials[2*chunk + (row == row_begin ? 0 : 1)].
struct DevChunk {
  int lo, hi;
  int row_begin, row_end;
  int blk_begin;
  int pad_[3];
};

// Segmented-tile carries (COO / merge-path / CSR5 kernels): per launch block,
// the partial of its first segment when that segment began in an earlier
// block (head) and of its last segment when it continues into the next block
// (tail).  Combined in block order by k_tile_fixup.
struct TileCarry {
  double head;
  double tail;
  int head_row;   // -1: first segment began in this block
  int tail_row;   // -1: last segment ends in this block
  int spanning;   // 1: the whole block lies inside one row (head only)
  int pad_;
};

// ---- launchers (sblas_kernels.hip) --------------------------------------------
struct VectorLaunch {
  int subwave;          // lanes per row: 2..64
  int rows_per_block;
  int total_blocks;     // logical blocks over all chunks
  int grid;             // launched blocks (>= total_blocks; multiple of 8 with remap)
  int xcd_remap;
};

int launch_csr_vector(int val_type, const VectorLaunch &L, const DevChunk *chunks, int nchunks,
                      const int *rp, const int *col, const void *val, const double *x,
                      double *y, double alpha, double beta, double *partials,
                      hipStream_t s);
int launch_csr_vector_mrhs(int val_type, int k, const VectorLaunch &L, const DevChunk *chunks,
                           int nchunks, const int *rp, const int *col, const void *val,
                           const double *X, double *Y, double alpha, double beta,
                           double *partials, hipStream_t s);
int launch_merge(const int *rows, const int *ptr, const int *slots, int groups,
                 const double *partials, int k, double *y, double alpha, double beta,
                 hipStream_t s);
int launch_scale(int64_t count, double beta, double *y, hipStream_t s);

// segmented tiles
int launch_coo_tiles(int val_type, int64_t nnz, int tiles, const int *row_idx,
                     const int *col_idx, const void *val, const double *x, double *y,
                     double alpha, double beta, TileCarry *carry, hipStream_t s);
int launch_tile_fixup(int tiles, const TileCarry *carry, double *y, double alpha, double beta,
                      hipStream_t s);
int launch_csr_merge_path(int val_type, int m, int64_t nnz, int tiles, const int *rp,
                          const int *col, const void *val, const double *x, double *y,
                          double alpha, double beta, TileCarry *carry, hipStream_t s);
int launch_csr5(int val_type, int m, int64_t nnz, int tiles, int sigma, const int *tile_row,
                const uint32_t *tile_flags, const int *rp, const int *col, const void *val,
                const double *x, double *y, double alpha, double beta, TileCarry *carry,
                hipStream_t s);
int launch_csr5_build(int m, int64_t nnz, int tiles, int sigma, const int *rp, int *tile_row,
                      uint32_t *tile_flags, hipStream_t s);
int launch_empty_rows(int m, const int *rp, double *y, double beta, hipStream_t s);

// CSC
int launch_csc_count(int m, int n, int64_t nnz, const int *cp, const int *ri, int *row_cnt,
                     hipStream_t s);

constexpr int kTileBlock = 256;   // threads per segmented-tile block
constexpr int kTileItems = 8;     // elements per thread per tile (COO / merge path)

}  // namespace sblas
