// partition.hpp -- the reference's chunking rules (bit-exact descriptors) and
// the execution ranges derived from them.
#pragma once

#include <cstdint>
#include <vector>

#include "sblas.h"

namespace sblas {

// Bisection searches of src/spmv_helper.cu (18-44, 646-664, 667-685).  `alen`
// bounds the one read the reference makes past the searched range (a[n]).
int search_row_of(int n, const int *a, int64_t alen, int idx);     // get_row_from_index
int search_first(const int *a, int n, int64_t alen, int key);      // findFirstInSorted
int search_last(const int *a, int n, int64_t alen, int key);       // findLastInSorted

int partition_rowblock(int m, const int *rp, int C, sblas_chunk_desc *out);
int partition_baseline_coo(int m, int nnz, const int *ri, int C, sblas_chunk_desc *out);
int partition_v1(int m, const int *rp, int nnz, int C, sblas_chunk_desc *out);
int partition_v1_numa(int m, const int *rp, int nnz, int C, const int *mapping,
                      sblas_chunk_desc *out, sblas_chunk_desc *node_out);
int partition_v1_numa_coo(int m, int nnz, const int *ri, int C, const int *mapping,
                          sblas_chunk_desc *out, sblas_chunk_desc *node_out);
int partition_v1_numa_csc(int n, const int *cp, long long nnz, int C, const int *mapping,
                          sblas_chunk_desc *out, sblas_chunk_desc *node_out);
int v2_num_tasks(int nnz, int nb, int q);
int partition_v2(int m, const int *rp, int nnz, int nb, int q, sblas_chunk_desc *out,
                 int cap);

// Execution range of one chunk: element range [lo, hi) and the row range
// [row_begin, row_end) it writes.  Rows in the range whose elements are not
// all inside [lo, hi) are "split" and produce partial sums that the device
// merge combines.
struct ExecChunk {
  int64_t lo, hi;
  int32_t row_begin, row_end;
};

// Builds execution chunks from element boundaries bounds[0..C] (bounds[0] = 0,
// bounds[C] = nnz, non-decreasing): every row of [0, m) lies in at least one
// chunk's row range, rows outside every chunk's elements (empty rows at the
// edges) included.  Empty element ranges are dropped.
std::vector<ExecChunk> exec_chunks_from_bounds(int m, const int *rp,
                                               const std::vector<int64_t> &bounds);

// Element boundaries of the v1 / v1_numa / v2 / rowblock rules at chunk count C
// computed with 64-bit arithmetic (equal to the reference wherever the
// reference is defined); v1_numa's double formula can drop trailing elements
// (the reference loses them); the returned bounds always end at nnz.
std::vector<int64_t> element_bounds(int part_mode, int m, const int *rp, int64_t nnz,
                                    int C, int nb);

}  // namespace sblas
