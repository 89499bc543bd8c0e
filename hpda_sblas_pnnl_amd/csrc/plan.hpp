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
int autotune_subwave(const RowStats &s);
int autotune_sigma(int64_t nnz, int m);
std::vector<RowBlock> build_row_blocks(const std::vector<ExecChunk> &ex, const int *rp,
                                       int cap);
const char *last_message();
// Vector-kernel (sub-wave) launch shape over execution chunks; dev receives
// one DevChunk per chunk.
VectorLaunch vector_shape(const std::vector<ExecChunk> &ex, double mean, int subwave,
                          int xcd_remap, std::vector<DevChunk> &dev);
// Rows that more than one chunk touches, with their partial slots in chunk
// order (slot 2c: chunk c's first row, 2c+1: its last row).
void merge_lists(const std::vector<ExecChunk> &ex, const int *rp, std::vector<int> &rows,
                 std::vector<int> &ptr, std::vector<int> &slots);

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

  // row-block (LDS-staged) form of kernel 1
  int variant = 1;                    // 1 sub-wave per row, 2 row blocks
  std::vector<RowBlock> rblocks;
  RowBlock *d_rblocks = nullptr;
  StreamLaunch sl{};
  int *d_bmrows = nullptr, *d_bmptr = nullptr, *d_bmslots = nullptr;
  int bgroups = 0;
  double *d_bpartials = nullptr;
  // multi-vector products over row blocks of 16 elements a thread where
  // those pay (large matrices with irregular rows); grid 0 otherwise, and
  // the row blocks above serve them
  StreamLaunch slk{};
  RowBlock *d_kblocks = nullptr;
  int *d_kmrows = nullptr, *d_kmptr = nullptr, *d_kmslots = nullptr;
  int kgroups = 0;
  double *d_kpartials = nullptr;

  // segmented tiles (COO / merge path / CSR5)
  int tiles = 0;
  int sigma = 0;
  TileCarry *d_carry = nullptr;
  int *d_tile_row = nullptr;
  uint32_t *d_tile_flags = nullptr;

  // column-panel form of kernel 1 (variant 3); rows cut by chunk edges
  PanelFormat panel;
  int *d_pmrows = nullptr, *d_pmptr = nullptr, *d_pmslots = nullptr;
  int pgroups = 0;

  // column-sorted row blocks of kernel 1 (variant 4); rows cut by block edges
  SortedFormat sorted;
  int *d_smrows = nullptr, *d_smptr = nullptr, *d_smslots = nullptr;
  int sgroups = 0;

  // row blocks over column segments of kernel 1 (variant 6): the format, the
  // row blocks over its virtual rows and the merge lists of the virtual rows
  // their windows cut
  SegmentFormat seg;
  StreamLaunch vsl{};
  RowBlock *d_vblocks = nullptr;
  int *d_vmrows = nullptr, *d_vmptr = nullptr, *d_vmslots = nullptr;
  int vgroups = 0;
  double *d_vpartials = nullptr;

  // CSR5-style format (kernel 3)
  uint32_t *d_c5_tile_nz = nullptr, *d_c5_flags = nullptr;
  int *d_c5_rows_nz = nullptr, *d_c5_empty = nullptr;
  int *d_c5_colT = nullptr;
  void *d_c5_valT = nullptr;
  int c5_empty = 0;

  // CSC: row-major copy built once on the device (stable by-row transpose)
  int *d_csr_rp = nullptr, *d_csr_col = nullptr;
  void *d_csr_val = nullptr;

  std::vector<hipStream_t> streams;
  std::vector<hipEvent_t> events;
  std::vector<void *> device_buffers;
  std::vector<int64_t> buffer_bytes;  // workspace_bytes before each buffer
  int64_t workspace_bytes = 0;
  // how the kernel was chosen: 0 asked for, 1 rule, 2 timed trial, 3 cached
  // trial; the trial's median times (row blocks, CSR5-style) in ms
  int tuned = 0;
  // row blocks, CSR5-style, column-sorted row blocks, column segments
  float trial_ms[4] = {0.f, 0.f, 0.f, 0.f};
  // FP32 copy of x for row blocks over narrow values (narrow_x = 2); nullptr: FP64 x
  float *x32_rows = nullptr;
  int launches = 0;

  // kernel timing (measurement only)
  bool timing = false;
  std::vector<hipEvent_t> tev;   // pairs: [2i] before, [2i+1] after
  int tcount = 0;

  ~Plan();
  void *dev_alloc(size_t bytes);
  int build_csr_chunks(const int *h_rp);
  int setup_vector(const int *h_rp);
  int setup_stream(const int *h_rp);
  bool merge_enabled = true;          // false: caller launches run_merge itself
  int tile_sub = 16;                  // lanes per row in the tile kernels
  bool uses_tiles(int k) const;
  // carries left for the ordered fix-up: one per tile, or one per workgroup
  // of four tiles for the CSR5-style kernel (its carries meet in LDS first)
  int carry_count() const { return kernel == SBLAS_KERNEL_CSR5 ? (tiles + 3) / 4 : tiles; }
  bool has_merge(int k) const;
  int setup_tiles();
  int setup_csr5(const std::vector<int> &h_rp);
  int finish_csr(const std::vector<int> &h_rp);
  int choose_csr(const std::vector<int> &h_rp);
  void free_buffers_from(size_t first);
  int setup_sorted(const int *h_rp);
  void drop_sorted();
  void drop_csr5(size_t mark);
  int setup_segments(int P);
  void drop_segments();
  // times the given (kernel, variant) forms on the plan's own matrix; ms[i]
  // receives the median of form i
  int trial(const std::vector<std::pair<int, int>> &forms, std::vector<float> &ms);
  int run_merge(int k, double alpha, double beta, double *Y, hipStream_t s);
  hipEvent_t *timing_slot();
  int run_spmv(double alpha, const double *x, double beta, double *y, hipStream_t s);
  int run_spmm(int k, double alpha, const double *X, double beta, double *Y, hipStream_t s);
};

}  // namespace sblas

struct sblas_plan_s {
  sblas::Plan impl;
};
