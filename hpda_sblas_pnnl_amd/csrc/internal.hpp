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
void clear_error();

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

// Row block of the LDS-staged CSR kernel: rows [row_begin, row_end) and the
// element window [lo, hi) they own (clamped at chunk edges and cut into
// pieces for rows longer than one block).  Split rows write partials to slot
// 2*block (first row) or 2*block+1 (last row).
struct RowBlock {
  int row_begin, row_end;
  int lo, hi;
};

constexpr int kStreamBlock = 256;      // threads per row-block launch block
constexpr int kStreamRowsMax = 1024;   // rows per row block (LDS y staging)

struct StreamLaunch {
  int items;        // elements per thread: block holds 256*items elements
  int subwave;      // lanes per row in the LDS reduction (1..64)
  int nblocks;
  int grid;
  int xcd_remap;
};

// Column-sorted row blocks (sblas_sorted.hip): each block's elements stored
// sorted by column with their offsets in the block's window.
constexpr int kSortedBlock = 512;   // threads per block
constexpr int kSortedItems = 16;    // elements per thread
constexpr int kSortedCap = kSortedBlock * kSortedItems;
struct SortedFormat {
  int nblocks = 0, grid = 0, subwave = 4, xcd_remap = 1;
  RowBlock *blocks = nullptr;   // device copy of hblocks
  int *scol = nullptr;          // columns, sorted within each block's window
  void *sval = nullptr;         // values in the same order
  uint16_t *spos = nullptr;     // each element's offset in its block's window
  double *partials = nullptr;   // 2 slots per block
  std::vector<RowBlock> hblocks;
  std::vector<void *> buffers;
  int64_t bytes = 0;
};
// Builds the format on the device from the caller's CSR (blocks of at most
// kSortedCap elements and kStreamRowsMax rows).
int sorted_build(int val_type, int n, const int *col, const void *val,
                 const std::vector<RowBlock> &hblocks, int subwave, int xcd_remap,
                 SortedFormat &f);
int launch_csr_sorted(int val_type, const SortedFormat &f, const int *rp, const double *x,
                      double *y, double alpha, double beta, hipStream_t s);

// x32 (FP32 / BF16 values only): x gathered from this FP32 copy instead of x.
int launch_csr_stream(int val_type, const StreamLaunch &L, const RowBlock *blocks,
                      const int *rp, const int *col, const void *val, const double *x,
                      double *y, double alpha, double beta, double *partials, hipStream_t s,
                      const float *x32 = nullptr);
// x32[i] = x[i] rounded to FP32, i < n
int launch_x_to_f32(int n, const double *x, float *x32, hipStream_t s);
// Row blocks over column segments (sblas_segment.hip): the elements of row r
// in column segment s (columns [ceil(s n/P), ceil((s+1) n/P))) form virtual
// row (s, r); virtual rows holding elements are numbered segment-major, rows
// ascending, and the matrix is copied in that order (vrp, vcol, vval).
constexpr int kSegmentsMax = 64;
constexpr int kSegMergeRows = 256;  // real rows per workgroup of the per-row merge
struct SegmentFormat {
  int P = 0;                    // column segments
  int m = 0;                    // real rows
  int M = 0;                    // virtual rows
  int *vrp = nullptr;           // [M + 1] virtual row pointers
  int *vcol = nullptr;          // [nnz] columns in virtual-row order
  void *vval = nullptr;         // [nnz] values in virtual-row order
  uint8_t *vrow = nullptr;      // [M] real row of each virtual row, less its merge
                                // block's first row (row mod kSegMergeRows)
  int *vtab = nullptr;          // [(ceil(m/kSegMergeRows) + 1) * P] first virtual row of
                                // segment s at or after row b*kSegMergeRows: [b*P + s]
  double *yv = nullptr;         // [M] one sum per virtual row
  std::vector<void *> buffers;
  int64_t bytes = 0;
};
// Builds the format on the device (synchronous); h_vrp receives the virtual
// row pointers.
int seg_build(int val_type, int m, int n, int64_t nnz, const int *rp, const int *col,
              const void *val, int P, SegmentFormat &f, std::vector<int> &h_vrp);
// y[r] = alpha * (f.yv over r's virtual rows, segments ascending) + beta*y[r]
int launch_seg_merge(const SegmentFormat &f, double alpha, double beta, double *y, hipStream_t s);
// k in {2, 4, 8, 16} right-hand sides over the same row blocks; partials hold
// 2 slots x k per block.
int launch_csr_blocks_mrhs(int val_type, int k, const StreamLaunch &L, const RowBlock *blocks,
                           const int *rp, const int *col, const void *val, const double *X,
                           double *Y, double alpha, double beta, double *partials,
                           hipStream_t s);

// Out-of-core CSR SpMV (stream.cpp): host-resident matrix streamed in the
// reference's v2 tasks (nb, q) through q device buffers; x, y on the device.
// Synchronous on s.  *tasks_out (may be null) = task count.
int stream_spmv_csr(int m, int n, int64_t nnz, const int *h_rp, const int *h_col,
                    const double *h_val, double alpha, const double *d_x, double beta,
                    double *d_y, int nb, int q, hipStream_t s, int *tasks_out);

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
// *ok = every a[i] in [lo, hi) (and non-decreasing when sorted); synchronous.
int check_indices(const int *a, int64_t count, int lo, int hi, bool sorted, hipStream_t s,
                  bool *ok);
// CSC -> CSR on the device, stable by row (sblas_csc.hip).
int csc_to_csr_device(int m, int n, int nnz, const int *cp, const int *ri, const void *val,
                      int val_type, int *rp, int *col, void *val_out, hipStream_t s);

// matrix ingestion (sblas_convert.hip).  Synchronous on s; scratch allocated
// per call.  Sorts (row, col, val) in place by (row, col) or, with by_col, by
// (col, row); equal pairs keep their input order.  Indices are validated.
int coo_sort_device(int m, int n, int64_t nnz, int *row, int *col, void *val, int val_type,
                    bool by_col, hipStream_t s);
// spMspV's filter: drops the elements whose x[col] == 0 (rows keep their
// order); outputs hold at least nnz elements; *nnz_out = elements kept.
// Synchronous on s.
int csr_drop_zero_x_device(int m, int n, int64_t nnz, const int *rp, const int *col,
                           const double *val, const double *x, int *rp_out, int *col_out,
                           double *val_out, int64_t *nnz_out, hipStream_t s);
// Asynchronous on s: column indices outside [0, n) are replaced by 0 and set
// *bad to 1 (the caller clears it), so a kernel reading x[col] stays in bounds.
int launch_sanitize_cols(int count, int *col, int n, int *bad, hipStream_t s);
// ptr[count + 1]: ptr[r] = first position of r in the sorted idx[nnz].
int ptr_from_sorted_device(int count, int64_t nnz, const int *idx, int *ptr, hipStream_t s);

// segmented tiles (sblas_tiles.hip): merge-path CSR and COO
constexpr int kTileBlock = 256;   // threads per tile block
constexpr int kTileItems = 8;     // merge items (or COO elements) per thread per tile
constexpr int kTileSize = kTileBlock * kTileItems;

int tile_subwave(double mean_row);  // 1, 4, 16 or 64 lanes per row
int launch_csr_merge_path(int val_type, int sub, int m, int64_t nnz, int tiles, const int *rp,
                          const int *col, const void *val, const double *x, double *y,
                          double alpha, double beta, TileCarry *carry, hipStream_t s);
int launch_coo_tiles(int val_type, int sub, int m, int64_t nnz, int tiles, const int *row_idx,
                     const int *col_idx, const void *val, const double *x, double *y,
                     double alpha, double beta, TileCarry *carry, hipStream_t s);
int launch_tile_fixup(int tiles, const TileCarry *carry, double *y, double alpha, double beta,
                      hipStream_t s);

// column-panel form of kernel 1 (sblas_panel.hip)
constexpr int kPanelCols = 4096;  // x panel staged in LDS (32 KB of FP64)
constexpr int kPanelRowsMax = 2048;  // rows per row panel (8 per thread)
constexpr int kPanelPad = 8;         // slices start on 8-element boundaries
constexpr int kPanelRowOrder = 0;    // column-panel slice layouts
constexpr int kPanelSorted = 1;
constexpr uint16_t kSortedPad = 0xFFFF;  // column offset of a padding element
// One row panel: rows [row0, row0 + nrow) of one execution chunk, its first
// column panel, and flags: bit 0 = the first row is cut by the chunk's
// element window (partial slot 2p), bit 1 = the last row is (slot 2p+1).
struct PanelInfo {
  int row0, nrow;
  int q_lo;
  int flags;
};
struct PanelFormat {
  int rows = 0;       // rows per row panel (at most)
  int rpt = 1;        // rows per thread: ceil(rows/256) rounded to 1, 2, 4, 6 or 8
  int npanels = 0, nslices = 0;
  double avg_panels_per_row_panel = 0, avg_slice = 0;
  PanelInfo *info = nullptr;    // per row panel
  int *slice_base = nullptr;    // first slice of each row panel (npanels+1)
  int64_t *slice_off = nullptr; // element offset of each slice (nslices+1)
  // per slice: R+1 row offsets inside the slice, 16-bit when every slice
  // holds at most 65535 elements, else 32-bit (srp32)
  uint16_t *srp = nullptr;
  uint32_t *srp32 = nullptr;
  void *pval = nullptr;         // values, slice-major, slices padded to kPanelPad
  uint16_t *pcol = nullptr;     // column offsets inside the column panel (padding: 0)
  // layout inside a slice: kPanelRowOrder (rows in order, row offsets srp) or
  // kPanelSorted (rows ordered by their element count in the slice, longest
  // first, in groups of 64: perm holds the R local rows of each slice in
  // that order, goff the groups' element offsets inside the slice, groups+1
  // per slice; group g stores element k of its i-th row at goff[g] + 64k + i,
  // rows shorter than the group's longest padded with column kSortedPad)
  int layout = 0;
  int groups = 0;
  uint16_t *perm = nullptr;
  uint32_t *goff = nullptr;
  double *partials = nullptr;   // 2 slots per row panel
  // narrow-x form (FP32 / BF16 values): FP32 copy of x written before each
  // launch, padded with zeros to whole column panels; nullptr: FP64 x
  float *x32 = nullptr;
  // BF16 values in the sorted-group layout: value (high half) and column
  // offset (low half) of each element in one word, in place of pval / pcol
  uint32_t *packed = nullptr;
  int64_t padded = 0;           // elements stored, slices padded
  std::vector<ExecChunk> hpanels;  // host copy: element window and rows of each panel
  std::vector<void *> buffers;
};
int panel_rpt(int rows);
// Rows per row panel so the panels fill whole rounds of resident workgroups.
int panel_rows_for(int val_type, int m, int device, bool narrow_x = false,
                   int layout = kPanelRowOrder, int chunks = 1);
// Builds the format on the device: each execution chunk of `ex` cut into row
// panels of at most `rows` rows (h_rp: host copy of the row pointers).  With
// force == false it stops after the per-panel column ranges and sets
// *chosen = false when the form would not help (few row panels, sparse
// slices, or rows whose x window is narrow).
// layout kPanelSorted rounds the rows per panel up to a multiple of 64.
int panel_build(int val_type, int rows, int m, int n, int64_t nnz, const int *rp, const int *col,
                const void *val, const std::vector<ExecChunk> &ex, const int *h_rp,
                bool force, int min_panels, PanelFormat &f, bool *chosen, hipStream_t s,
                int layout = kPanelRowOrder);
// x32_ready: the FP32 copy of x (narrow-x form) is already written.
int launch_panel_spmv(int val_type, const PanelFormat &f, int m, int n, int xcd_remap,
                      const double *x, double *y, double alpha, double beta, hipStream_t s,
                      bool x32_ready = false);

// batched matrices (sblas_batch.hip)
struct BatchMatrix {
  const int *rp;
  const int *col;
  const void *val;
  const double *x;
  double *y;
  double alpha, beta;
  int val_type;
  int subwave;  // lanes per row in the LDS reduction (1..64)
  int m, n;
  // column-panel form (pinfo == nullptr: the matrix runs as row blocks)
  const PanelInfo *pinfo;
  const int *slice_base;
  const int64_t *slice_off;
  const uint16_t *srp;
  const void *pval;
  const uint16_t *pcol;
  int prows;    // rows per row panel
  int k;        // right-hand sides (x, y are n x k, m x k row-major); 1 = one vector
  // row blocks over FP32 / BF16 values: x gathered from this FP32 copy,
  // written at the start of each run (null: from x)
  const float *x32;
};
// One x converted to its FP32 copy at the start of a batch run: entries
// [start, start + n) of the launch's index space.
struct BatchX32 {
  const double *x;
  float *x32;
  int64_t n;        // entries
  int first_chunk;  // its first chunk of kX32Chunk entries in the launch
};
constexpr int kX32Chunk = 2048;  // entries per workgroup of the conversion
// x32[i] = x[i] rounded to FP32 for every segment; chunk c (one workgroup)
// is chunk c - first_chunk of segment chunk_seg[c].
int launch_batch_x32(const BatchX32 *segs, const int *chunk_seg, int nchunks, hipStream_t s);
constexpr int kBatchRowBlock = 0;
constexpr int kBatchPanel = 1;
struct BatchItem {  // one row block, or one row panel, of one matrix
  int mat;
  int kind;     // kBatchRowBlock or kBatchPanel
  int row_begin, row_end;  // row block: rows; row panel: row_begin = panel index
  int lo, hi;              // row block: element window
  int pslot;               // first partial slot: 2*k slots (first row, last row)
  int pad_;
};
// Persistent launch over the items of one kind (kBatchPanel or
// kBatchRowBlock): the list is cut into kBatchQueues contiguous queues
// [qbound[q], qbound[q+1]); workgroups take items from per-queue counters
// ctr[q] (ctr[kBatchQueues] counts finished workgroups; the last one resets
// them all).  Each item's arithmetic does not depend on which workgroup runs it.
// flat: one item per workgroup instead (grid, qbound and ctr unused).
constexpr int kBatchQueues = 8;  // one per XCD
constexpr int kBatchPanelRowsMax = 1536;  // rows per row panel in a batch (6 per thread)
int launch_batch(int kind, int items_per_thread, int grid, int nitems, const BatchItem *items,
                 const BatchMatrix *mats, double *partials, const int *qbound, int *ctr,
                 hipStream_t s, bool flat = false);
// Resident workgroups per CU of the batched kernel for one item kind.
int batch_occupancy(int kind, int items_per_thread);
int launch_batch_merge(int groups, const int *gmat, const int *grow, const int *gptr,
                       const int *gslot, const BatchMatrix *mats, const double *partials,
                       hipStream_t s);

// CSR5-style tiles (sblas_csr5.hip): 64 lanes x sigma elements per tile
int launch_csr5(int val_type, int sigma, int ntiles, int64_t nnz, const uint32_t *tile_nz,
                const uint32_t *flags, const int *rows_nz, const int *colT, const void *valT,
                const double *x, double *y, double alpha, double beta, TileCarry *carry,
                hipStream_t s);
int launch_csr5_transpose(int val_type, int sigma, int64_t padded, int64_t nnz, const int *col,
                          const void *val, int *colT, void *valT, hipStream_t s);
int launch_scale_rows(int count, const int *rows, double beta, double *y, hipStream_t s);

}  // namespace sblas
