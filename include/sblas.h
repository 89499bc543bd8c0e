/*
 * sblas.h -- C ABI of the MI355X SpMV engine (libsblas.so).
 *
 * Four layers:
 *   1. Drop-in entry points with the reference's names and argument lists
 *      (include/spmv_kernel.h:211-296 of the reference).  Host pointers,
 *      synchronous, y = alpha*A*x + beta*y, ngpu must be 1.
 *   2. Partition API: the reference's chunking rules evaluated at any chunk
 *      count C, returning one descriptor per chunk.
 *   3. Plan API on device-resident arrays (the analogue of the reference's
 *      CSR5 handle, include/anonymouslib_cuda.h:11-53): analyse once, run
 *      many times on a caller stream.  Also the batched multi-matrix form.
 *   4. Matrix ingestion: the reference's COO sort / conversion helpers (on
 *      the device) and a Matrix Market reader.
 *
 * All functions are thread-safe.  Errors never exit the process: the
 * functions return a status (or NaN timings for the spmv_ret entry points)
 * and sblas_last_error() holds the thread's last status.
 */
#ifndef SBLAS_H
#define SBLAS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------------- */
/* status                                                                  */
/* ---------------------------------------------------------------------- */
enum {
  SBLAS_OK = 0,
  SBLAS_ERR_ARGS = 1,          /* invalid sizes, pointers or options          */
  SBLAS_ERR_NGPU = 2,          /* ngpu != 1                                   */
  SBLAS_ERR_KERNEL = 3,        /* kernel not in {1,2,3}                       */
  SBLAS_ERR_UNDEFINED = 4,     /* reference arithmetic undefined (overflow /  */
                               /* out-of-range read) for this partition input  */
  SBLAS_ERR_HIP = 5,           /* a HIP runtime call failed                   */
  SBLAS_ERR_ALLOC = 6,         /* host or device allocation failed            */
  SBLAS_ERR_UNSUPPORTED = 7,   /* entry point without a defined reference     */
  SBLAS_ERR_NO_DEVICE = 8      /* no HIP device visible                        */
};
int sblas_last_error(void);
const char *sblas_error_string(int status);
const char *sblas_version(void);

/* ---------------------------------------------------------------------- */
/* 1. drop-in entry points                                                 */
/* ---------------------------------------------------------------------- */

/* include/spmv_kernel.h:16-57: five timings in seconds, same field order.
 * The C++ compat header spmv_kernel.h defines the same struct with the
 * reference's helper methods (it sets SBLAS_SPMV_RET_DEFINED first). */
#ifndef SBLAS_SPMV_RET_DEFINED
typedef struct spmv_ret {
  double numa_part_time;
  double part_time;
  double comp_time;
  double comm_time;
  double merg_time;
} spmv_ret;
#endif

/* src/dspmv_mgpu_baseline.cu:14 -- row blocks */
spmv_ret spMV_mgpu_baseline(int m, int n, int nnz, double *alpha, double *csrVal,
                            int *csrRowPtr, int *csrColIndex, double *x, double *beta,
                            double *y, int ngpu);
/* src/dspmv_mgpu_baseline_csc.cu:14 -- column blocks */
spmv_ret spMV_mgpu_baseline_csc(int m, int n, int nnz, double *alpha, double *cscVal,
                                int *cscColPtr, int *cscRowIdx, double *x, double *beta,
                                double *y, int ngpu);
/* src/dspmv_mgpu_baseline_coo.cu:16 -- row blocks of row-sorted COO */
spmv_ret spMV_mgpu_baseline_coo(int m, int n, int nnz, double *alpha, double *cooVal,
                                int *cooRowIdx, int *cooColIdx, double *x, double *beta,
                                double *y, int ngpu);
/* src/dspmv_mgpu_v1.cu:76 -- nnz chunks, kernel 1 (vector CSR), 2 (merge
 * path), 3 (CSR5-style tiles) */
spmv_ret spMV_mgpu_v1(int m, int n, int nnz, double *alpha, double *csrVal,
                      int *csrRowPtr, int *csrColIndex, double *x, double *beta,
                      double *y, int ngpu, int kernel);
/* src/dspmv_mgpu_v1_numa.cu:98 -- two-level nnz chunks; kernel honoured */
spmv_ret spMV_mgpu_v1_numa(int m, int n, int nnz, double *alpha, double *csrVal,
                           int *csrRowPtr, int *csrColIndex, double *x, double *beta,
                           double *y, int ngpu, int kernel, int *numa_mapping,
                           int part_opt, int merg_opt);
/* src/dspmv_mgpu_v1_numa_csc.cu:43 */
spmv_ret spMV_mgpu_v1_numa_csc(int m, int n, long long nnz, double *alpha, double *cscVal,
                               int *cscColPtr, int *cscRowIndex, double *x, double *beta,
                               double *y, int ngpu, int kernel, int *numa_mapping,
                               int part_opt, int merg_opt);
/* src/dspmv_mgpu_v1_numa_coo.cu:35 */
spmv_ret spMV_mgpu_v1_numa_coo(int m, int n, int nnz, double *alpha, double *cooVal,
                               int *cooRowIdx, int *cooColIdx, double *x, double *beta,
                               double *y, int ngpu, int kernel, int *numa_mapping,
                               int part_opt, int merg_opt);
/* src/dspmv_mgpu_v2.cu:103 -- nb-sized tasks over copy_of_workspace streams */
spmv_ret spMV_mgpu_v2(int m, int n, int nnz, double *alpha, double *csrVal,
                      int *csrRowPtr, int *csrColIndex, double *x, double *beta,
                      double *y, int ngpu, int kernel, int nb, int copy_of_workspace);
/* src/dspmv_mgpu_v1.cu:19 -- drops entries whose x[col] == 0, then v1 */
spmv_ret spMspV_mgpu_v1(int m, int n, int nnz, double *alpha, double *csrVal,
                        int *csrRowPtr, int *csrColIndex, double *x, double *beta,
                        double *y, int ngpu, int kernel);
/* include/spmv_kernel.h:282 -- declared but never defined by the reference;
 * returns NaN timings with SBLAS_ERR_UNSUPPORTED and leaves y untouched. */
spmv_ret spMspV_mgpu_v1_numa(int m, int n, int nnz, double *alpha, double *csrVal,
                             int *csrRowPtr, int *csrColIndex, double *x, double *beta,
                             double *y, int ngpu, int kernel);
/* src/dspmv_mgpu_v2.cu:35 */
spmv_ret spMspV_mgpu_v2(int m, int n, int nnz, double *alpha, double *csrVal,
                        int *csrRowPtr, int *csrColIndex, double *x, double *beta,
                        double *y, int ngpu, int kernel, int nb, int copy_of_workspace);

/* src/spmv_helper.cu:18-44, :646-685 -- the reference's searches */
int get_row_from_index(int n, int *a, int idx);
int findFirstInSorted(int *a, int n, int key);
int findLastInSorted(int *a, int n, int key);

/* ---------------------------------------------------------------------- */
/* 2. partition API                                                        */
/* ---------------------------------------------------------------------- */

/* One chunk (pCSR / pCOO / pCSC / spmv_task fields, include/spmv_kernel.h:
 * 59-142, include/spmv_task.h:4-35) with indices made global.  For CSC
 * chunks start_row/end_row are columns and m_c is the column count. */
typedef struct sblas_chunk_desc {
  int64_t start_idx;
  int64_t end_idx;
  int32_t start_row;
  int32_t end_row;
  int32_t start_flag;
  int32_t end_flag;
  int32_t m_c;
  int32_t pad_;
  int64_t nnz_c;
} sblas_chunk_desc;

/* Host arrays.  Return SBLAS_OK, SBLAS_ERR_ARGS, or SBLAS_ERR_UNDEFINED where
 * the reference's own arithmetic is undefined (then out[] is unspecified). */
/* src/dspmv_mgpu_baseline.cu:64-70 */
int sblas_partition_rowblock(int m, const int *row_ptr, int C, sblas_chunk_desc *out);
/* src/dspmv_mgpu_baseline_coo.cu:65-73 */
int sblas_partition_baseline_coo(int m, int nnz, const int *row_idx, int C,
                                 sblas_chunk_desc *out);
/* src/dspmv_mgpu_baseline_csc.cu:61-67 */
int sblas_partition_baseline_csc(int n, const int *col_ptr, int C, sblas_chunk_desc *out);
/* src/dspmv_mgpu_v1.cu:119-157 */
int sblas_partition_v1(int m, const int *row_ptr, int nnz, int C, sblas_chunk_desc *out);
/* src/dspmv_mgpu_v1_numa.cu:210-255, :432-497.  numa_mapping: C entries or
 * NULL (one node).  node_out: optional, max(mapping)+1 entries. */
int sblas_partition_v1_numa(int m, const int *row_ptr, int nnz, int C,
                            const int *numa_mapping, sblas_chunk_desc *out,
                            sblas_chunk_desc *node_out);
/* src/dspmv_mgpu_v1_numa_coo.cu:75-111, :211-250 */
int sblas_partition_v1_numa_coo(int m, int nnz, const int *row_idx, int C,
                                const int *numa_mapping, sblas_chunk_desc *out,
                                sblas_chunk_desc *node_out);
/* src/dspmv_mgpu_v1_numa_csc.cu:79-108, :220-250 */
int sblas_partition_v1_numa_csc(int n, const int *col_ptr, long long nnz, int C,
                                const int *numa_mapping, sblas_chunk_desc *out,
                                sblas_chunk_desc *node_out);
/* src/dspmv_mgpu_v2.cu:114, :318-366.  Returns the task count T > 0 or
 * -status; out must hold T entries (T from sblas_v2_num_tasks). */
int sblas_v2_num_tasks(int nnz, int nb, int q);
int sblas_partition_v2(int m, const int *row_ptr, int nnz, int nb, int q,
                       sblas_chunk_desc *out, int out_cap);

/* ---------------------------------------------------------------------- */
/* 3. plan API (device pointers)                                           */
/* ---------------------------------------------------------------------- */

enum { SBLAS_VAL_F64 = 0, SBLAS_VAL_F32 = 1, SBLAS_VAL_BF16 = 2 };
enum { SBLAS_FMT_CSR = 0, SBLAS_FMT_COO = 1, SBLAS_FMT_CSC = 2 };
enum {
  SBLAS_KERNEL_AUTO = 0,
  SBLAS_KERNEL_VECTOR = 1,  /* vector CSR, wave64 split into sub-waves      */
  SBLAS_KERNEL_MERGE = 2,   /* merge-path CSR                               */
  SBLAS_KERNEL_CSR5 = 3,    /* CSR5-style tiles, wave64 segmented sum       */
  SBLAS_KERNEL_COO = 4      /* COO segmented reduction (COO plans)          */
};
enum {
  SBLAS_PART_V1 = 0,        /* src/dspmv_mgpu_v1.cu:119-130                 */
  SBLAS_PART_V1_NUMA = 1,   /* src/dspmv_mgpu_v1_numa.cu:210-214,:432-436    */
  SBLAS_PART_V2 = 2,        /* src/dspmv_mgpu_v2.cu:318-338 (nb per task)    */
  SBLAS_PART_ROWBLOCK = 3   /* src/dspmv_mgpu_baseline.cu:64-70             */
};

/* kernel AUTO: column panels (kernel 1) where the matrix's column structure
 * pays for them; otherwise row blocks or column-sorted row blocks (kernel 1)
 * or the CSR5-style kernel 3, whichever a timed trial on the matrix finds
 * fastest (matrices of at least 2^20 elements; smaller ones take row
 * blocks).  Trials are cached by row
 * profile; SBLAS_AUTOTUNE=0 turns them off.  sigma 0: the reference's rule
 * (include/anonymouslib_cuda.h:294-318) snapped to 8, 16 or 32.
 * chunks 0: SBLAS_CHUNKS when set, else one chunk per stream (one chunk for
 * a single launch, whose block order already gives each XCD a contiguous
 * range of rows). */
typedef struct sblas_opts {
  int kernel;      /* SBLAS_KERNEL_*; AUTO lets the autotuner choose        */
  int chunks;      /* chunk count C (0: see above)                          */
  int part_mode;   /* SBLAS_PART_*                                          */
  int nb;          /* v2 task size in nnz (part_mode V2)                    */
  int streams;     /* >1: one launch per chunk over this many streams       */
  int subwave;     /* vector kernel lanes per row (0 = from row lengths)    */
  int sigma;       /* CSR5 nnz per lane per tile (0 = autotuned)            */
  int xcd_remap;   /* 1: contiguous row ranges per XCD; 0: off; -1: auto   */
  int variant;     /* kernel 1 form: 0 auto, 1 sub-wave per row, 2 row blocks */
                   /* staged in LDS, 3 column panels with x staged in LDS,   */
                   /* 4 row blocks stored sorted by column (x read in column */
                   /* order, products put back in element order in LDS), 6  */
                   /* row blocks over column segments (each row cut at      */
                   /* column-segment edges, about 2 MB of x per segment; the */
                   /* pieces summed segment by segment, then merged per row) */
  int items;       /* row-block form: elements per thread (8 or 16; 0 auto) */
  int panel_rows;  /* column-panel form: rows per row panel (1..2048, or   */
                   /* 1..1536 in a batch; 0 auto)                           */
  int narrow_x;    /* FP32 / BF16 values, x in FP32: 0 auto (on for column  */
                   /* panels, and for row blocks of >= 2^20 elements), 1    */
                   /* off: x in FP64, each row summed in the host loop's    */
                   /* order; 2 on: x staged or gathered from an FP32 copy   */
                   /* written by each product (error within a few FP32      */
                   /* roundings per product); in a batch a narrow-x matrix  */
                   /* in column-panel form runs as a launch of its own      */
  int panel_layout;  /* column-panel form of a plan, layout inside a slice: */
                     /* 0 auto (2 where x is staged in FP32, else 1), 1    */
                     /* rows in order (one thread per row sums staged      */
                     /* products), 2 rows sorted by length in groups of 64 */
                     /* (one lane per row, padded to the group's longest); */
                     /* a batch uses 2 for narrow-x matrices, else 1       */
  int reserved[3];
} sblas_opts;

/* y = alpha*A*x + beta*y for a CSR matrix on the device, merge path (kernel
 * 2), WITHOUT ANY ALLOCATION: no plan, no hipMalloc, no workspace argument.
 * No counterpart in the reference; it is for callers that cannot allocate
 * (code under hipGraph capture, tight serving loops, harnesses that forbid
 * allocation).  The per-tile carries live in the library's own device array
 * of 262144 carries (8 MiB per GPU, held until the process ends), so
 * m + nnz must not exceed 2^29 (SBLAS_ERR_ARGS otherwise, before any device
 * work).  The same kernels and launch shape as a kernel-2 plan, so results
 * equal that plan's bit for bit.  Inputs are not validated (a plan created by
 * sblas_plan_create_csr validates them): row pointers must be non-decreasing
 * with d_row_ptr[m] == nnz and columns in [0, n).  Every call on one GPU
 * shares that GPU's array: calls that overlap on one GPU (on different
 * streams or threads) race.  Asynchronous on `stream`. */
int sblas_spmv_csr_noalloc(int m, int n, int64_t nnz, const int *d_row_ptr, const int *d_col,
                           const void *d_val, int val_type, double alpha, const double *d_x,
                           double beta, double *d_y, void *stream);

/* Fills opts with defaults (auto kernel, C=1, v1 partition, fused launch). */
void sblas_opts_default(sblas_opts *opts);

typedef struct sblas_plan_s *sblas_plan;

/* CSR: d_row_ptr[m+1], d_col[nnz], d_val[nnz] of val_type, device memory the
 * caller keeps alive and unchanged for the plan's lifetime. */
int sblas_plan_create_csr(int m, int n, int64_t nnz, const int *d_row_ptr,
                          const int *d_col, const void *d_val, int val_type,
                          const sblas_opts *opts, sblas_plan *plan);
/* COO sorted by row: d_row_idx[nnz], d_col_idx[nnz], d_val[nnz]. */
int sblas_plan_create_coo(int m, int n, int64_t nnz, const int *d_row_idx,
                          const int *d_col_idx, const void *d_val, int val_type,
                          const sblas_opts *opts, sblas_plan *plan);
/* CSC: d_col_ptr[n+1], d_row_idx[nnz], d_val[nnz]. */
int sblas_plan_create_csc(int m, int n, int64_t nnz, const int *d_col_ptr,
                          const int *d_row_idx, const void *d_val, int val_type,
                          const sblas_opts *opts, sblas_plan *plan);
/* y = alpha*A*x + beta*y on `stream` (a hipStream_t; NULL = default).
 * beta == 0 does not read y. Asynchronous. */
int sblas_plan_spmv(sblas_plan plan, double alpha, const double *d_x, double beta,
                    double *d_y, void *stream);
/* Y = alpha*A*X + beta*Y, X row-major n x k, Y row-major m x k (FP64). */
int sblas_plan_spmm(sblas_plan plan, int k, double alpha, const double *d_X,
                    double beta, double *d_Y, void *stream);
typedef struct sblas_plan_info {
  int format, val_type, kernel, chunks, subwave, sigma, xcd_remap, split_rows;
  int64_t m, n, nnz;
  int64_t workspace_bytes;
  double mean_row, cv_row;
  int64_t max_row;
  int launches;      /* kernel launches per sblas_plan_spmv                 */
  int variant;       /* kernel 1 form in use (1 sub-wave, 2 row blocks,     */
                     /* 3 column panels, 4 column-sorted row blocks, 6     */
                     /* row blocks over column segments)                   */
  int row_blocks;    /* row blocks of the row-block form                    */
  int items;         /* elements per thread of the row-block form           */
  int panel_rows;    /* column-panel form: rows per row panel               */
  int panel_slices;  /* column-panel form: (row panel, column panel) slices */
  int tuned;         /* how the kernel was chosen: 0 asked for (opts.kernel */
                     /* not AUTO), 1 by rule, 2 by a timed trial on this     */
                     /* matrix, 3 by a cached trial of the same row profile  */
  float trial_ms[3]; /* trial medians in ms: row blocks (kernel 1 variant   */
                     /* 2), CSR5-style (kernel 3), column-sorted row blocks  */
                     /* (kernel 1 variant 4); 0 without a trial              */
  int panel_layout;  /* column-panel form: 1 rows in order, 2 sorted groups */
  int narrow_x;      /* column-panel form: 1 when x is staged in FP32       */
  int64_t panel_elements; /* column-panel form: elements stored (padding    */
                          /* included)                                      */
  int segments;           /* variant 6: column segments                     */
  int virtual_rows;       /* variant 6: (segment, row) pieces with elements */
  float trial_seg_ms;     /* trial median of variant 6 (0: not tried)       */
  int reserved_info;
} sblas_plan_info;
int sblas_plan_get_info(sblas_plan plan, sblas_plan_info *info);
/* Chunk descriptors of the plan (reference rule of opts.part_mode). */
int sblas_plan_get_chunks(sblas_plan plan, sblas_chunk_desc *out, int cap);
int sblas_plan_destroy(sblas_plan plan);
/* Measurement: with timing on, every sblas_plan_spmv / _spmm records HIP
 * events around its main kernel launch (fused mode) on the call's stream;
 * sblas_plan_kernel_times waits for them, writes up to `cap` durations in
 * milliseconds, clears the record and returns how many it wrote. */
int sblas_plan_set_timing(sblas_plan plan, int on);
int sblas_plan_kernel_times(sblas_plan plan, float *ms, int cap);

/* ---------------------------------------------------------------------- */
/* batched matrices: launches over work lists, replayable as one hipGraph  */
/* ---------------------------------------------------------------------- */
/* No counterpart in the reference (its calls take one matrix).  Each item is
 * a CSR matrix on the device with its own x, y, alpha, beta; arrays must stay
 * alive and unchanged while the batch exists.  A run is a launch of its own
 * for each large FP32 / BF16 matrix in column-panel form (on the caller's
 * stream, one after another), a persistent launch over the other matrices'
 * row panels and one over their row blocks (one work item per workgroup when
 * launches of their own run beside it; each on a stream forked from the
 * caller's and joined back; narrow row-block matrices of 2^20 elements or
 * more gather x from FP32 copies written by one launch ahead of them), then
 * one merge launch.  A batch runs one call at a time. */
typedef struct sblas_batch_item {
  int m, n;
  int64_t nnz;
  const int *d_row_ptr;
  const int *d_col;
  const void *d_val;
  int val_type;
  int k;           /* right-hand sides: d_x n x k, d_y m x k, row-major;      */
                   /* 0 or 1 = one vector, else 2, 4, 8 or 16                  */
  const double *d_x;
  double *d_y;
  double alpha, beta;
} sblas_batch_item;

/* What the autotuner picked for one matrix. */
typedef struct sblas_batch_choice {
  int kernel;      /* SBLAS_KERNEL_VECTOR (kernel 1)                        */
  int subwave;     /* lanes per row                                          */
  int items;       /* elements per thread of a row block                     */
  int chunks;      /* nnz chunks the row blocks are cut from                 */
  int row_blocks;  /* work items of this matrix (row blocks or row panels)   */
  int variant;     /* 2 row blocks, 3 column panels (x staged in LDS)        */
  double mean_row, cv_row;
  int64_t max_row;
  int panel_rows;  /* column panels: rows per row panel (shorter than the    */
                   /* default when slices would overflow 16-bit offsets);    */
                   /* 0 for row blocks                                       */
  int reserved_;
} sblas_batch_choice;

typedef struct sblas_batch_s *sblas_batch;
int sblas_batch_create(int count, const sblas_batch_item *items, const sblas_opts *opts,
                       sblas_batch *batch);
/* y_i = alpha_i*A_i*x_i + beta_i*y_i for every item, launched directly. */
int sblas_batch_run(sblas_batch batch, void *stream);
/* The same work as one hipGraph (captured on first use, replayed after). */
int sblas_batch_graph_launch(sblas_batch batch, void *stream);
int sblas_batch_get_choice(sblas_batch batch, int index, sblas_batch_choice *out);
/* work items, the largest persistent grid of the work-list launches
 * (row-order row panels, row blocks), kernel launches per run */
int sblas_batch_get_shape(sblas_batch batch, int *work_items, int *grid, int *launches);
int sblas_batch_destroy(sblas_batch batch);

/* ---------------------------------------------------------------------- */
/* 4. matrix ingestion                                                     */
/* ---------------------------------------------------------------------- */
/* The reference's host helpers, same names and arguments.  Host arrays,
 * synchronous, the work done on the device.  The reference returns nothing
 * and exits on errors; these return nothing, leave sblas_last_error() ==
 * SBLAS_OK on success, and leave the arrays untouched on error (indices are
 * validated first: rows in [0, m), columns in [0, n)).  Sorts are stable:
 * elements with equal (row, column) keep their input order. */
/* src/spmv_helper.cu:277-331 -- in place, by row, then column */
void sortCOORow(int m, int n, int nnz, double *cooVal, int *cooRowIdx, int *cooColIdx);
/* src/spmv_helper.cu:333-387 -- in place, by column, then row */
void sortCOOCol(int m, int n, int nnz, double *cooVal, int *cooRowIdx, int *cooColIdx);
/* src/spmv_helper.cu:493-573 -- COO in any order -> CSR (csrRowPtr[m+1],
 * columns ascending within rows); the COO input is not modified */
void coo2csr(int m, int n, int nnz, double *cooVal, int *cooRowIdx, int *cooColIdx,
             double *csrVal, int *csrRowPtr, int *csrColIdx);
/* src/spmv_helper.cu:575-642 -- COO in any order -> CSC (cscColPtr[n+1]) */
void coo2csc(int m, int n, int nnz, double *cooVal, int *cooRowIdx, int *cooColIdx,
             double *cscVal, int *cscColPtr, int *cscRowIdx);

/* Device forms (cusparseXcoosortByRow / ByColumn with the value gather,
 * src/spmv_helper.cu:406-476, and cusparseXcoo2csr, :478-491).  Device
 * arrays; the call returns when the work on `stream` is done; scratch is
 * allocated per call.  sblas_coo_sort sorts in place by (row, column), or by
 * (column, row) with by_col != 0; d_val holds val_type values. */
int sblas_coo_sort(int m, int n, int64_t nnz, int *d_row_idx, int *d_col_idx, void *d_val,
                   int val_type, int by_col, void *stream);
/* d_ptr[count+1]: d_ptr[r] = position of the first index >= r in the sorted
 * d_sorted_idx[nnz] (row pointers of row-sorted COO, or column pointers of
 * column-sorted COO). */
int sblas_coo_ptr(int count, int64_t nnz, const int *d_sorted_idx, int *d_ptr, void *stream);

/* src/dspmv_mgpu_v2.cu:103-544 task structure, out of core: the CSR matrix
 * stays in host memory (h_*) and streams through the device in the v2 tasks
 * of nb elements (nb clamped to 8e8/q as in the reference) over q <= 64
 * streams; the device holds the row pointers, x, y (d_x, d_y) and q task
 * buffers, so the matrix may exceed device memory.  Kernel 1 on every task;
 * rows cut by task edges are merged on the device in task order.  Returns
 * when done.  An out-of-range column index gives SBLAS_ERR_ARGS (y is then
 * unspecified, the read of x stays in bounds).  spMV_mgpu_v2 takes this path
 * when SBLAS_V2_STREAM=1, or when the matrix does not fit the free device
 * memory (unless SBLAS_V2_STREAM=0). */
int sblas_stream_spmv_csr(int m, int n, int64_t nnz, const int *h_row_ptr, const int *h_col,
                          const double *h_val, double alpha, const double *d_x, double beta,
                          double *d_y, int nb, int q, void *stream);

/* spMspV's filter (src/dspmv_mgpu_v1.cu:19-74, src/dspmv_mgpu_v2.cu:35-101)
 * on the device: the CSR matrix without the elements whose d_x[col] == 0,
 * rows in order.  d_col_out / d_val_out hold at least nnz elements,
 * d_row_ptr_out m+1; *nnz_out = elements kept.  Returns when done. */
int sblas_csr_drop_zero_x(int m, int n, int64_t nnz, const int *d_row_ptr, const int *d_col,
                          const double *d_val, const double *d_x, int *d_row_ptr_out,
                          int *d_col_out, double *d_val_out, int64_t *nnz_out, void *stream);

/* include/mmio.h:254-367 and test/dspmv_test_dgx1.cu:133-199 -- a Matrix
 * Market coordinate file as COO: real, integer (stored as double) or pattern
 * (1.0) fields, 0-based indices, entries in file order; symmetric,
 * skew-symmetric and Hermitian files give their stored entries, not expanded
 * (*symmetry: 0 general, 1 symmetric, 2 skew-symmetric, 3 Hermitian; may be
 * NULL).  The arrays come from malloc: release them with sblas_free.  Complex
 * or dense (array) files: SBLAS_ERR_UNSUPPORTED; unreadable or malformed
 * files and out-of-range indices: SBLAS_ERR_ARGS. */
int sblas_mm_read_coo(const char *path, int *m, int *n, int *nnz, int **row_idx, int **col_idx,
                      double **val, int *symmetry);
void sblas_free(void *p);

#ifdef __cplusplus
}
#endif
#endif /* SBLAS_H */
