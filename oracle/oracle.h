/*
 * oracle.h -- CPU restatement of the reference's host-side algorithms for the
 * SpMV hot path.  TEST INFRASTRUCTURE ONLY: used by tests/, the smoke check and
 * the CPU-baseline leg of bench.py as the checker.  The product library never
 * links, loads or calls anything in this directory.
 *
 * Every function names the reference file:line it restates.  Parity status of
 * each function (what pins it) is listed in oracle/README.md.
 */
#ifndef SBLAS_ORACLE_H
#define SBLAS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Return codes of the partition restatements. */
#define ORA_OK 0
#define ORA_BAD_ARGS (-1)
/* The reference's own arithmetic is undefined for these inputs (int32
 * overflow of a product, or an array read outside the array). */
#define ORA_UNDEFINED (-2)

/* One chunk of a partition; field meaning follows pCSR/pCOO/pCSC
 * (include/spmv_kernel.h:59-142) with indices made global.  For CSC chunks
 * start_row/end_row hold columns and m_c holds the chunk's column count. */
typedef struct {
  int64_t start_idx;
  int64_t end_idx;
  int32_t start_row;
  int32_t end_row;
  int32_t start_flag;
  int32_t end_flag;
  int32_t m_c;
  int32_t pad_;
  int64_t nnz_c;
} ora_chunk;

/* --- searches (src/spmv_helper.cu) ------------------------------------- */
/* `alen` = number of readable entries of a[]; the reference reads a[r] with
 * r == n unconditionally in its final check; the restatement reads it only
 * when r < alen and otherwise treats it as "not equal". */
int ora_get_row_from_index(int n, const int *a, int64_t alen, int idx);
int ora_find_first_in_sorted(const int *a, int n, int64_t alen, int key);
int ora_find_last_in_sorted(const int *a, int n, int64_t alen, int key);

/* --- partitions --------------------------------------------------------- */
int ora_part_rowblock(int m, const int *rp, int C, ora_chunk *out);
int ora_part_baseline_coo(int m, int nnz, const int *row_idx, int C, ora_chunk *out);
int ora_part_baseline_csc(int n, const int *cp, int C, ora_chunk *out);
int ora_part_v1(int m, const int *rp, int nnz, int C, ora_chunk *out);
/* mapping: C entries (numa node of each chunk) or NULL for all-zero.
 * node_out: optional, one entry per numa node (max(mapping)+1 entries). */
int ora_part_v1_numa(int m, const int *rp, int nnz, int C, const int *mapping,
                     ora_chunk *out, ora_chunk *node_out);
int ora_part_v1_numa_coo(int m, int nnz, const int *row_idx, int C,
                         const int *mapping, ora_chunk *out, ora_chunk *node_out);
int ora_part_v1_numa_csc(int n, const int *cp, long long nnz, int C,
                         const int *mapping, ora_chunk *out, ora_chunk *node_out);
/* v2 task partition; returns the task count T (>0) or an error code.
 * nb is clamped as the reference clamps it with q streams. */
int ora_v2_num_tasks(int nnz, int nb, int q);
int ora_part_v2(int m, const int *rp, int nnz, int nb, int q, ora_chunk *out, int out_cap);

/* --- SpMV oracles -------------------------------------------------------- */
void ora_csr_spmv(int m, int n, int nnz, const double *alpha, const double *val,
                  const int *rp, const int *col, const double *x,
                  const double *beta, double *y);
void ora_csr_spmv_omp(int m, int n, int nnz, const double *alpha, const double *val,
                      const int *rp, const int *col, const double *x,
                      const double *beta, double *y, int nthreads);
void ora_coo_spmv(int m, int n, int nnz, const double *alpha, const double *val,
                  const int *row_idx, const int *col_idx, const double *x,
                  const double *beta, double *y);
/* CSR oracle with FP32 / BF16 (raw uint16) stored values, FP64 accumulate. */
void ora_csr_spmv_f32val(int m, const float *val, const int *rp, const int *col,
                         const double *x, double alpha, double beta, double *y);
void ora_csr_spmv_bf16val(int m, const uint16_t *val, const int *rp, const int *col,
                          const double *x, double alpha, double beta, double *y);
/* k right-hand sides, X row-major n x k, Y row-major m x k. */
void ora_csr_spmm(int m, int k, const double *val, const int *rp, const int *col,
                  const double *X, double alpha, double beta, double *Y);

/* --- the reference harness's 'g' generator (test/dspmv_test_smt.cu:176-232) */
int64_t ora_gen_g_nnz(int n);
/* Writes row-sorted COO with values 1.0; returns nnz or ORA_UNDEFINED when
 * n % 8 != 0 (the reference then writes rows >= m). */
int64_t ora_gen_g(int n, int *row_idx, int *col_idx, double *val, int64_t cap);

/* --- error metrics -------------------------------------------------------- */
/* normwise ||y - r||_2 / ||r||_2  (returns ||y-r||_2 when ||r||_2 == 0) */
double ora_err_normwise(int64_t m, const double *y, const double *r);
/* max_i |y_i - r_i| / (|alpha| * sum_j |a_ij x_j| + |beta y0_i|), rows with a
 * zero denominator compare |y_i - r_i| absolutely. */
double ora_err_maxrow(int m, const double *val, const int *rp, const int *col,
                      const double *x, double alpha, double beta,
                      const double *y0, const double *y, const double *r);

#ifdef __cplusplus
}
#endif
#endif
