/*
 * oracle.c -- plain-C restatement of the reference's host algorithms on the
 * SpMV path (searches, partitions, host SpMV oracle, the 'g' generator).
 *
 * TEST INFRASTRUCTURE ONLY.  Built by oracle/Makefile into oracle/liboracle.so
 * with -ffp-contract=off so that each multiply and add rounds separately, as
 * the reference's serial loops do.  Only tests/, the smoke check and bench.py's
 * CPU-baseline leg may load it, and only as the checker.
 *
 * Reference citations are file:line into the reference tree.
 */
#include "oracle.h"

#include <limits.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------ */
/* searches                                                                 */
/* ------------------------------------------------------------------------ */

/* get_row_from_index, src/spmv_helper.cu:18-44.  Bisection keeping
 * a[l] <= idx < a[r]; returns l, except that an exact hit on a[r] returns r. */
int ora_get_row_from_index(int n, const int *a, int64_t alen, int idx) {
  int l = 0;
  int r = n;
  while (l < r - 1) {
    int mid = l + (r - l) / 2;
    if (idx < a[mid])
      r = mid;
    else
      l = mid;
  }
  if (idx == a[l]) return l;
  if ((int64_t)r < alen && idx == a[r]) return r;
  return l;
}

/* findFirstInSorted, src/spmv_helper.cu:646-664. */
int ora_find_first_in_sorted(const int *a, int n, int64_t alen, int key) {
  int l = 0;
  int r = n;
  while (l < r - 1) {
    int mid = l + (r - l) / 2;
    if (key <= a[mid])
      r = mid;
    else
      l = mid;
  }
  if (key == a[l]) return l;
  if ((int64_t)r < alen && key == a[r]) return r;
  return l;
}

/* findLastInSorted, src/spmv_helper.cu:667-685. */
int ora_find_last_in_sorted(const int *a, int n, int64_t alen, int key) {
  int l = 0;
  int r = n;
  while (l < r - 1) {
    int mid = l + (r - l) / 2;
    if (key < a[mid])
      r = mid;
    else
      l = mid;
  }
  if (key == a[l]) return l;
  if ((int64_t)r < alen && key == a[r]) return r;
  return l;
}

/* ------------------------------------------------------------------------ */
/* helpers shared by the partition restatements                             */
/* ------------------------------------------------------------------------ */

static void set_chunk(ora_chunk *c, int64_t si, int64_t ei, int sr, int er,
                      int sf, int ef, int m_c, int64_t nnz_c) {
  memset(c, 0, sizeof(*c));
  c->start_idx = si;
  c->end_idx = ei;
  c->start_row = sr;
  c->end_row = er;
  c->start_flag = sf;
  c->end_flag = ef;
  c->m_c = m_c;
  c->nnz_c = nnz_c;
}

/* NumaContext, include/spmv_kernel.h:150-202: node count = max(mapping)+1,
 * GPUs per node, prefix workload[].  Returns the node count or -1. */
static int numa_context(const int *mapping, int C, int *num_gpus, int *workload,
                        int cap) {
  int nodes = 0;
  for (int i = 0; i < C; i++) {
    int v = mapping ? mapping[i] : 0;
    if (v < 0) return -1;
    if (v > nodes) nodes = v;
  }
  nodes += 1;
  if (nodes > cap) return -1;
  for (int k = 0; k < nodes; k++) num_gpus[k] = 0;
  for (int i = 0; i < C; i++) num_gpus[mapping ? mapping[i] : 0]++;
  workload[0] = 0;
  workload[1] = num_gpus[0];
  for (int k = 2; k < nodes + 1; k++) workload[k] = workload[k - 1] + num_gpus[k - 1];
  return nodes;
}

static int local_id(const int *mapping, int dev) {
  int node = mapping ? mapping[dev] : 0;
  int lid = 0;
  for (int i = 0; i < dev; i++)
    if ((mapping ? mapping[i] : 0) == node) lid++;
  return lid;
}

/* ------------------------------------------------------------------------ */
/* row blocks: src/dspmv_mgpu_baseline.cu:64-70 (CSR),                       */
/* src/dspmv_mgpu_baseline_coo.cu:65-73 (COO), _csc.cu:61-67 (CSC)           */
/* ------------------------------------------------------------------------ */

int ora_part_rowblock(int m, const int *rp, int C, ora_chunk *out) {
  if (m < 0 || C < 1) return ORA_BAD_ARGS;
  if ((int64_t)C * m > INT_MAX) return ORA_UNDEFINED; /* int (d+1)*m */
  for (int d = 0; d < C; d++) {
    int sr = (d * m) / C;
    int er = ((d + 1) * m) / C - 1;
    int mc = er - sr + 1;
    int64_t si = rp[sr];
    int64_t nnzc = (int64_t)rp[er + 1] - rp[sr];
    set_chunk(&out[d], si, si + nnzc - 1, sr, er, 0, 0, mc, nnzc);
  }
  return ORA_OK;
}

int ora_part_baseline_coo(int m, int nnz, const int *row_idx, int C, ora_chunk *out) {
  if (m < 0 || nnz < 1 || C < 1) return ORA_BAD_ARGS;
  if ((int64_t)C * m > INT_MAX) return ORA_UNDEFINED;
  for (int d = 0; d < C; d++) {
    int sr = (d * m) / C;
    int er = ((d + 1) * m) / C - 1;
    int si = ora_find_first_in_sorted(row_idx, nnz, nnz, sr);
    int ei = ora_find_last_in_sorted(row_idx, nnz, nnz, er);
    set_chunk(&out[d], si, ei, sr, er, 0, 0, er - sr + 1, (int64_t)ei - si + 1);
  }
  return ORA_OK;
}

int ora_part_baseline_csc(int n, const int *cp, int C, ora_chunk *out) {
  return ora_part_rowblock(n, cp, C, out); /* same form on columns */
}

/* ------------------------------------------------------------------------ */
/* v1: src/dspmv_mgpu_v1.cu:119-157                                          */
/* ------------------------------------------------------------------------ */

int ora_part_v1(int m, const int *rp, int nnz, int C, ora_chunk *out) {
  if (m < 1 || nnz < 0 || C < 1) return ORA_BAD_ARGS;
  /* int tmp2 = (i + 1) * nnz overflows once C * nnz > INT_MAX */
  if ((int64_t)C * nnz > INT_MAX) return ORA_UNDEFINED;
  for (int i = 0; i < C; i++) {
    int tmp1 = i * nnz;
    int tmp2 = (i + 1) * nnz;
    int si = (int)floor((double)tmp1 / C);
    int ei = (int)floor((double)tmp2 / C) - 1;
    int sr = ora_get_row_from_index(m, rp, (int64_t)m + 1, si);
    int sf = si > rp[sr];
    int er = ora_get_row_from_index(m, rp, (int64_t)m + 1, ei);
    if (er + 1 > m) return ORA_UNDEFINED;
    int ef = ei < rp[er + 1] - 1;
    int mc = er - sr + 1;
    if (mc < 0) return ORA_UNDEFINED;
    set_chunk(&out[i], si, ei, sr, er, sf, ef, mc, (int64_t)ei - si + 1);
  }
  return ORA_OK;
}

/* ------------------------------------------------------------------------ */
/* v1_numa: level 1 src/dspmv_mgpu_v1_numa.cu:210-255 (local rowPtr          */
/* :293-297), level 2 :432-497 with the chunk-0 / last-chunk flag overrides. */
/* ------------------------------------------------------------------------ */

#define ORA_MAX_NODES 1024

int ora_part_v1_numa(int m, const int *rp, int nnz, int C, const int *mapping,
                     ora_chunk *out, ora_chunk *node_out) {
  if (m < 1 || nnz < 1 || C < 1) return ORA_BAD_ARGS;
  int num_gpus[ORA_MAX_NODES], workload[ORA_MAX_NODES + 1];
  int nodes = numa_context(mapping, C, num_gpus, workload, ORA_MAX_NODES);
  if (nodes < 0) return ORA_BAD_ARGS;
  double tmp = (double)nnz / C;
  for (int k = 0; k < nodes; k++) {
    if (num_gpus[k] == 0) continue; /* no representative thread */
    int S = (int)floor(workload[k] * tmp);
    int E = (int)floor(workload[k + 1] * tmp) - 1;
    int SR = ora_get_row_from_index(m + 1, rp, (int64_t)m + 1, S);
    int SF = S > rp[SR];
    int ER = ora_get_row_from_index(m + 1, rp, (int64_t)m + 1, E);
    if (ER + 1 > m) return ORA_UNDEFINED;
    int EF = E < rp[ER + 1] - 1;
    int M = ER - SR + 1;
    int N1 = E - S + 1;
    if (M < 1 || N1 < 0) return ORA_UNDEFINED;
    if (node_out) set_chunk(&node_out[k], S, E, SR, ER, SF, EF, M, N1);
    /* local rowPtr of the node slice */
    int *L = (int *)malloc(sizeof(int) * ((size_t)M + 1));
    if (!L) return ORA_BAD_ARGS;
    L[0] = 0;
    L[M] = N1;
    for (int j = 1; j < M; j++) L[j] = rp[SR + j] - S;
    double t2 = (double)N1 / num_gpus[k];
    for (int d = 0; d < C; d++) {
      if ((mapping ? mapping[d] : 0) != k) continue;
      int lid = local_id(mapping, d);
      int si = (int)floor(lid * t2);
      int ei = (int)floor((lid + 1) * t2) - 1;
      int sr = ora_get_row_from_index(M, L, (int64_t)M + 1, si);
      int sf = si > L[sr];
      if (lid == 0) sf = SF;
      int er = ora_get_row_from_index(M, L, (int64_t)M + 1, ei);
      if (er + 1 > M) { free(L); return ORA_UNDEFINED; }
      int ef = ei < L[er + 1] - 1;
      if (lid + 1 == num_gpus[k]) ef = EF;
      int mc = er - sr + 1;
      if (mc < 0) { free(L); return ORA_UNDEFINED; }
      set_chunk(&out[d], (int64_t)S + si, (int64_t)S + ei, SR + sr, SR + er, sf, ef,
                mc, (int64_t)ei - si + 1);
    }
    free(L);
  }
  return ORA_OK;
}

/* ------------------------------------------------------------------------ */
/* v1_numa_coo: level 1 src/dspmv_mgpu_v1_numa_coo.cu:75-111, level 2        */
/* :211-250.  Level 2 searches the node's local row array over [0, nnz_c)    */
/* (not over the chunk), which the restatement keeps.                       */
/* ------------------------------------------------------------------------ */

int ora_part_v1_numa_coo(int m, int nnz, const int *row_idx, int C,
                         const int *mapping, ora_chunk *out, ora_chunk *node_out) {
  if (m < 1 || nnz < 1 || C < 1) return ORA_BAD_ARGS;
  int num_gpus[ORA_MAX_NODES], workload[ORA_MAX_NODES + 1];
  int nodes = numa_context(mapping, C, num_gpus, workload, ORA_MAX_NODES);
  if (nodes < 0) return ORA_BAD_ARGS;
  double tmp = (double)nnz / C;
  for (int k = 0; k < nodes; k++) {
    if (num_gpus[k] == 0) continue;
    int S = (int)floor(workload[k] * tmp);
    int E = (int)floor(workload[k + 1] * tmp) - 1;
    if (S < 0 || S >= nnz || E < 0 || E >= nnz) return ORA_UNDEFINED;
    int SR = row_idx[S];
    int ER = row_idx[E];
    int M = ER - SR + 1;
    int N1 = E - S + 1;
    if (N1 < 1) return ORA_UNDEFINED;
    int SF = S > ora_find_first_in_sorted(row_idx, nnz, nnz, SR);
    int EF = E < ora_find_last_in_sorted(row_idx, nnz, nnz, ER);
    if (node_out) set_chunk(&node_out[k], S, E, SR, ER, SF, EF, M, N1);
    int *R = (int *)malloc(sizeof(int) * (size_t)N1);
    if (!R) return ORA_BAD_ARGS;
    for (int i = 0; i < N1; i++) R[i] = row_idx[S + i] - SR;
    double t2 = (double)N1 / num_gpus[k];
    for (int d = 0; d < C; d++) {
      if ((mapping ? mapping[d] : 0) != k) continue;
      int lid = local_id(mapping, d);
      int si = (int)floor(lid * t2);
      int ei = (int)floor((lid + 1) * t2) - 1;
      if (si < 0 || si >= N1 || ei < 0 || ei >= N1) { free(R); return ORA_UNDEFINED; }
      int sr = R[si];
      int er = R[ei];
      int mc = er - sr + 1;
      int nnzc = ei - si + 1;
      int sf = si > ora_find_first_in_sorted(R, nnzc, N1, sr);
      if (lid == 0) sf = SF;
      int ef = ei < ora_find_last_in_sorted(R, nnzc, N1, er);
      if (lid + 1 == num_gpus[k]) ef = EF;
      set_chunk(&out[d], (int64_t)S + si, (int64_t)S + ei, SR + sr, SR + er, sf, ef,
                mc, nnzc);
    }
    free(R);
  }
  return ORA_OK;
}

/* ------------------------------------------------------------------------ */
/* v1_numa_csc: level 1 src/dspmv_mgpu_v1_numa_csc.cu:79-108, level 2        */
/* :220-250; columns instead of rows and no flag overrides.                 */
/* ------------------------------------------------------------------------ */

int ora_part_v1_numa_csc(int n, const int *cp, long long nnz, int C,
                         const int *mapping, ora_chunk *out, ora_chunk *node_out) {
  if (n < 1 || nnz < 1 || nnz > INT_MAX || C < 1) return ORA_BAD_ARGS;
  int num_gpus[ORA_MAX_NODES], workload[ORA_MAX_NODES + 1];
  int nodes = numa_context(mapping, C, num_gpus, workload, ORA_MAX_NODES);
  if (nodes < 0) return ORA_BAD_ARGS;
  double tmp = (double)nnz / C;
  for (int k = 0; k < nodes; k++) {
    if (num_gpus[k] == 0) continue;
    int S = (int)floor(workload[k] * tmp);
    int E = (int)floor(workload[k + 1] * tmp) - 1;
    int SC = ora_get_row_from_index(n + 1, cp, (int64_t)n + 1, S);
    int SF = S > cp[SC];
    int EC = ora_get_row_from_index(n + 1, cp, (int64_t)n + 1, E);
    if (EC + 1 > n) return ORA_UNDEFINED;
    int EF = E < cp[EC + 1] - 1;
    int N = EC - SC + 1;
    int N1 = E - S + 1;
    if (N < 1 || N1 < 0) return ORA_UNDEFINED;
    if (node_out) set_chunk(&node_out[k], S, E, SC, EC, SF, EF, N, N1);
    int *L = (int *)malloc(sizeof(int) * ((size_t)N + 1));
    if (!L) return ORA_BAD_ARGS;
    L[0] = 0;
    L[N] = N1;
    for (int j = 1; j < N; j++) L[j] = cp[SC + j] - S;
    double t2 = (double)N1 / num_gpus[k];
    for (int d = 0; d < C; d++) {
      if ((mapping ? mapping[d] : 0) != k) continue;
      int lid = local_id(mapping, d);
      int si = (int)floor(lid * t2);
      int ei = (int)floor((lid + 1) * t2) - 1;
      int sc = ora_get_row_from_index(N, L, (int64_t)N + 1, si);
      int sf = si > L[sc];
      int ec = ora_get_row_from_index(N, L, (int64_t)N + 1, ei);
      if (ec + 1 > N) { free(L); return ORA_UNDEFINED; }
      int ef = ei < L[ec + 1] - 1;
      int nc = ec - sc + 1;
      if (nc < 0) { free(L); return ORA_UNDEFINED; }
      set_chunk(&out[d], (int64_t)S + si, (int64_t)S + ei, SC + sc, SC + ec, sf, ef,
                nc, (int64_t)ei - si + 1);
    }
    free(L);
  }
  return ORA_OK;
}

/* ------------------------------------------------------------------------ */
/* v2 tasks: nb clamp src/dspmv_mgpu_v2.cu:114, tasks :318-366               */
/* ------------------------------------------------------------------------ */

static int v2_clamp_nb(int nb, int q) {
  int cap = (int)(0.8 * 16 * 1e9 / (double)(sizeof(double) + sizeof(int) + sizeof(int))) / q;
  return nb < cap ? nb : cap;
}

int ora_v2_num_tasks(int nnz, int nb, int q) {
  if (nnz < 0 || nb < 1 || q < 1) return ORA_BAD_ARGS;
  nb = v2_clamp_nb(nb, q);
  if (nb < 1) return ORA_BAD_ARGS;
  if ((int64_t)nnz + nb - 1 > INT_MAX) return ORA_UNDEFINED;
  return (nnz + nb - 1) / nb;
}

int ora_part_v2(int m, const int *rp, int nnz, int nb, int q, ora_chunk *out, int out_cap) {
  if (m < 1) return ORA_BAD_ARGS;
  int T = ora_v2_num_tasks(nnz, nb, q);
  if (T < 0) return T;
  if (T > out_cap) return ORA_BAD_ARGS;
  if ((int64_t)T * nnz > INT_MAX) return ORA_UNDEFINED; /* int (t+1)*nnz */
  for (int t = 0; t < T; t++) {
    int tmp1 = t * nnz;
    int tmp2 = (t + 1) * nnz;
    int si = (int)floor((double)(tmp1 / T));
    int ei = (int)floor((double)(tmp2 / T)) - 1;
    int sr = ora_get_row_from_index(m, rp, (int64_t)m + 1, si);
    int sf = si > rp[sr];
    int er = ora_get_row_from_index(m, rp, (int64_t)m + 1, ei);
    if (er + 1 > m) return ORA_UNDEFINED;
    int ef = ei < rp[er + 1] - 1;
    int mc = er - sr + 1;
    if (mc < 0) return ORA_UNDEFINED;
    set_chunk(&out[t], si, ei, sr, er, sf, ef, mc, (int64_t)ei - si + 1);
  }
  return T;
}

/* ------------------------------------------------------------------------ */
/* SpMV oracles                                                             */
/* ------------------------------------------------------------------------ */

/* csr_spmv_cpu, test/dspmv_test_smt.cu:39-52: per row, ascending j,
 * sum += val[j]*x[col[j]]; y = alpha*sum + beta*y. */
void ora_csr_spmv(int m, int n, int nnz, const double *alpha, const double *val,
                  const int *rp, const int *col, const double *x,
                  const double *beta, double *y) {
  (void)n;
  (void)nnz;
  for (int i = 0; i < m; i++) {
    double sum = 0.0;
    for (int j = rp[i]; j < rp[i + 1]; j++) sum += val[j] * x[col[j]];
    y[i] = (*alpha) * sum + (*beta) * y[i];
  }
}

/* Row-parallel form of the same loop: each row is summed in the same order,
 * so results are bitwise identical to ora_csr_spmv. */
void ora_csr_spmv_omp(int m, int n, int nnz, const double *alpha, const double *val,
                      const int *rp, const int *col, const double *x,
                      const double *beta, double *y, int nthreads) {
  (void)n;
  (void)nnz;
  const double a = *alpha, b = *beta;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static)
#endif
  for (int i = 0; i < m; i++) {
    double sum = 0.0;
    for (int j = rp[i]; j < rp[i + 1]; j++) sum += val[j] * x[col[j]];
    y[i] = a * sum + b * y[i];
  }
}

/* coo_spmv_cpu, test/dspmv_test_dgx1.cu:55-69: element order accumulation into
 * sum[row]; on row-sorted input this equals the CSR oracle bitwise. */
void ora_coo_spmv(int m, int n, int nnz, const double *alpha, const double *val,
                  const int *row_idx, const int *col_idx, const double *x,
                  const double *beta, double *y) {
  (void)n;
  double *sum = (double *)calloc((size_t)(m > 0 ? m : 1), sizeof(double));
  for (int i = 0; i < nnz; i++) sum[row_idx[i]] += val[i] * x[col_idx[i]];
  for (int i = 0; i < m; i++) y[i] = (*alpha) * sum[i] + (*beta) * y[i];
  free(sum);
}

/* The same per-row rule with reduced-precision stored values widened to FP64
 * before the multiply (the precision forms keep FP64 products and sums). */
void ora_csr_spmv_f32val(int m, const float *val, const int *rp, const int *col,
                         const double *x, double alpha, double beta, double *y) {
  for (int i = 0; i < m; i++) {
    double sum = 0.0;
    for (int j = rp[i]; j < rp[i + 1]; j++) sum += (double)val[j] * x[col[j]];
    y[i] = alpha * sum + beta * y[i];
  }
}

static double bf16_to_double(uint16_t h) {
  uint32_t u = (uint32_t)h << 16;
  float f;
  memcpy(&f, &u, sizeof(f));
  return (double)f;
}

void ora_csr_spmv_bf16val(int m, const uint16_t *val, const int *rp, const int *col,
                          const double *x, double alpha, double beta, double *y) {
  for (int i = 0; i < m; i++) {
    double sum = 0.0;
    for (int j = rp[i]; j < rp[i + 1]; j++) sum += bf16_to_double(val[j]) * x[col[j]];
    y[i] = alpha * sum + beta * y[i];
  }
}

/* k right-hand sides: the csr_spmv_cpu rule applied to each column of X. */
void ora_csr_spmm(int m, int k, const double *val, const int *rp, const int *col,
                  const double *X, double alpha, double beta, double *Y) {
  double *sum = (double *)malloc(sizeof(double) * (size_t)(k > 0 ? k : 1));
  for (int i = 0; i < m; i++) {
    for (int c = 0; c < k; c++) sum[c] = 0.0;
    for (int j = rp[i]; j < rp[i + 1]; j++)
      for (int c = 0; c < k; c++) sum[c] += val[j] * X[(size_t)col[j] * k + c];
    for (int c = 0; c < k; c++)
      Y[(size_t)i * k + c] = alpha * sum[c] + beta * Y[(size_t)i * k + c];
  }
  free(sum);
}

/* ------------------------------------------------------------------------ */
/* 'g' generator, test/dspmv_test_smt.cu:176-232                             */
/* ------------------------------------------------------------------------ */

int64_t ora_gen_g_nnz(int n) {
  if (n < 8 || n % 8 != 0) return ORA_UNDEFINED;
  int m = n, nb = m / 8;
  int64_t p = 0;
  for (int i = 0; i < m; i += nb) {
    double r = (i == 0) ? 0.9 : 0.01;
    for (int ii = i; ii < i + nb; ii++)
      for (int j = 0; j < n * r; j++) p++;
  }
  return p;
}

int64_t ora_gen_g(int n, int *row_idx, int *col_idx, double *val, int64_t cap) {
  int64_t total = ora_gen_g_nnz(n);
  if (total < 0) return total;
  if (total > cap) return ORA_BAD_ARGS;
  int m = n, nb = m / 8;
  int64_t p = 0;
  for (int i = 0; i < m; i += nb) {
    double r = (i == 0) ? 0.9 : 0.01;
    for (int ii = i; ii < i + nb; ii++)
      for (int j = 0; j < n * r; j++) {
        row_idx[p] = ii;
        col_idx[p] = j;
        val[p] = 1.0;
        p++;
      }
  }
  return p;
}

/* ------------------------------------------------------------------------ */
/* error metrics                                                            */
/* ------------------------------------------------------------------------ */

double ora_err_normwise(int64_t m, const double *y, const double *r) {
  long double num = 0.0L, den = 0.0L;
  for (int64_t i = 0; i < m; i++) {
    long double d = (long double)y[i] - (long double)r[i];
    num += d * d;
    den += (long double)r[i] * (long double)r[i];
  }
  if (den == 0.0L) return (double)sqrtl(num);
  return (double)sqrtl(num / den);
}

double ora_err_maxrow(int m, const double *val, const int *rp, const int *col,
                      const double *x, double alpha, double beta,
                      const double *y0, const double *y, const double *r) {
  double worst = 0.0;
  for (int i = 0; i < m; i++) {
    double s = 0.0;
    for (int j = rp[i]; j < rp[i + 1]; j++) s += fabs(val[j] * x[col[j]]);
    double den = fabs(alpha) * s + (y0 ? fabs(beta * y0[i]) : 0.0);
    double e = fabs(y[i] - r[i]);
    double q = den > 0.0 ? e / den : e;
    if (q > worst) worst = q;
  }
  return worst;
}
