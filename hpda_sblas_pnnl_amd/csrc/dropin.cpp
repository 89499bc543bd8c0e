// dropin.cpp -- the reference's C entry points (include/spmv_kernel.h:211-296)
// on one MI355X.
//
// Each call is synchronous and stateless like the reference's: the caller's
// host arrays are uploaded, a plan is built with the entry point's partition
// rule at the configured chunk count (ngpu must be 1; chunks come from
// SBLAS_CHUNKS, default 1), the kernel runs on a private stream, split rows
// are merged on the device and y is copied back.  spmv_ret holds the times of
// those steps in seconds: part_time (partition + plan), comm_time (uploads and
// download), comp_time (kernels), merg_time (device merge of split rows).
//
// Semantics follow y = alpha*A*x + beta*y for every row, including the rows
// the reference leaves unwritten or scales twice (empty rows at chunk edges,
// CSC/COO merges that add beta*y per chunk); beta == 0 does not read y.
#include <chrono>
#include <cstdio>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <vector>

#include "host_util.hpp"
#include "plan.hpp"

namespace sblas {
namespace {

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

// The failure return of every entry point: NaN timings, the status left in
// sblas_last_error().  With SBLAS_EXIT_ON_ERROR=1 the process exits instead,
// as the reference's checkCudaErrors does (include/detail/cuda/helper_cuda.h:
// 583-595), after printing the error.
spmv_ret nan_ret() {
  const char *e = std::getenv("SBLAS_EXIT_ON_ERROR");
  if (e && std::atoi(e) != 0) {
    std::fprintf(stderr, "sblas error %d (%s): %s\n", last_error(),
                 sblas_error_string(last_error()), last_message());
    std::exit(EXIT_FAILURE);
  }
  const double q = std::numeric_limits<double>::quiet_NaN();
  return spmv_ret{q, q, q, q, q};
}

int env_int(const char *name, int fallback) {
  const char *e = std::getenv(name);
  return (e && *e) ? std::atoi(e) : fallback;
}

struct EventPair {
  hipEvent_t a = nullptr, b = nullptr;
  EventPair() {
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&b);
  }
  ~EventPair() {
    if (a) (void)hipEventDestroy(a);
    if (b) (void)hipEventDestroy(b);
  }
  double seconds() {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.0;
    return ms / 1000.0;
  }
};

// Core of every entry point: host arrays in, y out.  format CSR: a0 = row
// pointers (m+1), a1 = columns; COO: a0 = row indices (nnz), a1 = columns;
// CSC: a0 = column pointers (n+1), a1 = row indices.
// With sparse_x (CSR only), elements whose x[col] == 0 are dropped on the
// device before the plan is built (spMspV).
spmv_ret run_format(int format, int m, int n, int nnz, const double *alpha, const double *val,
                    const int *a0, const int *a1, const double *x, const double *beta, double *y,
                    int kernel, int part_mode, int nb, int q, double host_prep_s,
                    bool sparse_x = false) {
  spmv_ret ret{0, 0, 0, 0, 0};
  DeviceGuard guard(env_int("SBLAS_DEVICE", 0));
  if (!guard.ok) {
    set_error(SBLAS_ERR_NO_DEVICE, "no HIP device");
    return nan_ret();
  }
  hipStream_t s = nullptr;
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) {
    set_error(SBLAS_ERR_HIP, "stream creation failed");
    return nan_ret();
  }
  struct StreamGuard {
    hipStream_t s;
    ~StreamGuard() { (void)hipStreamDestroy(s); }
  } sg{s};
  DeviceArrays mem;
  const size_t a0_len = format == SBLAS_FMT_CSR   ? static_cast<size_t>(m) + 1
                        : format == SBLAS_FMT_CSC ? static_cast<size_t>(n) + 1
                                                  : static_cast<size_t>(nnz);
  int *d_a0 = mem.alloc<int>(a0_len);
  int *d_col = mem.alloc<int>(nnz);
  double *d_val = mem.alloc<double>(nnz);
  double *d_x = mem.alloc<double>(n);
  double *d_y = mem.alloc<double>(m);
  if (!d_a0 || !d_col || !d_val || !d_x || !d_y) {
    set_error(SBLAS_ERR_ALLOC, "device allocation failed");
    return nan_ret();
  }
  EventPair up, comp, merge, down;
  {
    // uploads through the process-wide pinned staging ring (about twice the
    // pageable rate); its one-time set-up happens before the timed uploads
    StagingLease ring(sizeof(int) * a0_len + (sizeof(int) + sizeof(double)) * nnz +
                      sizeof(double) * (static_cast<size_t>(n) + m));
    (void)hipEventRecord(up.a, s);
    const bool ok = ring.to_device(d_a0, a0, sizeof(int) * a0_len, s) == SBLAS_OK &&
                    ring.to_device(d_col, a1, sizeof(int) * nnz, s) == SBLAS_OK &&
                    ring.to_device(d_val, val, sizeof(double) * nnz, s) == SBLAS_OK &&
                    ring.to_device(d_x, x, sizeof(double) * n, s) == SBLAS_OK &&
                    ring.to_device(d_y, y, sizeof(double) * m, s) == SBLAS_OK;
    (void)hipEventRecord(up.b, s);
    if (!ok || hipStreamSynchronize(s) != hipSuccess) {
      if (ok) set_error(SBLAS_ERR_HIP, "upload failed");
      return nan_ret();
    }
  }
  int64_t nz = nnz;
  if (sparse_x) {
    const double t_f0 = now_s();
    int *rp2 = mem.alloc<int>(static_cast<size_t>(m) + 1), *c2 = mem.alloc<int>(nnz);
    double *v2 = mem.alloc<double>(nnz);
    if (!rp2 || !c2 || !v2) {
      set_error(SBLAS_ERR_ALLOC, "device allocation failed");
      return nan_ret();
    }
    if (csr_drop_zero_x_device(m, n, nnz, d_a0, d_col, d_val, d_x, rp2, c2, v2, &nz, s) !=
        SBLAS_OK)
      return nan_ret();
    d_a0 = rp2;
    d_col = c2;
    d_val = v2;
    host_prep_s += now_s() - t_f0;
    if (part_mode == SBLAS_PART_V2) {  // the task count follows the kept elements
      const int T = sblas::v2_num_tasks(static_cast<int>(nz), nb, q);
      q = (T > 0 && T <= 64) ? q : 1;
    }
  }
  sblas_opts o;
  sblas_opts_default(&o);
  o.kernel = kernel;
  o.part_mode = part_mode;
  o.nb = nb;
  o.streams = q;
  o.chunks = std::max(1, env_int("SBLAS_CHUNKS", 1));
  // one SpMV per call: kernel 1 runs in the row-block form, which needs no
  // format build (the column-panel form's build costs far more than its
  // kernel saves in a single product); SBLAS_VARIANT picks another form
  o.variant = env_int("SBLAS_VARIANT", 2);
  sblas_plan plan = nullptr;
  const double t_plan0 = now_s();
  int st = format == SBLAS_FMT_CSR
               ? sblas_plan_create_csr(m, n, nz, d_a0, d_col, d_val, SBLAS_VAL_F64, &o, &plan)
           : format == SBLAS_FMT_COO
               ? sblas_plan_create_coo(m, n, nz, d_a0, d_col, d_val, SBLAS_VAL_F64, &o, &plan)
               : sblas_plan_create_csc(m, n, nz, d_a0, d_col, d_val, SBLAS_VAL_F64, &o, &plan);
  if (st != SBLAS_OK) return nan_ret();
  struct PlanGuard {
    sblas_plan p;
    ~PlanGuard() { sblas_plan_destroy(p); }
  } pg{plan};
  ret.part_time = host_prep_s + (now_s() - t_plan0);
  // kernels, then the split-row merge, each between its own events
  Plan &p = plan->impl;
  const bool merged = p.has_merge(1);
  p.merge_enabled = false;
  (void)hipEventRecord(comp.a, s);
  st = p.run_spmv(*alpha, d_x, *beta, d_y, s);
  (void)hipEventRecord(comp.b, s);
  if (st == SBLAS_OK && merged) {
    (void)hipEventRecord(merge.a, s);
    st = p.run_merge(1, *alpha, *beta, d_y, s);
    (void)hipEventRecord(merge.b, s);
  }
  if (st != SBLAS_OK) return nan_ret();
  (void)hipEventRecord(down.a, s);
  if (m > 0 && hipMemcpyAsync(y, d_y, sizeof(double) * m, hipMemcpyDeviceToHost, s) != hipSuccess) {
    set_error(SBLAS_ERR_HIP, "download failed");
    return nan_ret();
  }
  (void)hipEventRecord(down.b, s);
  if (hipStreamSynchronize(s) != hipSuccess) {
    set_error(SBLAS_ERR_HIP, "kernel execution failed");
    return nan_ret();
  }
  ret.comm_time = up.seconds() + down.seconds();
  ret.comp_time = comp.seconds();
  ret.merg_time = merged ? merge.seconds() : 0.0;
  ret.numa_part_time = 0.0;
  return ret;
}

// v2 with the matrix left in host memory: tasks streamed through the device
// (stream.cpp).  comm_time = x and y transfers, comp_time = the streamed tasks
// (their copies overlap their kernels), merg_time is inside comp_time.
spmv_ret run_v2_streamed(int m, int n, int nnz, const double *alpha, const double *val,
                         const int *rp, const int *col, const double *x, const double *beta,
                         double *y, int nb, int q) {
  spmv_ret ret{0, 0, 0, 0, 0};
  DeviceGuard guard(env_int("SBLAS_DEVICE", 0));
  if (!guard.ok) {
    set_error(SBLAS_ERR_NO_DEVICE, "no HIP device");
    return nan_ret();
  }
  OwnedStream os;
  if (!os.s) {
    set_error(SBLAS_ERR_HIP, "stream creation failed");
    return nan_ret();
  }
  hipStream_t s = os.s;
  DeviceArrays mem;
  double *d_x = mem.alloc<double>(n), *d_y = mem.alloc<double>(m);
  if (!d_x || !d_y) {
    set_error(SBLAS_ERR_ALLOC, "device allocation failed");
    return nan_ret();
  }
  EventPair up, down;
  (void)hipEventRecord(up.a, s);
  bool ok = n == 0 || hipMemcpyAsync(d_x, x, sizeof(double) * n, hipMemcpyHostToDevice, s) ==
                          hipSuccess;
  ok = ok && (m == 0 || hipMemcpyAsync(d_y, y, sizeof(double) * m, hipMemcpyHostToDevice, s) ==
                            hipSuccess);
  (void)hipEventRecord(up.b, s);
  if (!ok || hipStreamSynchronize(s) != hipSuccess) {
    set_error(SBLAS_ERR_HIP, "upload failed");
    return nan_ret();
  }
  const double t0 = now_s();
  if (stream_spmv_csr(m, n, nnz, rp, col, val, *alpha, d_x, *beta, d_y, nb, q, s, nullptr) !=
      SBLAS_OK)
    return nan_ret();
  ret.comp_time = now_s() - t0;
  (void)hipEventRecord(down.a, s);
  if (m > 0 && hipMemcpyAsync(y, d_y, sizeof(double) * m, hipMemcpyDeviceToHost, s) != hipSuccess) {
    set_error(SBLAS_ERR_HIP, "download failed");
    return nan_ret();
  }
  (void)hipEventRecord(down.b, s);
  if (hipStreamSynchronize(s) != hipSuccess) {
    set_error(SBLAS_ERR_HIP, "download failed");
    return nan_ret();
  }
  ret.comm_time = up.seconds() + down.seconds();
  return ret;
}

// The v2 drop-in streams its tasks from host memory when SBLAS_V2_STREAM=1,
// or, unless SBLAS_V2_STREAM=0, when the in-core form would not fit in the
// device memory that is free.
bool v2_streams(int m, int n, int64_t nnz) {
  const char *e = std::getenv("SBLAS_V2_STREAM");
  if (e && *e) return std::atoi(e) != 0;
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return false;
  const double need = 1.1 * (12.0 * nnz + 4.0 * (m + 1.0) + 8.0 * n + 8.0 * m);
  return need > 0.85 * static_cast<double>(free_b);
}

spmv_ret run_csr(int m, int n, int nnz, const double *alpha, const double *val, const int *rp,
                 const int *col, const double *x, const double *beta, double *y, int kernel,
                 int part_mode, int nb, int q, double host_prep_s) {
  return run_format(SBLAS_FMT_CSR, m, n, nnz, alpha, val, rp, col, x, beta, y, kernel, part_mode,
                    nb, q, host_prep_s);
}

bool check_common(int m, int n, int64_t nnz, const double *alpha, const double *beta,
                  const void *val, const void *a0, const void *a1, const double *x,
                  const double *y, int ngpu) {
  clear_error();
  if (ngpu != 1) {
    set_error(SBLAS_ERR_NGPU, "ngpu must be 1 on this engine");
    return false;
  }
  if (m < 0 || n < 0 || nnz < 0 || nnz > INT_MAX - 4096 || !alpha || !beta || !a0 ||
      (nnz > 0 && (!val || !a1 || !x)) || (m > 0 && !y)) {
    set_error(SBLAS_ERR_ARGS, "invalid sizes or null pointers");
    return false;
  }
  return true;
}

bool check_kernel(int kernel) {
  if (kernel < 1 || kernel > 3) {
    set_error(SBLAS_ERR_KERNEL, "kernel must be 1, 2 or 3");
    return false;
  }
  return true;
}

bool check_csr(int m, int nnz, const int *rp) {
  if (rp[0] != 0 || rp[m] != nnz) {
    set_error(SBLAS_ERR_ARGS, "csrRowPtr must start at 0 and end at nnz");
    return false;
  }
  return true;
}

}  // namespace
}  // namespace sblas

using namespace sblas;

extern "C" {

spmv_ret spMV_mgpu_baseline(int m, int n, int nnz, double *alpha, double *csrVal,
                            int *csrRowPtr, int *csrColIndex, double *x, double *beta,
                            double *y, int ngpu) {
  if (!check_common(m, n, nnz, alpha, beta, csrVal, csrRowPtr, csrColIndex, x, y, ngpu) ||
      !check_csr(m, nnz, csrRowPtr))
    return nan_ret();
  return run_csr(m, n, nnz, alpha, csrVal, csrRowPtr, csrColIndex, x, beta, y, 1,
                 SBLAS_PART_ROWBLOCK, 0, 1, 0.0);
}

spmv_ret spMV_mgpu_v1(int m, int n, int nnz, double *alpha, double *csrVal, int *csrRowPtr,
                      int *csrColIndex, double *x, double *beta, double *y, int ngpu,
                      int kernel) {
  if (!check_common(m, n, nnz, alpha, beta, csrVal, csrRowPtr, csrColIndex, x, y, ngpu) ||
      !check_kernel(kernel) || !check_csr(m, nnz, csrRowPtr))
    return nan_ret();
  return run_csr(m, n, nnz, alpha, csrVal, csrRowPtr, csrColIndex, x, beta, y, kernel,
                 SBLAS_PART_V1, 0, 1, 0.0);
}

spmv_ret spMV_mgpu_v1_numa(int m, int n, int nnz, double *alpha, double *csrVal,
                           int *csrRowPtr, int *csrColIndex, double *x, double *beta, double *y,
                           int ngpu, int kernel, int *numa_mapping, int part_opt, int merg_opt) {
  (void)part_opt;  // host vs device rebase: no effect on results here
  (void)merg_opt;  // host vs device-assisted merge: the merge is always on the device
  if (!check_common(m, n, nnz, alpha, beta, csrVal, csrRowPtr, csrColIndex, x, y, ngpu) ||
      !check_kernel(kernel) || !check_csr(m, nnz, csrRowPtr))
    return nan_ret();
  if (numa_mapping && numa_mapping[0] < 0) {
    set_error(SBLAS_ERR_ARGS, "numa_mapping entries must be >= 0");
    return nan_ret();
  }
  return run_csr(m, n, nnz, alpha, csrVal, csrRowPtr, csrColIndex, x, beta, y, kernel,
                 SBLAS_PART_V1_NUMA, 0, 1, 0.0);
}

spmv_ret spMV_mgpu_v2(int m, int n, int nnz, double *alpha, double *csrVal, int *csrRowPtr,
                      int *csrColIndex, double *x, double *beta, double *y, int ngpu,
                      int kernel, int nb, int copy_of_workspace) {
  if (!check_common(m, n, nnz, alpha, beta, csrVal, csrRowPtr, csrColIndex, x, y, ngpu) ||
      !check_kernel(kernel) || !check_csr(m, nnz, csrRowPtr))
    return nan_ret();
  if (nb < 1 || copy_of_workspace < 1) {
    set_error(SBLAS_ERR_ARGS, "nb and copy_of_workspace must be >= 1");
    return nan_ret();
  }
  if (copy_of_workspace <= 64 && v2_streams(m, n, nnz))
    return run_v2_streamed(m, n, nnz, alpha, csrVal, csrRowPtr, csrColIndex, x, beta, y, nb,
                           copy_of_workspace);
  // tasks go to copy_of_workspace streams when there are few of them;
  // a large task count runs as one fused launch (same results)
  const int T = sblas::v2_num_tasks(nnz, nb, copy_of_workspace);
  const int q = (T > 0 && T <= 64) ? copy_of_workspace : 1;
  return run_csr(m, n, nnz, alpha, csrVal, csrRowPtr, csrColIndex, x, beta, y, kernel,
                 SBLAS_PART_V2, nb, q, 0.0);
}

// x-sparse forms: entries whose x[col] == 0 are dropped first
// (src/dspmv_mgpu_v1.cu:19-74, src/dspmv_mgpu_v2.cu:35-101), here on the
// device after the upload; the filter's time counts in part_time.
spmv_ret spMspV_mgpu_v1(int m, int n, int nnz, double *alpha, double *csrVal, int *csrRowPtr,
                        int *csrColIndex, double *x, double *beta, double *y, int ngpu,
                        int kernel) {
  if (!check_common(m, n, nnz, alpha, beta, csrVal, csrRowPtr, csrColIndex, x, y, ngpu) ||
      !check_kernel(kernel) || !check_csr(m, nnz, csrRowPtr))
    return nan_ret();
  return run_format(SBLAS_FMT_CSR, m, n, nnz, alpha, csrVal, csrRowPtr, csrColIndex, x, beta, y,
                    kernel, SBLAS_PART_V1, 0, 1, 0.0, true);
}

spmv_ret spMspV_mgpu_v1_numa(int m, int n, int nnz, double *alpha, double *csrVal,
                             int *csrRowPtr, int *csrColIndex, double *x, double *beta,
                             double *y, int ngpu, int kernel) {
  (void)m; (void)n; (void)nnz; (void)alpha; (void)csrVal; (void)csrRowPtr; (void)csrColIndex;
  (void)x; (void)beta; (void)y; (void)ngpu; (void)kernel;
  clear_error();
  set_error(SBLAS_ERR_UNSUPPORTED,
            "spMspV_mgpu_v1_numa has no definition in the reference (include/spmv_kernel.h:282)");
  return nan_ret();
}

spmv_ret spMspV_mgpu_v2(int m, int n, int nnz, double *alpha, double *csrVal, int *csrRowPtr,
                        int *csrColIndex, double *x, double *beta, double *y, int ngpu,
                        int kernel, int nb, int copy_of_workspace) {
  if (!check_common(m, n, nnz, alpha, beta, csrVal, csrRowPtr, csrColIndex, x, y, ngpu) ||
      !check_kernel(kernel) || !check_csr(m, nnz, csrRowPtr))
    return nan_ret();
  if (nb < 1 || copy_of_workspace < 1) {
    set_error(SBLAS_ERR_ARGS, "nb and copy_of_workspace must be >= 1");
    return nan_ret();
  }
  return run_format(SBLAS_FMT_CSR, m, n, nnz, alpha, csrVal, csrRowPtr, csrColIndex, x, beta, y,
                    kernel, SBLAS_PART_V2, nb, copy_of_workspace, 0.0, true);
}

spmv_ret spMV_mgpu_baseline_coo(int m, int n, int nnz, double *alpha, double *cooVal,
                                int *cooRowIdx, int *cooColIdx, double *x, double *beta,
                                double *y, int ngpu) {
  if (!check_common(m, n, nnz, alpha, beta, cooVal, cooRowIdx, cooColIdx, x, y, ngpu))
    return nan_ret();
  // as the reference: row pointers from the sorted row indices (on the
  // device), then kernel 1 (src/dspmv_mgpu_baseline_coo.cu:146-165)
  return run_format(SBLAS_FMT_COO, m, n, nnz, alpha, cooVal, cooRowIdx, cooColIdx, x, beta, y,
                    SBLAS_KERNEL_VECTOR, SBLAS_PART_ROWBLOCK, 0, 1, 0.0);
}

spmv_ret spMV_mgpu_v1_numa_coo(int m, int n, int nnz, double *alpha, double *cooVal,
                               int *cooRowIdx, int *cooColIdx, double *x, double *beta,
                               double *y, int ngpu, int kernel, int *numa_mapping, int part_opt,
                               int merg_opt) {
  (void)numa_mapping; (void)part_opt; (void)merg_opt;
  if (!check_common(m, n, nnz, alpha, beta, cooVal, cooRowIdx, cooColIdx, x, y, ngpu) ||
      !check_kernel(kernel))
    return nan_ret();
  // row pointers from the sorted row indices (on the device), then the
  // kernel asked for (the reference runs kernel 1 whatever kernel says,
  // src/dspmv_mgpu_v1_numa_coo.cu:472)
  return run_format(SBLAS_FMT_COO, m, n, nnz, alpha, cooVal, cooRowIdx, cooColIdx, x, beta, y,
                    kernel, SBLAS_PART_V1_NUMA, 0, 1, 0.0);
}

spmv_ret spMV_mgpu_baseline_csc(int m, int n, int nnz, double *alpha, double *cscVal,
                                int *cscColPtr, int *cscRowIdx, double *x, double *beta,
                                double *y, int ngpu) {
  if (!check_common(m, n, nnz, alpha, beta, cscVal, cscColPtr, cscRowIdx, x, y, ngpu))
    return nan_ret();
  // stable device transpose to row-major form, then the CSR kernels
  return run_format(SBLAS_FMT_CSC, m, n, nnz, alpha, cscVal, cscColPtr, cscRowIdx, x, beta, y,
                    SBLAS_KERNEL_AUTO, SBLAS_PART_ROWBLOCK, 0, 1, 0.0);
}

spmv_ret spMV_mgpu_v1_numa_csc(int m, int n, long long nnz, double *alpha, double *cscVal,
                               int *cscColPtr, int *cscRowIndex, double *x, double *beta,
                               double *y, int ngpu, int kernel, int *numa_mapping, int part_opt,
                               int merg_opt) {
  (void)numa_mapping; (void)part_opt; (void)merg_opt;
  if (!check_common(m, n, nnz, alpha, beta, cscVal, cscColPtr, cscRowIndex, x, y, ngpu) ||
      !check_kernel(kernel))
    return nan_ret();
  return run_format(SBLAS_FMT_CSC, m, n, static_cast<int>(nnz), alpha, cscVal, cscColPtr,
                    cscRowIndex, x, beta, y, kernel, SBLAS_PART_V1_NUMA, 0, 1, 0.0);
}

// ---- searches and partitions ---------------------------------------------------

int get_row_from_index(int n, int *a, int idx) {
  return sblas::search_row_of(n, a, static_cast<int64_t>(n) + 1, idx);
}
int findFirstInSorted(int *a, int n, int key) {
  return sblas::search_first(a, n, static_cast<int64_t>(n), key);
}
int findLastInSorted(int *a, int n, int key) {
  return sblas::search_last(a, n, static_cast<int64_t>(n), key);
}

int sblas_partition_rowblock(int m, const int *row_ptr, int C, sblas_chunk_desc *out) {
  return sblas::partition_rowblock(m, row_ptr, C, out);
}
int sblas_partition_baseline_coo(int m, int nnz, const int *row_idx, int C,
                                 sblas_chunk_desc *out) {
  return sblas::partition_baseline_coo(m, nnz, row_idx, C, out);
}
int sblas_partition_baseline_csc(int n, const int *col_ptr, int C, sblas_chunk_desc *out) {
  return sblas::partition_rowblock(n, col_ptr, C, out);
}
int sblas_partition_v1(int m, const int *row_ptr, int nnz, int C, sblas_chunk_desc *out) {
  return sblas::partition_v1(m, row_ptr, nnz, C, out);
}
int sblas_partition_v1_numa(int m, const int *row_ptr, int nnz, int C, const int *numa_mapping,
                            sblas_chunk_desc *out, sblas_chunk_desc *node_out) {
  return sblas::partition_v1_numa(m, row_ptr, nnz, C, numa_mapping, out, node_out);
}
int sblas_partition_v1_numa_coo(int m, int nnz, const int *row_idx, int C,
                                const int *numa_mapping, sblas_chunk_desc *out,
                                sblas_chunk_desc *node_out) {
  return sblas::partition_v1_numa_coo(m, nnz, row_idx, C, numa_mapping, out, node_out);
}
int sblas_partition_v1_numa_csc(int n, const int *col_ptr, long long nnz, int C,
                                const int *numa_mapping, sblas_chunk_desc *out,
                                sblas_chunk_desc *node_out) {
  return sblas::partition_v1_numa_csc(n, col_ptr, nnz, C, numa_mapping, out, node_out);
}
int sblas_v2_num_tasks(int nnz, int nb, int q) { return sblas::v2_num_tasks(nnz, nb, q); }
int sblas_partition_v2(int m, const int *row_ptr, int nnz, int nb, int q, sblas_chunk_desc *out,
                       int out_cap) {
  return sblas::partition_v2(m, row_ptr, nnz, nb, q, out, out_cap);
}

}  // extern "C"
