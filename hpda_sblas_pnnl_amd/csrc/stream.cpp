// stream.cpp -- out-of-core CSR SpMV in the task structure of the reference's
// v2 driver (src/dspmv_mgpu_v2.cu:103-544).
//
// The reference cuts the matrix into T = ceil(nnz / nb) tasks by its v2 rule
// (src/dspmv_mgpu_v2.cu:311-404), and q workspaces per GPU take tasks from a
// queue; each task uploads its values, columns, rebased row pointers, the
// whole x and y, runs one csrmv and is gathered on the host.  Here the matrix
// stays in the caller's host memory; the device holds the row pointers, x, y
// and q task buffers, so a matrix larger than device memory runs.  Task t goes
// to stream t mod q: its columns and values are copied into that stream's
// buffer (through pinned staging), out-of-range columns are replaced by 0 and
// flagged (the kernel then reads x in bounds; the call reports
// SBLAS_ERR_ARGS), and the sub-wave CSR kernel runs the task's rows over the
// task's element window.  Copies on one
// stream overlap kernels on the others.  Rows cut by task edges leave
// partials that one merge adds in task order after every stream has
// finished, so the results equal those of the in-core plan at the same tasks.
#include <algorithm>
#include <climits>
#include <cstdint>
#include <vector>

#include "host_util.hpp"
#include "internal.hpp"
#include "partition.hpp"
#include "plan.hpp"

namespace sblas {

int stream_spmv_csr(int m, int n, int64_t nnz, const int *h_rp, const int *h_col,
                    const double *h_val, double alpha, const double *d_x, double beta,
                    double *d_y, int nb, int q, hipStream_t s, int *tasks_out) {
  if (m < 0 || n < 0 || nnz < 0 || nnz > INT_MAX - 4096 || nb < 1 || q < 1 || q > 64 || !h_rp ||
      (nnz > 0 && (!h_col || !h_val || !d_x)) || (m > 0 && !d_y)) {
    set_error(SBLAS_ERR_ARGS, "stream spmv: bad sizes, options or pointers");
    return SBLAS_ERR_ARGS;
  }
  if (h_rp[0] != 0 || h_rp[m] != nnz) {
    set_error(SBLAS_ERR_ARGS, "stream spmv: row pointers must start at 0 and end at nnz");
    return SBLAS_ERR_ARGS;
  }
  for (int r = 0; r < m; r++)
    if (h_rp[r + 1] < h_rp[r]) {
      set_error(SBLAS_ERR_ARGS, "stream spmv: row pointers decrease");
      return SBLAS_ERR_ARGS;
    }
  if (m == 0) return SBLAS_OK;
  if (nnz == 0) {  // y = beta * y
    if (tasks_out) *tasks_out = 0;
    const int st = launch_scale(m, beta, d_y, s);
    if (st != SBLAS_OK) return st;
    SBLAS_HIP_TRY(hipStreamSynchronize(s));
    return SBLAS_OK;
  }
  // tasks of the v2 rule at the reference's clamp of nb (min(nb, 8e8 / q))
  const int cap = static_cast<int>(0.8 * 16 * 1e9 / 16.0) / q;
  const int nb_eff = std::max(1, std::min(nb, cap));
  const std::vector<ExecChunk> ex =
      exec_chunks_from_bounds(m, h_rp, element_bounds(SBLAS_PART_V2, m, h_rp, nnz, 1, nb_eff));
  const int T = static_cast<int>(ex.size());
  if (tasks_out) *tasks_out = T;
  std::vector<int> mrows, mptr, mslots;
  merge_lists(ex, h_rp, mrows, mptr, mslots);
  const RowStats stats = row_stats(m, h_rp);
  const int sub = autotune_subwave(stats);
  std::vector<VectorLaunch> shape(T);
  std::vector<DevChunk> dchunks(T);
  int64_t max_task = 1;
  for (int t = 0; t < T; t++) {
    std::vector<DevChunk> one;
    shape[t] = vector_shape(std::vector<ExecChunk>(1, ex[t]), stats.mean, sub, 1, one);
    dchunks[t] = one[0];
    max_task = std::max<int64_t>(max_task, ex[t].hi - ex[t].lo);
  }
  const int nq = std::min(q, std::max(T, 1));

  DeviceArrays mem;
  int *d_rp = mem.alloc<int>(static_cast<size_t>(m) + 1);
  DevChunk *d_chunks = mem.alloc<DevChunk>(T);
  double *d_partials = mem.alloc<double>(2 * static_cast<size_t>(std::max(T, 1)));
  int *d_mrows = mem.alloc<int>(mrows.size()), *d_mptr = mem.alloc<int>(mptr.size());
  int *d_mslots = mem.alloc<int>(mslots.size());
  int *d_bad = mem.alloc<int>(1);
  std::vector<int *> d_col(nq);
  std::vector<double *> d_val(nq);
  for (int k = 0; k < nq; k++) {
    d_col[k] = mem.alloc<int>(max_task);
    d_val[k] = mem.alloc<double>(max_task);
  }
  bool ok = d_rp && d_chunks && d_partials && d_mrows && d_mptr && d_mslots && d_bad;
  for (int k = 0; k < nq; k++) ok = ok && d_col[k] && d_val[k];
  if (!ok) {
    set_error(SBLAS_ERR_ALLOC, "stream spmv: device allocation failed");
    return SBLAS_ERR_ALLOC;
  }
  auto up = [&](void *dst, const void *src, size_t bytes) {
    return bytes == 0 || hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s) == hipSuccess;
  };
  const int zero = 0;
  ok = up(d_rp, h_rp, sizeof(int) * (static_cast<size_t>(m) + 1)) &&
       up(d_chunks, dchunks.data(), sizeof(DevChunk) * T) &&
       up(d_mrows, mrows.data(), sizeof(int) * mrows.size()) &&
       up(d_mptr, mptr.data(), sizeof(int) * mptr.size()) &&
       up(d_mslots, mslots.data(), sizeof(int) * mslots.size()) && up(d_bad, &zero, sizeof(int));
  if (!ok || hipStreamSynchronize(s) != hipSuccess) {
    set_error(SBLAS_ERR_HIP, "stream spmv: upload failed");
    return SBLAS_ERR_HIP;
  }

  struct Streams {
    std::vector<hipStream_t> st;
    std::vector<hipEvent_t> ev;
    ~Streams() {
      for (hipStream_t x : st) (void)hipStreamDestroy(x);
      for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    }
  } S;
  for (int k = 0; k < nq; k++) {
    hipStream_t x = nullptr;
    hipEvent_t e = nullptr;
    if (hipStreamCreateWithFlags(&x, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
      if (x) (void)hipStreamDestroy(x);
      set_error(SBLAS_ERR_HIP, "stream spmv: stream creation failed");
      return SBLAS_ERR_HIP;
    }
    S.st.push_back(x);
    S.ev.push_back(e);
  }
  // matrix copies through the process-wide staging ring; the lease waits for
  // its transfers when it ends, before the streams are destroyed
  StagingLease ring(static_cast<size_t>(nnz) * (sizeof(int) + sizeof(double)));
  hipEvent_t start = nullptr;
  if (hipEventCreateWithFlags(&start, hipEventDisableTiming) != hipSuccess) {
    set_error(SBLAS_ERR_HIP, "stream spmv: event creation failed");
    return SBLAS_ERR_HIP;
  }
  S.ev.push_back(start);
  // the task streams start after the caller's stream (x, y ready)
  SBLAS_HIP_TRY(hipEventRecord(start, s));
  for (int k = 0; k < nq; k++) SBLAS_HIP_TRY(hipStreamWaitEvent(S.st[k], start, 0));
  for (int t = 0; t < T; t++) {
    const int k = t % nq;
    hipStream_t ts = S.st[k];
    const int64_t lo = ex[t].lo, len = ex[t].hi - ex[t].lo;
    // a buffer is reused by task t + nq on the same stream, after task t's
    // kernel; the copies go through pinned staging (host threads fill the
    // next slots while this task's transfers and the other streams' kernels run)
    int st = ring.to_device(d_col[k], h_col + lo, sizeof(int) * len, ts);
    if (st == SBLAS_OK) st = ring.to_device(d_val[k], h_val + lo, sizeof(double) * len, ts);
    if (st == SBLAS_OK) st = launch_sanitize_cols(static_cast<int>(len), d_col[k], n, d_bad, ts);
    // the kernel addresses elements by their global index: the buffers are
    // passed rebased by the task's first element
    if (st == SBLAS_OK)
      st = launch_csr_vector(SBLAS_VAL_F64, shape[t], d_chunks + t, 1, d_rp, d_col[k] - lo,
                             d_val[k] - lo, d_x, d_y, alpha, beta, d_partials + 2 * t, ts);
    if (st != SBLAS_OK) {
      for (hipStream_t x : S.st) (void)hipStreamSynchronize(x);
      return st;
    }
  }
  for (int k = 0; k < nq; k++) {
    SBLAS_HIP_TRY(hipEventRecord(S.ev[k], S.st[k]));
    SBLAS_HIP_TRY(hipStreamWaitEvent(s, S.ev[k], 0));
  }
  if (!mrows.empty()) {
    const int st = launch_merge(d_mrows, d_mptr, d_mslots, static_cast<int>(mrows.size()),
                                d_partials, 1, d_y, alpha, beta, s);
    if (st != SBLAS_OK) return st;
  }
  int bad = 0;
  SBLAS_HIP_TRY(hipMemcpyAsync(&bad, d_bad, sizeof(int), hipMemcpyDeviceToHost, s));
  SBLAS_HIP_TRY(hipStreamSynchronize(s));
  if (bad) {
    set_error(SBLAS_ERR_ARGS, "stream spmv: a column index is out of range (y is unspecified)");
    return SBLAS_ERR_ARGS;
  }
  return SBLAS_OK;
}

}  // namespace sblas

extern "C" int sblas_stream_spmv_csr(int m, int n, int64_t nnz, const int *h_row_ptr,
                                     const int *h_col, const double *h_val, double alpha,
                                     const double *d_x, double beta, double *d_y, int nb, int q,
                                     void *stream) {
  sblas::clear_error();
  return sblas::stream_spmv_csr(m, n, nnz, h_row_ptr, h_col, h_val, alpha, d_x, beta, d_y, nb, q,
                                static_cast<hipStream_t>(stream), nullptr);
}
