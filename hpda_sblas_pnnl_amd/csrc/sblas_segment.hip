// sblas_segment.hip -- row blocks over column segments (kernel 1, variant 6)
// for matrices whose x gathers miss L2: the format build and the merge.
//
// Replaces the reference's kernel 1 (cusparseDcsrmv, call sites
// src/dspmv_mgpu_v1.cu:261, src/dspmv_mgpu_baseline.cu:125) on matrices such
// as R-MAT graphs, where x (32 MB on R-MAT scale 22) is far larger than an
// XCD's 4 MB L2 and a row block's gathers spread over all of it: a wave's
// gather waits for its slowest lane, and nearly every one has lanes that miss.
// The columns are cut into P segments of equal width (n/P entries of x, about
// 2 MB of FP64 each).  The elements of row r in segment s (a contiguous run,
// columns being sorted within rows) form a virtual row (s, r); the virtual
// rows that hold elements are numbered segment by segment, rows ascending, and
// the matrix is copied in that order.  The row-block kernel runs over the
// virtual rows as over any CSR matrix, writing one sum per virtual row; with
// the XCD remap each XCD works through a contiguous run of them, so through a
// few segments of x, one after another, each held in its L2 while it is read.
// A merge then forms
// y[r] = alpha * (sum over its virtual rows, segments ascending) + beta*y[r],
// each 256-row block of y reading every segment's contiguous run of sums.
//
// On R-MAT scale 22 at P = 8 (6.75 M virtual rows) L2 misses fall from 22.8 M
// to 8.2 M per product and reads from HBM from 2.9 to 1.0 GB; P = 16 (9.3 M
// virtual rows) runs faster still.
#include <hipcub/hipcub.hpp>

#include <algorithm>

#include "internal.hpp"

namespace sblas {
namespace {

// first column of segment s: ceil(s * n / P), so that column c lies in segment
// floor(c * P / n)
__host__ __device__ inline int64_t seg_begin(int s, int n, int P) {
  return (static_cast<int64_t>(s) * n + P - 1) / P;
}

// cnt[s * m + r] = elements of row r in segment s (a binary search per
// segment edge over the row's sorted columns)
__global__ void k_seg_count(int m, int n, int P, const int *__restrict__ rp,
                            const int *__restrict__ col, int *__restrict__ cnt) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= m) return;
  const int hi = rp[r + 1];
  int prev = rp[r];
  for (int s = 1; s <= P; s++) {
    int a = prev, b = hi;
    if (s < P) {
      const int64_t bound = seg_begin(s, n, P);
      while (a < b) {
        const int mid = (a + b) >> 1;
        if (col[mid] < bound) a = mid + 1; else b = mid;
      }
    } else {
      a = hi;
    }
    cnt[static_cast<int64_t>(s - 1) * m + r] = a - prev;
    prev = a;
  }
}

__global__ void k_seg_flags(int64_t count, const int *__restrict__ cnt, int *__restrict__ flag) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < count) flag[i] = cnt[i] > 0;
}

static_assert(kSegMergeRows <= 256, "row offsets in a merge block are 8-bit");

// virtual row v of every (s, r) that holds elements: its first element
// (voff, the exclusive scan of cnt in segment-major order) and its real row's
// place in its merge block
__global__ void k_seg_rows(int64_t count, int m, const int *__restrict__ cnt,
                           const int *__restrict__ vindex, const int *__restrict__ voff,
                           int *__restrict__ vrp, uint8_t *__restrict__ vrow) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < count && cnt[i] > 0) {
    vrp[vindex[i]] = voff[i];
    vrow[vindex[i]] = static_cast<uint8_t>((i % m) % kSegMergeRows);
  }
}

// element e of row r moves to its virtual row's run: the run of (s, r) starts
// at rp[r] + (elements of r in segments before s) in the original order
template <typename VT>
__global__ void k_seg_scatter(int64_t nnz, int m, int n, int P, const int *__restrict__ rp,
                              const int *__restrict__ col, const VT *__restrict__ val,
                              const int *__restrict__ cnt, const int *__restrict__ voff,
                              int *__restrict__ vcol, VT *__restrict__ vval) {
  const int64_t e = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (e >= nnz) return;
  int a = 0, b = m - 1;  // the row holding e: last r with rp[r] <= e
  while (a < b) {
    const int mid = (a + b + 1) >> 1;
    if (rp[mid] <= e) a = mid; else b = mid - 1;
  }
  const int r = a;
  const int c = col[e];
  const int s = static_cast<int>(static_cast<int64_t>(c) * P / n);
  int64_t start = rp[r];
  for (int q = 0; q < s; q++) start += cnt[static_cast<int64_t>(q) * m + r];
  const int64_t i = static_cast<int64_t>(s) * m + r;
  const int64_t pos = voff[i] + (e - start);
  vcol[pos] = c;
  vval[pos] = val[e];
}

// vtab[b * P + s] = first virtual row of segment s whose real row is at
// least b * kSegMergeRows (b <= the merge blocks): an entry of the dense
// scan, row index clamped to m
__global__ void k_seg_tab(int nb, int m, int P, const int *__restrict__ vindex,
                          int *__restrict__ vtab) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (nb + 1) * P) return;
  const int b = i / P, s = i % P;
  const int64_t r = std::min<int64_t>(static_cast<int64_t>(b) * kSegMergeRows, m);
  vtab[i] = vindex[static_cast<int64_t>(s) * m + r];
}

// The per-row merge.  Rows [b*256, b*256 + 256), segments in groups of
// kSegGroup: the group's virtual rows of these rows (per segment a contiguous
// run of yv, each real row at most once) are loaded together, as one list
// over the runs, into an LDS table [segment][row] that is zero where a row has
// no elements in the segment; then each thread adds its row's entries,
// segments ascending, to a sum that starts from 0 (adding a zero entry leaves
// the sum's bits as they are), and writes y[r] = alpha * sum + beta * y[r].
// The run bounds of all segments are loaded once, at the start, and each
// thread issues the loads of up to kSegUnroll list entries before it waits
// for any.  The kernel waits on its dependent loads (run bounds, then runs),
// so it wants many workgroups per CU: at 8 segments a group the table
// takes 16 KB and eight workgroups fit (on R-MAT scale 22 at 16 segments the
// whole product ran 0.402 ms; 16 segments a group, 32 KB: 0.429 ms; one
// barrier per segment over LDS sums: 0.417 ms; persistent workgroups loading
// the next group's runs during the current one, four per CU: 0.428-0.442 ms).
constexpr int kSegGroup = 8;
constexpr int kSegUnroll = 4;
__global__ __launch_bounds__(kSegMergeRows) void k_seg_merge(
    int m, int P, const int *__restrict__ vtab, const uint8_t *__restrict__ vrow,
    const double *__restrict__ yv, double *__restrict__ y, double alpha, double beta) {
  __shared__ double tab[kSegGroup][kSegMergeRows];
  __shared__ int run_lo[kSegmentsMax], run_len[kSegmentsMax];
  const int t = threadIdx.x, b = blockIdx.x;
  const int r = b * kSegMergeRows + t;
  const double y_old = (beta != 0.0 && r < m) ? y[r] : 0.0;
  if (t < P) {  // every segment's run bounds at once, read after the first barrier
    const int lo = vtab[b * P + t];
    run_lo[t] = lo;
    run_len[t] = vtab[(b + 1) * P + t] - lo;
  }
  double acc = 0.0;
  for (int s0 = 0; s0 < P; s0 += kSegGroup) {
    const int g = min(kSegGroup, P - s0);
#pragma unroll
    for (int k = 0; k < kSegGroup; k++) tab[k][t] = 0.0;
    __syncthreads();
    int pre[kSegGroup + 1];  // list position where each run starts
    pre[0] = 0;
#pragma unroll
    for (int k = 0; k < kSegGroup; k++) pre[k + 1] = pre[k] + (k < g ? run_len[s0 + k] : 0);
    const int total = pre[kSegGroup];
    for (int j0 = t; j0 < total; j0 += kSegUnroll * kSegMergeRows) {
      int seg[kSegUnroll], idx[kSegUnroll];
#pragma unroll
      for (int u = 0; u < kSegUnroll; u++) {
        const int j = j0 + u * kSegMergeRows;  // entries past the list load entry j0's
        const int jj = j < total ? j : j0;     // place again and store nothing
        int k = 0, base = 0;                   // static indexes keep pre in registers
#pragma unroll
        for (int q = 1; q < kSegGroup; q++)
          if (jj >= pre[q]) {
            k = q;
            base = pre[q];
          }
        seg[u] = j < total ? k : -1;
        idx[u] = run_lo[s0 + k] + (jj - base);
      }
      uint8_t row[kSegUnroll];
      double v[kSegUnroll];
#pragma unroll
      for (int u = 0; u < kSegUnroll; u++) {
        row[u] = vrow[idx[u]];
        v[u] = yv[idx[u]];
      }
#pragma unroll
      for (int u = 0; u < kSegUnroll; u++)
        if (seg[u] >= 0) tab[seg[u]][row[u]] = v[u];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kSegGroup; k++)
      if (k < g) acc += tab[k][t];
    __syncthreads();  // the table is read before the next group clears it
  }
  if (r < m) y[r] = (beta == 0.0) ? alpha * acc : alpha * acc + beta * y_old;
}

int exclusive_sum(const int *in, int *out, int64_t count, hipStream_t s) {
  if (count > INT32_MAX) {
    set_error(SBLAS_ERR_UNSUPPORTED, "column segments: too many (segment, row) pairs");
    return SBLAS_ERR_UNSUPPORTED;
  }
  size_t temp_bytes = 0;
  SBLAS_HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, temp_bytes, in, out,
                                                 static_cast<int>(count), s));
  void *temp = nullptr;
  SBLAS_HIP_TRY(hipMalloc(&temp, std::max<size_t>(temp_bytes, 16)));
  const hipError_t e =
      hipcub::DeviceScan::ExclusiveSum(temp, temp_bytes, in, out, static_cast<int>(count), s);
  const hipError_t e2 = hipStreamSynchronize(s);
  (void)hipFree(temp);
  SBLAS_HIP_TRY(e);
  SBLAS_HIP_TRY(e2);
  return SBLAS_OK;
}

}  // namespace

int seg_build(int val_type, int m, int n, int64_t nnz, const int *rp, const int *col,
              const void *val, int P, SegmentFormat &f, std::vector<int> &h_vrp) {
  f = SegmentFormat{};
  f.P = P;
  f.m = m;
  if (m <= 0 || n <= 0 || nnz <= 0 || P < 1 || P > kSegmentsMax) {
    set_error(SBLAS_ERR_ARGS, "column segments: bad sizes");
    return SBLAS_ERR_ARGS;
  }
  const int64_t pairs = static_cast<int64_t>(P) * m;
  if (pairs + 1 > INT32_MAX) {  // the scans and virtual row numbers are int
    set_error(SBLAS_ERR_UNSUPPORTED, "column segments: too many (segment, row) pairs");
    return SBLAS_ERR_UNSUPPORTED;
  }
  std::vector<void *> temps;
  auto tmp = [&](size_t bytes) -> void * {
    void *p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(bytes, 16)) != hipSuccess) return nullptr;
    temps.push_back(p);
    return p;
  };
  auto alloc = [&](size_t bytes) -> void * {
    void *p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(bytes, 16)) != hipSuccess) return nullptr;
    f.buffers.push_back(p);
    f.bytes += static_cast<int64_t>(std::max<size_t>(bytes, 16));
    return p;
  };
  auto done = [&](int st) {
    for (void *p : temps) (void)hipFree(p);
    if (st != SBLAS_OK) {
      for (void *p : f.buffers) (void)hipFree(p);
      f.buffers.clear();
      f.bytes = 0;
    }
    return st;
  };
  int *cnt = static_cast<int *>(tmp(sizeof(int) * (pairs + 1)));
  int *flag = static_cast<int *>(tmp(sizeof(int) * (pairs + 1)));
  int *vindex = static_cast<int *>(tmp(sizeof(int) * (pairs + 1)));
  int *voff = static_cast<int *>(tmp(sizeof(int) * (pairs + 1)));
  if (!cnt || !flag || !vindex || !voff) {
    set_error(SBLAS_ERR_ALLOC, "column segments: build buffers");
    return done(SBLAS_ERR_ALLOC);
  }
  hipStream_t s = nullptr;
  const dim3 rows_grid((m + 255) / 256), block(256);
  const dim3 pairs_grid(static_cast<unsigned>((pairs + 1 + 255) / 256));
  hipLaunchKernelGGL(k_seg_count, rows_grid, block, 0, s, m, n, P, rp, col, cnt);
  // one entry past the pairs closes each scan
  int zero = 0;
  if (hipMemcpy(cnt + pairs, &zero, sizeof(int), hipMemcpyHostToDevice) != hipSuccess) {
    set_error(SBLAS_ERR_HIP, "column segments: set-up copy failed");
    return done(SBLAS_ERR_HIP);
  }
  hipLaunchKernelGGL(k_seg_flags, pairs_grid, block, 0, s, pairs + 1, cnt, flag);
  int st;
  if ((st = exclusive_sum(flag, vindex, pairs + 1, s)) != SBLAS_OK ||
      (st = exclusive_sum(cnt, voff, pairs + 1, s)) != SBLAS_OK)
    return done(st);
  int M = 0;
  int64_t total = 0;
  int last = 0;
  if (hipMemcpy(&M, vindex + pairs, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(&last, voff + pairs, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) {
    set_error(SBLAS_ERR_HIP, "column segments: count download failed");
    return done(SBLAS_ERR_HIP);
  }
  total = last;
  if (total != nnz || M <= 0) {
    set_error(SBLAS_ERR_HIP, "column segments: element count mismatch");
    return done(SBLAS_ERR_HIP);
  }
  f.M = M;
  const size_t vb = val_type == SBLAS_VAL_F64 ? 8 : val_type == SBLAS_VAL_F32 ? 4 : 2;
  f.vrp = static_cast<int *>(alloc(sizeof(int) * (static_cast<size_t>(M) + 1)));
  f.vcol = static_cast<int *>(alloc(sizeof(int) * static_cast<size_t>(nnz)));
  f.vval = alloc(vb * static_cast<size_t>(nnz));
  const int nb = (m + kSegMergeRows - 1) / kSegMergeRows;
  f.vrow = static_cast<uint8_t *>(alloc(static_cast<size_t>(M)));
  f.vtab = static_cast<int *>(alloc(sizeof(int) * (static_cast<size_t>(nb) + 1) * P));
  f.yv = static_cast<double *>(alloc(sizeof(double) * static_cast<size_t>(M)));
  if (!f.vrp || !f.vcol || !f.vval || !f.vrow || !f.vtab || !f.yv) {
    set_error(SBLAS_ERR_ALLOC, "column segments: format allocation failed");
    return done(SBLAS_ERR_ALLOC);
  }
  hipLaunchKernelGGL(k_seg_rows, pairs_grid, block, 0, s, pairs, m, cnt, vindex, voff, f.vrp,
                     f.vrow);
  const int nnz_i = static_cast<int>(nnz);
  if (hipMemcpy(f.vrp + M, &nnz_i, sizeof(int), hipMemcpyHostToDevice) != hipSuccess) {
    set_error(SBLAS_ERR_HIP, "column segments: set-up copy failed");
    return done(SBLAS_ERR_HIP);
  }
  const dim3 elems_grid(static_cast<unsigned>((nnz + 255) / 256));
  if (val_type == SBLAS_VAL_F64)
    hipLaunchKernelGGL(k_seg_scatter<double>, elems_grid, block, 0, s, nnz, m, n, P, rp, col,
                       static_cast<const double *>(val), cnt, voff, f.vcol,
                       static_cast<double *>(f.vval));
  else if (val_type == SBLAS_VAL_F32)
    hipLaunchKernelGGL(k_seg_scatter<float>, elems_grid, block, 0, s, nnz, m, n, P, rp, col,
                       static_cast<const float *>(val), cnt, voff, f.vcol,
                       static_cast<float *>(f.vval));
  else
    hipLaunchKernelGGL(k_seg_scatter<uint16_t>, elems_grid, block, 0, s, nnz, m, n, P, rp, col,
                       static_cast<const uint16_t *>(val), cnt, voff, f.vcol,
                       static_cast<uint16_t *>(f.vval));
  hipLaunchKernelGGL(k_seg_tab, dim3(((nb + 1) * P + 255) / 256), block, 0, s, nb, m, P, vindex,
                     f.vtab);
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess) {
    set_error(SBLAS_ERR_HIP, "column segments: format build failed");
    return done(SBLAS_ERR_HIP);
  }
  h_vrp.resize(static_cast<size_t>(M) + 1);
  if (hipMemcpy(h_vrp.data(), f.vrp, sizeof(int) * h_vrp.size(), hipMemcpyDeviceToHost) !=
      hipSuccess) {
    set_error(SBLAS_ERR_HIP, "column segments: row pointer download failed");
    return done(SBLAS_ERR_HIP);
  }
  return done(SBLAS_OK);
}

int launch_seg_merge(const SegmentFormat &f, double alpha, double beta, double *y, hipStream_t s) {
  if (f.m <= 0) return SBLAS_OK;
  hipLaunchKernelGGL(k_seg_merge, dim3((f.m + kSegMergeRows - 1) / kSegMergeRows),
                     dim3(kSegMergeRows), 0, s, f.m, f.P, f.vtab, f.vrow, f.yv, y, alpha, beta);
  SBLAS_HIP_TRY(hipGetLastError());
  return SBLAS_OK;
}

}  // namespace sblas
