// sblas_vector.hip -- vector-CSR SpMV / SpMM for gfx950, the split-row merge
// and small helpers.
//
// Replaces the reference's kernel 1 (cusparseDcsrmv, call sites
// src/dspmv_mgpu_v1.cu:261, src/dspmv_mgpu_baseline.cu:125,
// src/dspmv_mgpu_v1_numa.cu:696) and its host merges
// (src/dspmv_mgpu_v1.cu:304-319, src/dspmv_mgpu_v2.cu:486-544).
//
// Design: a 64-lane wave is cut into sub-waves of SUBW lanes, one row per
// sub-wave at a time; SUBW follows the mean row length.  A launch block owns a
// contiguous run of rows of one chunk (rows_per_block), so consecutive blocks
// touch neighbouring parts of x; with xcd_remap the logical block order is
// re-dealt so that each XCD (blocks are dealt round-robin over the 8 XCDs)
// owns a contiguous 1/8 of the rows and keeps its own x window in its L2.
// Matrix values and column indices are streamed with non-temporal loads so
// they do not evict x from L2 / MALL.  Products and sums are FP64 for every
// stored value type.
#include <type_traits>

#include "internal.hpp"
#include "rows_device.hpp"

namespace sblas {
namespace {

__device__ __forceinline__ double load_val(const double *p, int j) {
  return __builtin_nontemporal_load(p + j);
}
__device__ __forceinline__ double load_val(const float *p, int j) {
  return static_cast<double>(__builtin_nontemporal_load(p + j));
}
__device__ __forceinline__ double load_val(const uint16_t *p, int j) {
  const uint32_t h = __builtin_nontemporal_load(p + j);
  return static_cast<double>(__uint_as_float(h << 16));
}
__device__ __forceinline__ int load_col(const int *p, int j) {
  return __builtin_nontemporal_load(p + j);
}

// Sum over the SUBW lanes of each aligned sub-wave (xor butterfly).
template <int SUBW>
__device__ __forceinline__ double subwave_sum(double v) {
#pragma unroll
  for (int off = SUBW / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, SUBW);
  return v;
}

__device__ __forceinline__ int logical_block(int total_blocks, int xcd_remap) {
  int b = blockIdx.x;
  if (xcd_remap) {
    const int per = gridDim.x >> 3;  // gridDim.x is a multiple of 8
    b = (b & 7) * per + (b >> 3);
  }
  return b < total_blocks ? b : -1;
}

__device__ __forceinline__ int chunk_of_block(const DevChunk *chunks, int nchunks, int b) {
  int lo = 0, hi = nchunks - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (chunks[mid].blk_begin <= b) lo = mid; else hi = mid - 1;
  }
  return lo;
}

template <typename VT, int SUBW, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_csr_vector(
    const DevChunk *__restrict__ chunks, int nchunks, int rows_per_block, int total_blocks,
    int xcd_remap, const int *__restrict__ rp, const int *__restrict__ col,
    const VT *__restrict__ val, const double *__restrict__ x, double *__restrict__ y,
    double alpha, double beta, double *__restrict__ partials) {
  const int b = logical_block(total_blocks, xcd_remap);
  if (b < 0) return;
  const int c = nchunks > 1 ? chunk_of_block(chunks, nchunks, b) : 0;
  const DevChunk ch = chunks[c];
  constexpr int SUBS = BLOCK / SUBW;
  const int sub = threadIdx.x / SUBW;
  const int lane = threadIdx.x % SUBW;
  const int row0 = ch.row_begin + (b - ch.blk_begin) * rows_per_block;
  const int row1 = min(row0 + rows_per_block, ch.row_end);
  for (int r = row0 + sub; r < row1; r += SUBS) {
    const int rs = rp[r], re = rp[r + 1];
    const int s = max(rs, ch.lo), e = min(re, ch.hi);
    double sum = 0.0;
    int j = s + lane;
    // two independent accumulators keep two loads in flight per lane
    double sum2 = 0.0;
    for (; j + SUBW < e; j += 2 * SUBW) {
      const double v0 = load_val(val, j);
      const int c0 = load_col(col, j);
      const double v1 = load_val(val, j + SUBW);
      const int c1 = load_col(col, j + SUBW);
      sum += v0 * x[c0];
      sum2 += v1 * x[c1];
    }
    if (j < e) sum += load_val(val, j) * x[load_col(col, j)];
    sum = subwave_sum<SUBW>(sum + sum2);
    if (lane == 0) {
      if (rs < ch.lo || re > ch.hi) {
        partials[2 * c + (r == ch.row_begin ? 0 : 1)] = sum;
      } else {
        y[r] = (beta == 0.0) ? alpha * sum : alpha * sum + beta * y[r];
      }
    }
  }
}

// k right-hand sides, X row-major (n x K), Y row-major (m x K).  Lane l of a
// sub-wave handles vector l % K and every (SUBW/K)-th element of the row, so
// K neighbouring lanes read one contiguous K*8-byte row of X.
template <typename VT, int SUBW, int K, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_csr_vector_mrhs(
    const DevChunk *__restrict__ chunks, int nchunks, int rows_per_block, int total_blocks,
    int xcd_remap, const int *__restrict__ rp, const int *__restrict__ col,
    const VT *__restrict__ val, const double *__restrict__ X, double *__restrict__ Y,
    double alpha, double beta, double *__restrict__ partials) {
  static_assert(SUBW >= K, "sub-wave narrower than K");
  const int b = logical_block(total_blocks, xcd_remap);
  if (b < 0) return;
  const int c = nchunks > 1 ? chunk_of_block(chunks, nchunks, b) : 0;
  const DevChunk ch = chunks[c];
  constexpr int SUBS = BLOCK / SUBW;
  constexpr int STEP = SUBW / K;
  const int sub = threadIdx.x / SUBW;
  const int lane = threadIdx.x % SUBW;
  const int vec = lane % K;
  const int grp = lane / K;
  const int row0 = ch.row_begin + (b - ch.blk_begin) * rows_per_block;
  const int row1 = min(row0 + rows_per_block, ch.row_end);
  for (int r = row0 + sub; r < row1; r += SUBS) {
    const int rs = rp[r], re = rp[r + 1];
    const int s = max(rs, ch.lo), e = min(re, ch.hi);
    double sum = 0.0;
    for (int j = s + grp; j < e; j += STEP) {
      const double v = load_val(val, j);
      const int cc = load_col(col, j);
      sum += v * X[static_cast<int64_t>(cc) * K + vec];
    }
#pragma unroll
    for (int off = SUBW / 2; off >= K; off >>= 1) sum += __shfl_xor(sum, off, SUBW);
    if (grp == 0) {
      if (rs < ch.lo || re > ch.hi) {
        partials[(2 * c + (r == ch.row_begin ? 0 : 1)) * K + vec] = sum;
      } else {
        double *yp = Y + static_cast<int64_t>(r) * K + vec;
        *yp = (beta == 0.0) ? alpha * sum : alpha * sum + beta * (*yp);
      }
    }
  }
}

// LDS-staged row blocks.  Phase 1 issues every global load the block needs
// that does not depend on another: ITEMS values and columns per thread
// (consecutive threads on consecutive elements, so each load instruction of a
// wave covers whole cache lines), the block's row pointers and, with
// beta != 0, its y entries; then the x gathers.  Row pointers go to LDS and y
// stays in registers, so the row sums (phase 2, SUBW lanes per row, products
// read from LDS) and the y writes (phase 3, consecutive threads on consecutive
// rows) wait on no further memory round trip.  A block that holds a piece of
// one long row sums in registers and reduces across the whole workgroup
// instead.  XT float: x gathered from its FP32 copy (narrow values: half the
// bytes per gather and per cached line of x; an FP32 value times an FP32 x is
// exact in FP64, so x's rounding is the only one added).
template <typename VT, int ITEMS, int SUBW, int BLOCK, typename XT = double>
__global__ __launch_bounds__(BLOCK) void k_csr_stream(
    const RowBlock *__restrict__ blocks, int nblocks, int xcd_remap,
    const int *__restrict__ rp, const int *__restrict__ col, const VT *__restrict__ val,
    const XT *__restrict__ x, double *__restrict__ y, double alpha, double beta,
    double *__restrict__ partials) {
  constexpr int CAP = BLOCK * ITEMS;
  constexpr int RPT = (kStreamRowsMax + BLOCK) / BLOCK;  // row pointers per thread
  constexpr int YPT = kStreamRowsMax / BLOCK;            // rows per thread in phase 3
  __shared__ double prod[CAP];
  __shared__ double ysh[kStreamRowsMax];
  __shared__ int rps[kStreamRowsMax + 1];
  __shared__ double wsum[BLOCK / 64];
  const int b = logical_block(nblocks, xcd_remap);
  if (b < 0) return;
  const RowBlock rb = blocks[b];
  const int lo = rb.lo, hi = rb.hi;
  const int t = threadIdx.x;
  const int r0 = rb.row_begin, nrows = rb.row_end - rb.row_begin;

  // phase 1: all independent loads first.  Lanes past the window re-read its
  // last element (same cache line) and their product is dropped by a select;
  // row pointer and y addresses are clamped into the block's rows, so no load
  // sits under a branch.
  double p[ITEMS];
  // the first row's bounds, the same for every thread (the long-row test
  // below must be uniform over the workgroup)
  const int rp_first = rp[r0], rp_second = rp[r0 + 1];
  int rpv[RPT];
#pragma unroll
  for (int k = 0; k < RPT; k++) rpv[k] = rp[r0 + min(t + k * BLOCK, nrows)];
  double yv[YPT];
  if (beta != 0.0) {
#pragma unroll
    for (int k = 0; k < YPT; k++) yv[k] = y[r0 + min(t + k * BLOCK, nrows - 1)];
  } else {
#pragma unroll
    for (int k = 0; k < YPT; k++) yv[k] = 0.0;
  }
  if (hi > lo) {
    double v[ITEMS];
    int c[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
      const int e = min(lo + t + k * BLOCK, hi - 1);
      v[k] = load_val(val, e);
      c[k] = load_col(col, e);
    }
    XT xv[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS; k++) xv[k] = x[c[k]];
#pragma unroll
    for (int k = 0; k < ITEMS; k++)
      p[k] = (lo + t + k * BLOCK < hi) ? v[k] * static_cast<double>(xv[k]) : 0.0;
  } else {
#pragma unroll
    for (int k = 0; k < ITEMS; k++) p[k] = 0.0;
  }

  if (nrows == 1 && (rp_first < lo || rp_second > hi)) {
    // piece of a long row: register sum, wave reduction, then across waves
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < ITEMS; k++) s += p[k];
    s = subwave_sum<64>(s);
    if ((t & 63) == 0) wsum[t >> 6] = s;
    __syncthreads();
    if (t == 0) {
      double tot = 0.0;
#pragma unroll
      for (int w = 0; w < BLOCK / 64; w++) tot += wsum[w];
      partials[2 * b] = tot;
    }
    return;
  }
#pragma unroll
  for (int k = 0; k < ITEMS; k++) {
    const int e = t + k * BLOCK;
    if (e < hi - lo) prod[e] = p[k];
  }
#pragma unroll
  for (int k = 0; k < RPT; k++)
    if (t + k * BLOCK <= nrows) rps[t + k * BLOCK] = rpv[k];
  __syncthreads();

  // phase 2: one sub-wave of SUBW lanes per row, whole waves for long rows
  rows::block_row_sums<SUBW, BLOCK>(prod, rps, nrows, lo, hi, ysh);
  __syncthreads();

  // phase 3: y for whole rows, partials for rows cut by the window
#pragma unroll
  for (int k = 0; k < YPT; k++) {
    const int i = t + k * BLOCK;
    if (i < nrows) {
      const double s = ysh[i];
      const bool edge = (i == 0 || i == nrows - 1);
      if (edge && (rps[i] < lo || rps[i + 1] > hi)) {
        partials[2 * b + (i == 0 ? 0 : 1)] = s;
      } else {
        y[r0 + i] = (beta == 0.0) ? alpha * s : alpha * s + beta * yv[k];
      }
    }
  }
}

// K right-hand sides over the row-block list (X, Y row-major, K values per
// row).  A block of several rows gives each row SUBW lanes, K of which share
// an element (one lane per vector); a block holding one row, or a piece of a
// long row, spreads that row over all BLOCK/K lane groups and reduces across
// the workgroup in a fixed order.  Rows cut by the block's element window
// leave partials (slot 2b for the block's first row, 2b+1 for its last) for
// the ordered merge, so a row is never the work of a single sub-wave however
// long it is.
// sum over j = j0, j0+step, ... < j1 of val[j] * X[col[j]*K + vec]; four
// elements' loads are issued before their gathers, so each lane keeps several
// X rows in flight instead of one dependent chain per element
template <typename VT, int K>
__device__ __forceinline__ double strided_dot(int j0, int j1, int step,
                                              const VT *__restrict__ val,
                                              const int *__restrict__ col,
                                              const double *__restrict__ X, int vec) {
  double sum = 0.0;
  int j = j0;
  for (; j + 3 * step < j1; j += 4 * step) {
    double v[4];
    int c[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      v[i] = load_val(val, j + i * step);
      c[i] = load_col(col, j + i * step);
    }
    double xv[4];
#pragma unroll
    for (int i = 0; i < 4; i++) xv[i] = X[static_cast<int64_t>(c[i]) * K + vec];
#pragma unroll
    for (int i = 0; i < 4; i++) sum += v[i] * xv[i];
  }
  for (; j < j1; j += step) sum += load_val(val, j) * X[static_cast<int64_t>(load_col(col, j)) * K + vec];
  return sum;
}

template <typename VT, int K, int SUBW, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_csr_blocks_mrhs(
    const RowBlock *__restrict__ blocks, int nblocks, int xcd_remap,
    const int *__restrict__ rp, const int *__restrict__ col, const VT *__restrict__ val,
    const double *__restrict__ X, double *__restrict__ Y, double alpha, double beta,
    double *__restrict__ partials) {
  static_assert(SUBW >= K && BLOCK % 64 == 0, "lane layout");
  __shared__ double red[BLOCK / 64][K];
  const int b = logical_block(nblocks, xcd_remap);
  if (b < 0) return;
  const RowBlock rb = blocks[b];
  const int t = threadIdx.x;
  const int vec = t % K;
  if (rb.row_end - rb.row_begin == 1) {
    constexpr int G = BLOCK / K;
    const int g = t / K;
    double sum = strided_dot<VT, K>(rb.lo + g, rb.hi, G, val, col, X, vec);
#pragma unroll
    for (int off = 32; off >= K; off >>= 1) sum += __shfl_xor(sum, off, 64);
    if ((t & 63) < K) red[t >> 6][vec] = sum;
    __syncthreads();
    if (t < K) {
      double s = 0.0;
#pragma unroll
      for (int w = 0; w < BLOCK / 64; w++) s += red[w][t];
      const int r = rb.row_begin;
      if (rp[r] < rb.lo || rp[r + 1] > rb.hi) {
        partials[static_cast<int64_t>(2 * b) * K + t] = s;
      } else {
        double *yp = Y + static_cast<int64_t>(r) * K + t;
        *yp = (beta == 0.0) ? alpha * s : alpha * s + beta * (*yp);
      }
    }
    return;
  }
  constexpr int SUBS = BLOCK / SUBW;
  constexpr int STEP = SUBW / K;
  // a row of more than 16 steps of its lane groups goes to a second pass
  // where a whole wave (64/K groups per vector) sums it: one long row no
  // longer holds its sub-wave for len/STEP dependent steps
  constexpr int LONG = 16 * STEP;
  const int sub = t / SUBW;
  const int grp = (t % SUBW) / K;
  auto finish = [&](int r, int rs, int re, double sum) {
    if (rs < rb.lo || re > rb.hi) {
      partials[static_cast<int64_t>(2 * b + (r == rb.row_begin ? 0 : 1)) * K + vec] = sum;
    } else {
      double *yp = Y + static_cast<int64_t>(r) * K + vec;
      *yp = (beta == 0.0) ? alpha * sum : alpha * sum + beta * (*yp);
    }
  };
  for (int r = rb.row_begin + sub; r < rb.row_end; r += SUBS) {
    const int rs = rp[r], re = rp[r + 1];
    const int s = max(rs, rb.lo), e = min(re, rb.hi);
    if (SUBW < 64 && e - s > LONG) continue;
    double sum = strided_dot<VT, K>(s + grp, e, STEP, val, col, X, vec);
#pragma unroll
    for (int off = SUBW / 2; off >= K; off >>= 1) sum += __shfl_xor(sum, off, SUBW);
    if (grp == 0) finish(r, rs, re, sum);
  }
  if constexpr (SUBW < 64) {
    // long rows in block order, found by a ballot (the same list in every
    // wave) and dealt out to the waves in turn
    constexpr int WAVES = BLOCK / 64;
    const int w = t >> 6, l = t & 63, g64 = l / K;
    const int nrows = rb.row_end - rb.row_begin;
    int seen = 0;
    for (int base = 0; base < nrows; base += 64) {
      bool is_long = false;
      if (base + l < nrows) {
        const int r = rb.row_begin + base + l;
        is_long = min(rp[r + 1], rb.hi) - max(rp[r], rb.lo) > LONG;
      }
      uint64_t mask = __ballot(is_long);
      while (mask) {
        const int bit = __ffsll(static_cast<unsigned long long>(mask)) - 1;
        mask &= mask - 1;
        if (seen % WAVES == w) {
          const int r = rb.row_begin + base + bit;
          const int rs = rp[r], re = rp[r + 1];
          double sum = strided_dot<VT, K>(max(rs, rb.lo) + g64, min(re, rb.hi), 64 / K, val, col,
                                          X, vec);
#pragma unroll
          for (int off = 32; off >= K; off >>= 1) sum += __shfl_xor(sum, off, 64);
          if (g64 == 0) finish(r, rs, re, sum);
        }
        seen++;
      }
    }
  }
}

// Split rows: y[row] = alpha * (partials of the row in chunk order) + beta*y.
__global__ void k_merge(const int *__restrict__ rows, const int *__restrict__ ptr,
                        const int *__restrict__ slots, int groups,
                        const double *__restrict__ partials, int k, double *__restrict__ y,
                        double alpha, double beta) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  const int g = t / k, v = t % k;
  if (g >= groups) return;
  double s = 0.0;
  for (int q = ptr[g]; q < ptr[g + 1]; q++) s += partials[slots[q] * k + v];
  double *yp = y + static_cast<int64_t>(rows[g]) * k + v;
  *yp = (beta == 0.0) ? alpha * s : alpha * s + beta * (*yp);
}

// Index validation at plan time: flag[0] = 1 when some a[i] is outside
// [lo, hi) or, with `sorted`, a[i] > a[i+1].  Every offending thread stores
// the same value.
__global__ void k_check_indices(int64_t count, const int *__restrict__ a, int lo, int hi,
                                int sorted, int *__restrict__ flag) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const int v = a[i];
  bool bad = v < lo || v >= hi;
  if (sorted && i + 1 < count) bad = bad || v > a[i + 1];
  if (bad) flag[0] = 1;
}

__global__ void k_scale(int64_t count, double beta, double *__restrict__ y) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < count) y[i] = (beta == 0.0) ? 0.0 : beta * y[i];
}

template <typename VT>
int launch_vector_typed(const VectorLaunch &L, const DevChunk *chunks, int nchunks,
                        const int *rp, const int *col, const VT *val, const double *x,
                        double *y, double alpha, double beta, double *partials,
                        hipStream_t s) {
  constexpr int B = 256;
  const dim3 grid(L.grid), block(B);
#define SBLAS_VEC_CASE(W)                                                                 \
  case W:                                                                                 \
    hipLaunchKernelGGL((k_csr_vector<VT, W, B>), grid, block, 0, s, chunks, nchunks,      \
                       L.rows_per_block, L.total_blocks, L.xcd_remap, rp, col, val, x, y, \
                       alpha, beta, partials);                                            \
    break;
  switch (L.subwave) {
    SBLAS_VEC_CASE(2)
    SBLAS_VEC_CASE(4)
    SBLAS_VEC_CASE(8)
    SBLAS_VEC_CASE(16)
    SBLAS_VEC_CASE(32)
    SBLAS_VEC_CASE(64)
    default:
      set_error(SBLAS_ERR_ARGS, "vector kernel: sub-wave must be a power of two in [2,64]");
      return SBLAS_ERR_ARGS;
  }
#undef SBLAS_VEC_CASE
  SBLAS_HIP_TRY(hipGetLastError());
  return SBLAS_OK;
}

template <typename VT, int K>
int launch_mrhs_k(const VectorLaunch &L, const DevChunk *chunks, int nchunks, const int *rp,
                  const int *col, const VT *val, const double *X, double *Y, double alpha,
                  double beta, double *partials, hipStream_t s) {
  constexpr int B = 256;
  const dim3 grid(L.grid), block(B);
  const int w = L.subwave < K ? K : L.subwave;
#define SBLAS_MRHS_CASE(W)                                                                 \
  case W:                                                                                  \
    if constexpr (W >= K)                                                                  \
      hipLaunchKernelGGL((k_csr_vector_mrhs<VT, W, K, B>), grid, block, 0, s, chunks,      \
                         nchunks, L.rows_per_block, L.total_blocks, L.xcd_remap, rp, col,  \
                         val, X, Y, alpha, beta, partials);                                \
    break;
  switch (w) {
    SBLAS_MRHS_CASE(2)
    SBLAS_MRHS_CASE(4)
    SBLAS_MRHS_CASE(8)
    SBLAS_MRHS_CASE(16)
    SBLAS_MRHS_CASE(32)
    SBLAS_MRHS_CASE(64)
    default:
      set_error(SBLAS_ERR_ARGS, "spmm kernel: bad sub-wave");
      return SBLAS_ERR_ARGS;
  }
#undef SBLAS_MRHS_CASE
  SBLAS_HIP_TRY(hipGetLastError());
  return SBLAS_OK;
}

template <typename VT>
int launch_mrhs_typed(int k, const VectorLaunch &L, const DevChunk *chunks, int nchunks,
                      const int *rp, const int *col, const VT *val, const double *X, double *Y,
                      double alpha, double beta, double *partials, hipStream_t s) {
  switch (k) {
    case 1: return launch_mrhs_k<VT, 1>(L, chunks, nchunks, rp, col, val, X, Y, alpha, beta, partials, s);
    case 2: return launch_mrhs_k<VT, 2>(L, chunks, nchunks, rp, col, val, X, Y, alpha, beta, partials, s);
    case 4: return launch_mrhs_k<VT, 4>(L, chunks, nchunks, rp, col, val, X, Y, alpha, beta, partials, s);
    case 8: return launch_mrhs_k<VT, 8>(L, chunks, nchunks, rp, col, val, X, Y, alpha, beta, partials, s);
    case 16: return launch_mrhs_k<VT, 16>(L, chunks, nchunks, rp, col, val, X, Y, alpha, beta, partials, s);
    default:
      set_error(SBLAS_ERR_ARGS, "spmm: k must be 1, 2, 4, 8 or 16");
      return SBLAS_ERR_ARGS;
  }
}

template <typename VT, int ITEMS>
int launch_stream_items(const StreamLaunch &L, const RowBlock *blocks, const int *rp,
                        const int *col, const VT *val, const double *x, double *y,
                        double alpha, double beta, double *partials, hipStream_t s,
                        const float *x32) {
  constexpr int B = kStreamBlock;
  const dim3 grid(L.grid), block(B);
#define SBLAS_STREAM_CASE(W)                                                              \
  case W:                                                                                 \
    if constexpr (!std::is_same<VT, double>::value) {                                     \
      if (x32) {                                                                          \
        hipLaunchKernelGGL((k_csr_stream<VT, ITEMS, W, B, float>), grid, block, 0, s,     \
                           blocks, L.nblocks, L.xcd_remap, rp, col, val, x32, y, alpha,   \
                           beta, partials);                                               \
        break;                                                                            \
      }                                                                                   \
    }                                                                                     \
    hipLaunchKernelGGL((k_csr_stream<VT, ITEMS, W, B>), grid, block, 0, s, blocks,        \
                       L.nblocks, L.xcd_remap, rp, col, val, x, y, alpha, beta, partials); \
    break;
  switch (L.subwave) {
    SBLAS_STREAM_CASE(1)
    SBLAS_STREAM_CASE(2)
    SBLAS_STREAM_CASE(4)
    SBLAS_STREAM_CASE(8)
    SBLAS_STREAM_CASE(16)
    SBLAS_STREAM_CASE(32)
    SBLAS_STREAM_CASE(64)
    default:
      set_error(SBLAS_ERR_ARGS, "stream kernel: sub-wave must be a power of two in [1,64]");
      return SBLAS_ERR_ARGS;
  }
#undef SBLAS_STREAM_CASE
  SBLAS_HIP_TRY(hipGetLastError());
  return SBLAS_OK;
}

template <typename VT>
int launch_stream_typed(const StreamLaunch &L, const RowBlock *blocks, const int *rp,
                        const int *col, const VT *val, const double *x, double *y,
                        double alpha, double beta, double *partials, hipStream_t s,
                        const float *x32) {
  switch (L.items) {
    case 8:
      return launch_stream_items<VT, 8>(L, blocks, rp, col, val, x, y, alpha, beta, partials, s,
                                        x32);
    case 16:
      return launch_stream_items<VT, 16>(L, blocks, rp, col, val, x, y, alpha, beta, partials,
                                         s, x32);
    default:
      set_error(SBLAS_ERR_ARGS, "stream kernel: items per thread must be 8 or 16");
      return SBLAS_ERR_ARGS;
  }
}

template <typename VT, int K>
int launch_blocks_mrhs_k(const StreamLaunch &L, const RowBlock *blocks, const int *rp,
                         const int *col, const VT *val, const double *X, double *Y,
                         double alpha, double beta, double *partials, hipStream_t s) {
  constexpr int B = kStreamBlock;
  const dim3 grid(L.grid), block(B);
  const int w = L.subwave < K ? K : L.subwave;
#define SBLAS_BMRHS_CASE(W)                                                                \
  case W:                                                                                  \
    if constexpr (W >= K)                                                                  \
      hipLaunchKernelGGL((k_csr_blocks_mrhs<VT, K, W, B>), grid, block, 0, s, blocks,      \
                         L.nblocks, L.xcd_remap, rp, col, val, X, Y, alpha, beta,          \
                         partials);                                                        \
    break;
  switch (w) {
    SBLAS_BMRHS_CASE(2)
    SBLAS_BMRHS_CASE(4)
    SBLAS_BMRHS_CASE(8)
    SBLAS_BMRHS_CASE(16)
    SBLAS_BMRHS_CASE(32)
    SBLAS_BMRHS_CASE(64)
    default:
      set_error(SBLAS_ERR_ARGS, "spmm kernel: bad sub-wave");
      return SBLAS_ERR_ARGS;
  }
#undef SBLAS_BMRHS_CASE
  SBLAS_HIP_TRY(hipGetLastError());
  return SBLAS_OK;
}

template <typename VT>
int launch_blocks_mrhs_typed(int k, const StreamLaunch &L, const RowBlock *blocks,
                             const int *rp, const int *col, const VT *val, const double *X,
                             double *Y, double alpha, double beta, double *partials,
                             hipStream_t s) {
  switch (k) {
    case 2: return launch_blocks_mrhs_k<VT, 2>(L, blocks, rp, col, val, X, Y, alpha, beta, partials, s);
    case 4: return launch_blocks_mrhs_k<VT, 4>(L, blocks, rp, col, val, X, Y, alpha, beta, partials, s);
    case 8: return launch_blocks_mrhs_k<VT, 8>(L, blocks, rp, col, val, X, Y, alpha, beta, partials, s);
    case 16: return launch_blocks_mrhs_k<VT, 16>(L, blocks, rp, col, val, X, Y, alpha, beta, partials, s);
    default:
      set_error(SBLAS_ERR_ARGS, "spmm: k must be 1, 2, 4, 8 or 16");
      return SBLAS_ERR_ARGS;
  }
}

}  // namespace

int launch_csr_blocks_mrhs(int val_type, int k, const StreamLaunch &L, const RowBlock *blocks,
                           const int *rp, const int *col, const void *val, const double *X,
                           double *Y, double alpha, double beta, double *partials,
                           hipStream_t s) {
  if (L.grid <= 0) return SBLAS_OK;
  switch (val_type) {
    case SBLAS_VAL_F64:
      return launch_blocks_mrhs_typed(k, L, blocks, rp, col, static_cast<const double *>(val), X,
                                      Y, alpha, beta, partials, s);
    case SBLAS_VAL_F32:
      return launch_blocks_mrhs_typed(k, L, blocks, rp, col, static_cast<const float *>(val), X,
                                      Y, alpha, beta, partials, s);
    case SBLAS_VAL_BF16:
      return launch_blocks_mrhs_typed(k, L, blocks, rp, col, static_cast<const uint16_t *>(val),
                                      X, Y, alpha, beta, partials, s);
    default:
      set_error(SBLAS_ERR_ARGS, "unknown value type");
      return SBLAS_ERR_ARGS;
  }
}

int launch_csr_stream(int val_type, const StreamLaunch &L, const RowBlock *blocks,
                      const int *rp, const int *col, const void *val, const double *x,
                      double *y, double alpha, double beta, double *partials, hipStream_t s,
                      const float *x32) {
  if (L.grid <= 0) return SBLAS_OK;
  switch (val_type) {
    case SBLAS_VAL_F64:
      return launch_stream_typed(L, blocks, rp, col, static_cast<const double *>(val), x, y,
                                 alpha, beta, partials, s, nullptr);
    case SBLAS_VAL_F32:
      return launch_stream_typed(L, blocks, rp, col, static_cast<const float *>(val), x, y,
                                 alpha, beta, partials, s, x32);
    case SBLAS_VAL_BF16:
      return launch_stream_typed(L, blocks, rp, col, static_cast<const uint16_t *>(val), x, y,
                                 alpha, beta, partials, s, x32);
    default:
      set_error(SBLAS_ERR_ARGS, "unknown value type");
      return SBLAS_ERR_ARGS;
  }
}

int launch_csr_vector(int val_type, const VectorLaunch &L, const DevChunk *chunks, int nchunks,
                      const int *rp, const int *col, const void *val, const double *x,
                      double *y, double alpha, double beta, double *partials,
                      hipStream_t s) {
  if (L.grid <= 0) return SBLAS_OK;
  switch (val_type) {
    case SBLAS_VAL_F64:
      return launch_vector_typed(L, chunks, nchunks, rp, col, static_cast<const double *>(val),
                                 x, y, alpha, beta, partials, s);
    case SBLAS_VAL_F32:
      return launch_vector_typed(L, chunks, nchunks, rp, col, static_cast<const float *>(val),
                                 x, y, alpha, beta, partials, s);
    case SBLAS_VAL_BF16:
      return launch_vector_typed(L, chunks, nchunks, rp, col,
                                 static_cast<const uint16_t *>(val), x, y, alpha, beta,
                                 partials, s);
    default:
      set_error(SBLAS_ERR_ARGS, "unknown value type");
      return SBLAS_ERR_ARGS;
  }
}

int launch_csr_vector_mrhs(int val_type, int k, const VectorLaunch &L, const DevChunk *chunks,
                           int nchunks, const int *rp, const int *col, const void *val,
                           const double *X, double *Y, double alpha, double beta,
                           double *partials, hipStream_t s) {
  if (L.grid <= 0) return SBLAS_OK;
  switch (val_type) {
    case SBLAS_VAL_F64:
      return launch_mrhs_typed(k, L, chunks, nchunks, rp, col, static_cast<const double *>(val),
                               X, Y, alpha, beta, partials, s);
    case SBLAS_VAL_F32:
      return launch_mrhs_typed(k, L, chunks, nchunks, rp, col, static_cast<const float *>(val),
                               X, Y, alpha, beta, partials, s);
    case SBLAS_VAL_BF16:
      return launch_mrhs_typed(k, L, chunks, nchunks, rp, col,
                               static_cast<const uint16_t *>(val), X, Y, alpha, beta,
                               partials, s);
    default:
      set_error(SBLAS_ERR_ARGS, "unknown value type");
      return SBLAS_ERR_ARGS;
  }
}

int launch_merge(const int *rows, const int *ptr, const int *slots, int groups,
                 const double *partials, int k, double *y, double alpha, double beta,
                 hipStream_t s) {
  if (groups <= 0) return SBLAS_OK;
  const int threads = groups * k;
  hipLaunchKernelGGL(k_merge, dim3((threads + 255) / 256), dim3(256), 0, s, rows, ptr, slots,
                     groups, partials, k, y, alpha, beta);
  SBLAS_HIP_TRY(hipGetLastError());
  return SBLAS_OK;
}

int check_indices(const int *a, int64_t count, int lo, int hi, bool sorted, hipStream_t s,
                  bool *ok) {
  *ok = true;
  if (count <= 0) return SBLAS_OK;
  int *flag = nullptr;
  SBLAS_HIP_TRY(hipMalloc(&flag, sizeof(int)));
  int h = 0;
  hipError_t e = hipMemcpyAsync(flag, &h, sizeof(int), hipMemcpyHostToDevice, s);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_check_indices, dim3(static_cast<unsigned>((count + 255) / 256)), dim3(256),
                       0, s, count, a, lo, hi, sorted ? 1 : 0, flag);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(&h, flag, sizeof(int), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(flag);
  if (e != hipSuccess) {
    set_error(SBLAS_ERR_HIP, std::string("index check: ") + hipGetErrorString(e));
    return SBLAS_ERR_HIP;
  }
  *ok = (h == 0);
  return SBLAS_OK;
}

int launch_scale(int64_t count, double beta, double *y, hipStream_t s) {
  if (count <= 0) return SBLAS_OK;
  hipLaunchKernelGGL(k_scale, dim3(static_cast<unsigned>((count + 255) / 256)), dim3(256), 0, s,
                     count, beta, y);
  SBLAS_HIP_TRY(hipGetLastError());
  return SBLAS_OK;
}

}  // namespace sblas
