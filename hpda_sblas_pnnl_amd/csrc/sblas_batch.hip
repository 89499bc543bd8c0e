// sblas_batch.hip -- many matrices of uneven sizes in one persistent launch.
//
// Every matrix is cut into work items of the form its autotuner choice
// names: row panels of the column-panel form (matrices whose rows draw x from
// a wide window) or row blocks of the LDS-staged row form, and all items of
// all matrices form one list in matrix order, cut into eight queues of equal
// work, one per XCD.  A persistent grid sized to the resident workgroups
// takes items from per-queue device counters (its own XCD's queue first,
// then the others), so the launch balances itself across the GPU while
// neighbouring items share that XCD's L2; an item's arithmetic (and so every sum) is
// the same whichever workgroup runs it.  Each item carries its matrix's
// parameters (arrays, value type, lanes per row), so matrices with different
// profiles share the launch.  Rows cut by an item's element window leave
// partials that one merge launch combines in item order.
#include "internal.hpp"
#include "panel_device.hpp"

namespace sblas {
namespace {

__device__ __forceinline__ double t_val(const double *p, int j) {
  return __builtin_nontemporal_load(p + j);
}
__device__ __forceinline__ double t_val(const float *p, int j) {
  return static_cast<double>(__builtin_nontemporal_load(p + j));
}
__device__ __forceinline__ double t_val(const uint16_t *p, int j) {
  const uint32_t h = __builtin_nontemporal_load(p + j);
  return static_cast<double>(__uint_as_float(h << 16));
}

__device__ __forceinline__ double b_axpby(double alpha, double s, double beta, double yv) {
  return (beta == 0.0) ? alpha * s : alpha * s + beta * yv;
}

// One row block (the LDS-staged form of kernel 1): products of the block's
// element window into LDS, sub-wave row sums, y written (or the partials of
// rows the window cuts).  lds holds BLOCK*ITEMS + kStreamRowsMax doubles.
template <typename VT, int ITEMS, int BLOCK>
__device__ __forceinline__ void row_block(int w, const BatchItem &it, const BatchMatrix &M,
                                          double *__restrict__ partials, double *lds,
                                          double *wsum) {
  const VT *__restrict__ val = static_cast<const VT *>(M.val);
  constexpr int CAP = BLOCK * ITEMS;
  double *prod = lds;
  double *ysh = lds + CAP;
  const int t = threadIdx.x;
  const int *__restrict__ rp = M.rp;
  const int *__restrict__ col = M.col;
  const double *__restrict__ x = M.x;
  double *__restrict__ y = M.y;
  const int lo = it.lo, hi = it.hi;
  double p[ITEMS];
  if (hi > lo) {
    double v[ITEMS];
    int c[ITEMS];
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
      const int e = min(lo + t + k * BLOCK, hi - 1);
      v[k] = t_val(val, e);
      c[k] = __builtin_nontemporal_load(col + e);
    }
    double xv[ITEMS];
    if (sizeof(VT) < 8 && M.x32) {
      // FP32 copy: 4-byte gathers, half the lines of x; an FP32 value times
      // an FP32 x is exact in FP64, so x's rounding is the only one added
      float xf[ITEMS];
#pragma unroll
      for (int k = 0; k < ITEMS; k++) xf[k] = M.x32[c[k]];
#pragma unroll
      for (int k = 0; k < ITEMS; k++) xv[k] = static_cast<double>(xf[k]);
    } else {
#pragma unroll
      for (int k = 0; k < ITEMS; k++) xv[k] = x[c[k]];
    }
#pragma unroll
    for (int k = 0; k < ITEMS; k++) p[k] = (lo + t + k * BLOCK < hi) ? v[k] * xv[k] : 0.0;
  } else {
#pragma unroll
    for (int k = 0; k < ITEMS; k++) p[k] = 0.0;
  }
  const int nrows = it.row_end - it.row_begin;
  if (nrows == 1 && (rp[it.row_begin] < lo || rp[it.row_begin + 1] > hi)) {
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < ITEMS; k++) s += p[k];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if ((t & 63) == 0) wsum[t >> 6] = s;
    __syncthreads();
    if (t == 0) {
      double tot = 0.0;
#pragma unroll
      for (int q = 0; q < BLOCK / 64; q++) tot += wsum[q];
      partials[it.pslot] = tot;
    }
    __syncthreads();
    return;
  }
#pragma unroll
  for (int k = 0; k < ITEMS; k++) {
    const int e = t + k * BLOCK;
    if (e < hi - lo) prod[e] = p[k];
  }
  __syncthreads();
  // row sums: one sub-wave per row; a row of more than 16 sub-wave widths is
  // left to whole waves afterwards (found by a ballot over the block's rows,
  // the same list in every wave, dealt out in turn), so a long row never
  // holds the block for len / sub steps
  const int sub = M.subwave;  // 1..64, power of two
  const int subs = BLOCK / sub;
  const int g = t / sub, lane = t % sub;
  const int longest = 16 * sub;
  for (int i = g; i < nrows; i += subs) {
    const int r = it.row_begin + i;
    const int s0 = max(rp[r], lo) - lo, e0 = min(rp[r + 1], hi) - lo;
    if (sub < 64 && e0 - s0 > longest) continue;
    double s = 0.0;
    for (int j = s0 + lane; j < e0; j += sub) s += prod[j];
    for (int off = sub >> 1; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) ysh[i] = s;
  }
  if (sub < 64) {
    const int w = t >> 6, l = t & 63;
    int seen = 0;
    for (int base = 0; base < nrows; base += 64) {
      const int i = base + l;
      bool is_long = false;
      if (i < nrows) {
        const int r = it.row_begin + i;
        is_long = min(rp[r + 1], hi) - max(rp[r], lo) > longest;
      }
      uint64_t mask = __ballot(is_long);
      while (mask) {
        const int bit = __ffsll(static_cast<unsigned long long>(mask)) - 1;
        mask &= mask - 1;
        if (seen % (BLOCK / 64) == w) {
          const int r = it.row_begin + base + bit;
          const int s0 = max(rp[r], lo) - lo, e0 = min(rp[r + 1], hi) - lo;
          double s = 0.0;
          for (int j = s0 + l; j < e0; j += 64) s += prod[j];
#pragma unroll
          for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
          if (l == 0) ysh[base + bit] = s;
        }
        seen++;
      }
    }
  }
  __syncthreads();
  for (int i = t; i < nrows; i += BLOCK) {
    const int r = it.row_begin + i;
    const double s = ysh[i];
    const bool edge = (i == 0 || i == nrows - 1);
    if (edge && (rp[r] < lo || rp[r + 1] > hi)) {
      partials[it.pslot + (i == 0 ? 0 : 1)] = s;
    } else {
      y[r] = b_axpby(M.alpha, s, M.beta, y[r]);
    }
  }
  __syncthreads();
}

// One row block with K right-hand sides (X n x K, Y m x K, row-major): lane
// t works on vector t % K.  A block of one row (or a piece of one) spreads it
// over all BLOCK/K lane groups and reduces across the workgroup in a fixed
// order; otherwise each row gets max(subwave, K) lanes.  Cut rows leave K
// partials per slot: partials[pslot + i*K + vec], i = 0 first row, 1 last.
template <typename VT, int K, int BLOCK>
__device__ __forceinline__ void row_block_mrhs(const BatchItem &it, const BatchMatrix &M,
                                               double *__restrict__ partials, double *lds) {
  const VT *__restrict__ val = static_cast<const VT *>(M.val);
  const int t = threadIdx.x;
  const int vec = t % K;
  const int *__restrict__ rp = M.rp;
  const int *__restrict__ col = M.col;
  const double *__restrict__ X = M.x;
  double *__restrict__ Y = M.y;
  const int lo = it.lo, hi = it.hi;
  // sum over j = j0, j0+step, ... < hi, four elements' loads ahead of their gathers
  auto dot = [&](int j0, int j1, int step) {
    double sum = 0.0;
    int j = j0;
    for (; j + 3 * step < j1; j += 4 * step) {
      double v[4];
      int c[4];
#pragma unroll
      for (int i = 0; i < 4; i++) {
        v[i] = t_val(val, j + i * step);
        c[i] = __builtin_nontemporal_load(col + j + i * step);
      }
      double xv[4];
#pragma unroll
      for (int i = 0; i < 4; i++) xv[i] = X[static_cast<int64_t>(c[i]) * K + vec];
#pragma unroll
      for (int i = 0; i < 4; i++) sum += v[i] * xv[i];
    }
    for (; j < j1; j += step)
      sum += t_val(val, j) * X[static_cast<int64_t>(col[j]) * K + vec];
    return sum;
  };
  const int nrows = it.row_end - it.row_begin;
  if (nrows == 1) {
    constexpr int G = BLOCK / K;
    double sum = dot(lo + t / K, hi, G);
#pragma unroll
    for (int off = 32; off >= K; off >>= 1) sum += __shfl_xor(sum, off, 64);
    if ((t & 63) < K) lds[(t >> 6) * K + vec] = sum;
    __syncthreads();
    if (t < K) {
      double s = 0.0;
#pragma unroll
      for (int w = 0; w < BLOCK / 64; w++) s += lds[w * K + t];
      const int r = it.row_begin;
      if (rp[r] < lo || rp[r + 1] > hi) {
        partials[it.pslot + t] = s;
      } else {
        double *yp = Y + static_cast<int64_t>(r) * K + t;
        *yp = b_axpby(M.alpha, s, M.beta, *yp);
      }
    }
    __syncthreads();
    return;
  }
  const int sub = M.subwave < K ? K : M.subwave;  // power of two in [K, 64]
  const int subs = BLOCK / sub;
  const int grp = (t % sub) / K;
  // a row of more than 16 steps of its lane groups is left to whole waves
  // afterwards (64/K groups per vector), as in the plans' multi-RHS kernel
  const int longest = 16 * (sub / K);
  auto finish = [&](int r, int rs, int re, double sum) {
    if (rs < lo || re > hi) {
      partials[it.pslot + (r == it.row_begin ? 0 : K) + vec] = sum;
    } else {
      double *yp = Y + static_cast<int64_t>(r) * K + vec;
      *yp = b_axpby(M.alpha, sum, M.beta, *yp);
    }
  };
  for (int r = it.row_begin + t / sub; r < it.row_end; r += subs) {
    const int rs = rp[r], re = rp[r + 1];
    if (sub < 64 && min(re, hi) - max(rs, lo) > longest) continue;
    double sum = dot(max(rs, lo) + grp, min(re, hi), sub / K);
    for (int off = sub >> 1; off >= K; off >>= 1) sum += __shfl_xor(sum, off, 64);
    if (grp == 0) finish(r, rs, re, sum);
  }
  if (sub < 64) {
    // long rows in block order, found by a ballot (the same list in every
    // wave) and dealt out to the waves in turn
    const int w = t >> 6, l = t & 63, g64 = l / K;
    int seen = 0;
    for (int base = 0; base < nrows; base += 64) {
      bool is_long = false;
      if (base + l < nrows) {
        const int r = it.row_begin + base + l;
        is_long = min(rp[r + 1], hi) - max(rp[r], lo) > longest;
      }
      uint64_t mask = __ballot(is_long);
      while (mask) {
        const int bit = __ffsll(static_cast<unsigned long long>(mask)) - 1;
        mask &= mask - 1;
        if (seen % (BLOCK / 64) == w) {
          const int r = it.row_begin + base + bit;
          const int rs = rp[r], re = rp[r + 1];
          double sum = dot(max(rs, lo) + g64, min(re, hi), 64 / K);
#pragma unroll
          for (int off = 32; off >= K; off >>= 1) sum += __shfl_xor(sum, off, 64);
          if (g64 == 0) finish(r, rs, re, sum);
        }
        seen++;
      }
    }
  }
}

// One row panel of a matrix in column-panel form (16-bit slice offsets).
__device__ __forceinline__ void row_panel_any(const BatchItem &it, const BatchMatrix &M,
                                              double *__restrict__ partials, double *lds) {
  constexpr int RPT = kBatchPanelRowsMax / 256;
  const int p = it.row_begin;
  const PanelInfo pi = M.pinfo[p];
  double *slots = partials + it.pslot;
  switch (M.val_type) {
    case SBLAS_VAL_F64:
      panel::row_panel<double, RPT, uint16_t>(p, M.n, M.prows, pi, M.slice_base, M.slice_off,
                                              M.srp, static_cast<const double *>(M.pval), M.pcol,
                                              M.x, M.y, M.alpha, M.beta, slots, lds);
      break;
    case SBLAS_VAL_F32:
      panel::row_panel<float, RPT, uint16_t>(p, M.n, M.prows, pi, M.slice_base, M.slice_off,
                                             M.srp, static_cast<const float *>(M.pval), M.pcol,
                                             M.x, M.y, M.alpha, M.beta, slots, lds);
      break;
    default:
      panel::row_panel<uint16_t, RPT, uint16_t>(p, M.n, M.prows, pi, M.slice_base, M.slice_off,
                                                M.srp, static_cast<const uint16_t *>(M.pval),
                                                M.pcol, M.x, M.y, M.alpha, M.beta, slots, lds);
      break;
  }
}

// A row-block item of one value type: one vector or K right-hand sides.
template <typename VT, int ITEMS, int BLOCK>
__device__ __forceinline__ void rows_typed(int w, const BatchItem &it, const BatchMatrix &M,
                                           double *__restrict__ partials, double *lds,
                                           double *wsum) {
  switch (M.k) {  // 1, 2, 4, 8 or 16 (checked when the batch is created)
    case 2: row_block_mrhs<VT, 2, BLOCK>(it, M, partials, lds); break;
    case 4: row_block_mrhs<VT, 4, BLOCK>(it, M, partials, lds); break;
    case 8: row_block_mrhs<VT, 8, BLOCK>(it, M, partials, lds); break;
    case 16: row_block_mrhs<VT, 16, BLOCK>(it, M, partials, lds); break;
    default: row_block<VT, ITEMS, BLOCK>(w, it, M, partials, lds, wsum); break;
  }
}

// LDS of one workgroup of a batch launch, in doubles
template <int MODE, int ITEMS, int BLOCK>
__host__ __device__ constexpr int batch_lds() {
  return MODE == kBatchPanel ? panel::lds_doubles<uint16_t, kBatchPanelRowsMax / 256>()
                             : BLOCK * ITEMS + kStreamRowsMax;
}

// work item w of the list
template <int MODE, int ITEMS, int BLOCK>
__device__ __forceinline__ void batch_item(int w, const BatchItem *__restrict__ items,
                                           const BatchMatrix *__restrict__ mats,
                                           double *__restrict__ partials, double *lds,
                                           double *wsum) {
  const BatchItem it = items[w];
  const BatchMatrix &M = mats[it.mat];
  if constexpr (MODE == kBatchPanel) {
    row_panel_any(it, M, partials, lds);
  } else if (M.val_type == SBLAS_VAL_F64) {  // value type chosen once per item
    rows_typed<double, ITEMS, BLOCK>(w, it, M, partials, lds, wsum);
  } else if (M.val_type == SBLAS_VAL_F32) {
    rows_typed<float, ITEMS, BLOCK>(w, it, M, partials, lds, wsum);
  } else {
    rows_typed<uint16_t, ITEMS, BLOCK>(w, it, M, partials, lds, wsum);
  }
}

// The persistent loop of one item kind (MODE kBatchPanel: row panels, else
// row blocks).  Each kind is its own kernel compiled with only its own body,
// so its registers and LDS (52 KB for panels, 20-36 KB for row blocks) set
// its own occupancy.
template <int MODE, int ITEMS, int BLOCK>
__device__ __forceinline__ void batch_loop(const BatchItem *__restrict__ items,
                                           const BatchMatrix *__restrict__ mats,
                                           double *__restrict__ partials,
                                           const int *__restrict__ qbound,
                                           int *__restrict__ ctr) {
  static_assert(BLOCK == 256, "the panel body assumes 256 threads");
  __shared__ double lds[batch_lds<MODE, ITEMS, BLOCK>()];
  __shared__ double wsum[BLOCK / 64];
  __shared__ int next;
  const int t = threadIdx.x;
  // workgroups are dispatched to the eight XCDs in turn: start on this XCD's
  // queue (a contiguous run of the list, so neighbouring row panels share x
  // in that XCD's L2), then help the other queues in order
  const int home = static_cast<int>(blockIdx.x) & (kBatchQueues - 1);
  for (int d = 0; d < kBatchQueues; d++) {
    const int q = (home + d) & (kBatchQueues - 1);
    const int end = qbound[q + 1];
    if (t == 0) next = qbound[q] + atomicAdd(&ctr[q], 1);
    __syncthreads();
    int w = next;
    for (;;) {
      __syncthreads();  // every thread has read `next` before it is taken again
      if (w >= end) break;  // counters only grow: every workgroup gets here
      // claim the following item now; its counter round trip overlaps this item
      int claimed = 0;
      if (t == 0) claimed = qbound[q] + atomicAdd(&ctr[q], 1);
      batch_item<MODE, ITEMS, BLOCK>(w, items, mats, partials, lds, wsum);
      if (t == 0) next = claimed;  // last read before this iteration's first barrier
      __syncthreads();
      w = next;
    }
  }
  if (t == 0) {
    // the last workgroup to finish rewinds the counters for the next launch
    const int done = atomicAdd(&ctr[kBatchQueues], 1);
    if (done == static_cast<int>(gridDim.x) - 1) {
      for (int q = 0; q <= kBatchQueues; q++) atomicExch(&ctr[q], 0);
    }
  }
}

template <int MODE, int ITEMS, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_batch(int nitems, const BatchItem *__restrict__ items,
                                                 const BatchMatrix *__restrict__ mats,
                                                 double *__restrict__ partials,
                                                 const int *__restrict__ qbound,
                                                 int *__restrict__ ctr) {
  batch_loop<MODE, ITEMS, BLOCK>(items, mats, partials, qbound, ctr);
  (void)nitems;
}

// One chunk of kX32Chunk entries of one segment per workgroup (its segment
// read from the chunk table), x32 = FP32 of x: every thread issues its eight
// loads before it stores.  A thread per entry that found its segment by a
// binary search over the segments' starts took 0.10 ms for the 1.5e7 entries
// of the 64-matrix FP32 batch, its dependent loads leaving the GPU idle.
__global__ __launch_bounds__(256) void k_batch_x32(const BatchX32 *__restrict__ segs,
                                                   const int *__restrict__ chunk_seg) {
  constexpr int K = kX32Chunk / 256;
  const BatchX32 S = segs[chunk_seg[blockIdx.x]];
  const int64_t base = static_cast<int64_t>(blockIdx.x - S.first_chunk) * kX32Chunk + threadIdx.x;
  double v[K];
#pragma unroll
  for (int k = 0; k < K; k++) {
    const int64_t j = base + k * 256;
    v[k] = j < S.n ? S.x[j] : 0.0;
  }
#pragma unroll
  for (int k = 0; k < K; k++) {
    const int64_t j = base + k * 256;
    if (j < S.n) S.x32[j] = static_cast<float>(v[k]);
  }
}

// One work item per workgroup, for a launch that shares the GPU with others:
// the hardware hands out its workgroups as CUs free up, where a persistent
// launch would hold every CU until its queues drain.  Workgroup b runs on
// XCD b % 8 and takes item (b % 8) * G/8 + b / 8, so each XCD works through a
// contiguous run of the list (the grid is a multiple of 8).
template <int MODE, int ITEMS, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_batch_flat(int nitems,
                                                      const BatchItem *__restrict__ items,
                                                      const BatchMatrix *__restrict__ mats,
                                                      double *__restrict__ partials) {
  static_assert(BLOCK == 256, "the panel body assumes 256 threads");
  __shared__ double lds[batch_lds<MODE, ITEMS, BLOCK>()];
  __shared__ double wsum[BLOCK / 64];
  const int per = static_cast<int>(gridDim.x) >> 3;
  const int w = (static_cast<int>(blockIdx.x) & 7) * per + (static_cast<int>(blockIdx.x) >> 3);
  if (w >= nitems) return;
  batch_item<MODE, ITEMS, BLOCK>(w, items, mats, partials, lds, wsum);
}

// Split rows of all matrices: one thread per row (all its vectors), partials
// in item order.
__global__ void k_batch_merge(int groups, const int *__restrict__ gmat,
                              const int *__restrict__ grow, const int *__restrict__ gptr,
                              const int *__restrict__ gslot, const BatchMatrix *__restrict__ mats,
                              const double *__restrict__ partials) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= groups) return;
  const BatchMatrix &M = mats[gmat[g]];
  const int k = M.k <= 1 ? 1 : M.k;
  for (int v = 0; v < k; v++) {
    double s = 0.0;
    for (int q = gptr[g]; q < gptr[g + 1]; q++) s += partials[gslot[q] + v];
    double *yp = M.y + static_cast<int64_t>(grow[g]) * k + v;
    *yp = b_axpby(M.alpha, s, M.beta, *yp);
  }
}

}  // namespace

int launch_batch(int kind, int items_per_thread, int grid, int nitems, const BatchItem *items,
                 const BatchMatrix *mats, double *partials, const int *qbound, int *ctr,
                 hipStream_t s, bool flat) {
  if (nitems <= 0 || grid <= 0) return SBLAS_OK;
  constexpr int B = kStreamBlock;
  constexpr int P = kBatchPanel, R = kBatchRowBlock;
  if (flat) {
    const dim3 g(static_cast<unsigned>((nitems + 7) / 8 * 8));
    if (kind == kBatchPanel)
      hipLaunchKernelGGL((k_batch_flat<P, 8, B>), g, dim3(B), 0, s, nitems, items, mats, partials);
    else if (items_per_thread == 16)
      hipLaunchKernelGGL((k_batch_flat<R, 16, B>), g, dim3(B), 0, s, nitems, items, mats,
                         partials);
    else
      hipLaunchKernelGGL((k_batch_flat<R, 8, B>), g, dim3(B), 0, s, nitems, items, mats, partials);
    SBLAS_HIP_TRY(hipGetLastError());
    return SBLAS_OK;
  }
  if (kind == kBatchPanel)
    hipLaunchKernelGGL((k_batch<P, 8, B>), dim3(grid), dim3(B), 0, s, nitems, items, mats,
                       partials, qbound, ctr);
  else if (items_per_thread == 16)
    hipLaunchKernelGGL((k_batch<R, 16, B>), dim3(grid), dim3(B), 0, s, nitems, items, mats,
                       partials, qbound, ctr);
  else
    hipLaunchKernelGGL((k_batch<R, 8, B>), dim3(grid), dim3(B), 0, s, nitems, items, mats,
                       partials, qbound, ctr);
  SBLAS_HIP_TRY(hipGetLastError());
  return SBLAS_OK;
}

int batch_occupancy(int kind, int items_per_thread) {
  int per_cu = 0;
  constexpr int P = kBatchPanel, R = kBatchRowBlock;
  const void *fn =
      kind == kBatchPanel      ? reinterpret_cast<const void *>(&k_batch<P, 8, kStreamBlock>)
      : items_per_thread == 16 ? reinterpret_cast<const void *>(&k_batch<R, 16, kStreamBlock>)
                               : reinterpret_cast<const void *>(&k_batch<R, 8, kStreamBlock>);
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, kStreamBlock, 0) != hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  return per_cu;
}

int launch_batch_x32(const BatchX32 *segs, const int *chunk_seg, int nchunks, hipStream_t s) {
  if (nchunks <= 0) return SBLAS_OK;
  hipLaunchKernelGGL(k_batch_x32, dim3(static_cast<unsigned>(nchunks)), dim3(256), 0, s, segs,
                     chunk_seg);
  SBLAS_HIP_TRY(hipGetLastError());
  return SBLAS_OK;
}

int launch_batch_merge(int groups, const int *gmat, const int *grow, const int *gptr,
                       const int *gslot, const BatchMatrix *mats, const double *partials,
                       hipStream_t s) {
  if (groups <= 0) return SBLAS_OK;
  hipLaunchKernelGGL(k_batch_merge, dim3((groups + 255) / 256), dim3(256), 0, s, groups, gmat,
                     grow, gptr, gslot, mats, partials);
  SBLAS_HIP_TRY(hipGetLastError());
  return SBLAS_OK;
}

}  // namespace sblas
