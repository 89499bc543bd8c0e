// sblas_batch.hip -- many matrices of uneven sizes in one persistent launch.
//
// The work of every matrix is cut into row blocks (the LDS-staged form of
// kernel 1) and all blocks of all matrices form one list.  A persistent grid
// of a few workgroups per CU walks the list with a fixed stride (workgroup g
// takes items g, g + grid, g + 2*grid, ...), so the assignment -- and every
// sum -- is the same on every run; no counters or atomics are involved.
// Each item carries its matrix's parameters (arrays, value type, lanes per
// row chosen by the autotuner), so matrices with different row-length
// profiles share the launch.  Pieces of rows longer than one block leave
// partials that one merge launch combines in block order.
#include "internal.hpp"

namespace sblas {
namespace {

__device__ __forceinline__ double b_val(const void *p, int type, int j) {
  if (type == SBLAS_VAL_F64) return __builtin_nontemporal_load(static_cast<const double *>(p) + j);
  if (type == SBLAS_VAL_F32)
    return static_cast<double>(__builtin_nontemporal_load(static_cast<const float *>(p) + j));
  const uint32_t h = __builtin_nontemporal_load(static_cast<const uint16_t *>(p) + j);
  return static_cast<double>(__uint_as_float(h << 16));
}

__device__ __forceinline__ double b_axpby(double alpha, double s, double beta, double yv) {
  return (beta == 0.0) ? alpha * s : alpha * s + beta * yv;
}

template <int ITEMS, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_batch_rowblocks(
    int nitems, const BatchItem *__restrict__ items, const BatchMatrix *__restrict__ mats,
    double *__restrict__ partials) {
  constexpr int CAP = BLOCK * ITEMS;
  __shared__ double prod[CAP];
  __shared__ double ysh[kStreamRowsMax];
  __shared__ double wsum[BLOCK / 64];
  const int t = threadIdx.x;
  for (int w = blockIdx.x; w < nitems; w += gridDim.x) {
    const BatchItem it = items[w];
    const BatchMatrix &M = mats[it.mat];
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
        v[k] = b_val(M.val, M.val_type, e);
        c[k] = __builtin_nontemporal_load(col + e);
      }
      double xv[ITEMS];
#pragma unroll
      for (int k = 0; k < ITEMS; k++) xv[k] = x[c[k]];
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
        partials[2 * w] = tot;
      }
      __syncthreads();
      continue;
    }
#pragma unroll
    for (int k = 0; k < ITEMS; k++) {
      const int e = t + k * BLOCK;
      if (e < hi - lo) prod[e] = p[k];
    }
    __syncthreads();
    const int sub = M.subwave;  // 1..64, power of two
    const int subs = BLOCK / sub;
    const int g = t / sub, lane = t % sub;
    for (int i = g; i < nrows; i += subs) {
      const int r = it.row_begin + i;
      const int s0 = max(rp[r], lo) - lo, e0 = min(rp[r + 1], hi) - lo;
      double s = 0.0;
      for (int j = s0 + lane; j < e0; j += sub) s += prod[j];
      for (int off = sub >> 1; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
      if (lane == 0) ysh[i] = s;
    }
    __syncthreads();
    for (int i = t; i < nrows; i += BLOCK) {
      const int r = it.row_begin + i;
      const double s = ysh[i];
      const bool edge = (i == 0 || i == nrows - 1);
      if (edge && (rp[r] < lo || rp[r + 1] > hi)) {
        partials[2 * w + (i == 0 ? 0 : 1)] = s;
      } else {
        y[r] = b_axpby(M.alpha, s, M.beta, y[r]);
      }
    }
    __syncthreads();
  }
}

// Split rows of all matrices: one thread per row, partials in block order.
__global__ void k_batch_merge(int groups, const int *__restrict__ gmat,
                              const int *__restrict__ grow, const int *__restrict__ gptr,
                              const int *__restrict__ gslot, const BatchMatrix *__restrict__ mats,
                              const double *__restrict__ partials) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= groups) return;
  double s = 0.0;
  for (int q = gptr[g]; q < gptr[g + 1]; q++) s += partials[gslot[q]];
  const BatchMatrix &M = mats[gmat[g]];
  double *yp = M.y + grow[g];
  *yp = b_axpby(M.alpha, s, M.beta, *yp);
}

}  // namespace

int launch_batch_rowblocks(int items_per_thread, int grid, int nitems, const BatchItem *items,
                           const BatchMatrix *mats, double *partials, hipStream_t s) {
  if (nitems <= 0 || grid <= 0) return SBLAS_OK;
  constexpr int B = kStreamBlock;
  if (items_per_thread == 16)
    hipLaunchKernelGGL((k_batch_rowblocks<16, B>), dim3(grid), dim3(B), 0, s, nitems, items,
                       mats, partials);
  else
    hipLaunchKernelGGL((k_batch_rowblocks<8, B>), dim3(grid), dim3(B), 0, s, nitems, items, mats,
                       partials);
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
