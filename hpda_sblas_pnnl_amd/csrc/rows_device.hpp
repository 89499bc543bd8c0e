// rows_device.hpp -- the row sums of an LDS-staged row block, shared by the
// row-block kernel (sblas_vector.hip) and the column-sorted row blocks
// (sblas_sorted.hip).
#pragma once

#include "internal.hpp"

namespace sblas {
namespace rows {

// Sum over the W lanes of each aligned group of W lanes (xor butterfly).
template <int W>
__device__ __forceinline__ double lanes_sum(double v) {
#pragma unroll
  for (int off = W / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, W);
  return v;
}

// Rows longer than this many elements are summed by whole waves.
template <int SUBW>
__host__ __device__ constexpr int long_row() {
  return 16 * SUBW;
}

// ysh[i] = sum of the block's products of row i (prod holds the products of
// the block's element window [lo, hi) in element order; rps[i] .. rps[i+1] is
// row i's element range).  A row is summed by one sub-wave of SUBW lanes,
// lane l taking its elements l, l + SUBW, ..., then a butterfly; a row of more
// than long_row<SUBW>() elements is left to a second pass where one whole
// wave sums it the same way with 64 lanes, so a long row never holds the
// block for len / SUBW steps.  Each wave scans the block's rows 64 at a time,
// finds the long ones with a ballot (the same list, in the same order, in
// every wave) and takes every WAVES-th of them.  The order of every sum is
// fixed by the row's element range alone.
template <int SUBW, int BLOCK>
__device__ __forceinline__ void block_row_sums(const double *prod, const int *rps, int nrows,
                                               int lo, int hi, double *ysh) {
  constexpr int SUBS = BLOCK / SUBW;
  constexpr int LONG = long_row<SUBW>();
  const int t = threadIdx.x;
  const int sub = t / SUBW, lane = t % SUBW;
  for (int i = sub; i < nrows; i += SUBS) {
    const int s0 = max(rps[i], lo) - lo, e0 = min(rps[i + 1], hi) - lo;
    if (SUBW < 64 && e0 - s0 > LONG) continue;  // whole-wave pass below
    double s = 0.0;
    for (int j = s0 + lane; j < e0; j += SUBW) s += prod[j];
    if constexpr (SUBW > 1) s = lanes_sum<SUBW>(s);
    if (lane == 0) ysh[i] = s;
  }
  if constexpr (SUBW < 64) {
    constexpr int WAVES = BLOCK / 64;
    const int w = t >> 6, l = t & 63;
    int seen = 0;  // long rows before this scan step, the same in every wave
    for (int base = 0; base < nrows; base += 64) {
      const int i = base + l;
      bool is_long = false;
      if (i < nrows) is_long = min(rps[i + 1], hi) - max(rps[i], lo) > LONG;
      uint64_t mask = __ballot(is_long);
      while (mask) {
        const int bit = __ffsll(static_cast<unsigned long long>(mask)) - 1;
        mask &= mask - 1;
        if (seen % WAVES == w) {
          const int r = base + bit;
          const int s0 = max(rps[r], lo) - lo, e0 = min(rps[r + 1], hi) - lo;
          double s = 0.0;
          for (int j = s0 + l; j < e0; j += 64) s += prod[j];
          s = lanes_sum<64>(s);
          if (l == 0) ysh[r] = s;
        }
        seen++;
      }
    }
  }
}

}  // namespace rows
}  // namespace sblas
