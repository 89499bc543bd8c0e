// sorted_panel_device.hpp -- device body of the sorted-group layout of the
// column-panel form (PanelFormat::layout == kPanelSorted), shared by the
// single-matrix launch (sblas_panel.hip) and the batched persistent launch
// (sblas_batch.hip).
//
// Inside a (row panel, column panel) slice the rows are ordered by their
// element count in the slice, longest first, and cut into groups of 64; group
// g keeps element k of its i-th row at goff[g] + 64k + i, and rows shorter
// than the group's first are padded with column kSortedPad.  One lane owns one
// row of a group: for k = 0, 1, ... a wave loads 64 consecutive values and
// column offsets (whole lines), gathers x from the slice's window in LDS and
// adds the product to the row's running sum in a register.  No products pass
// through LDS, no row boundaries are looked up, and a wave's loop length is
// its group's first row: sorting keeps the padding to the spread of lengths
// inside a group.  The running sums live in LDS between slices (the lane that
// owns a row changes from slice to slice).
//
// Each wave w takes groups w, w + 4, w + 8, ... of a slice, so the four waves
// get long and short groups alike.  The slice's element count and the group
// lengths are the same for every lane, so the loops branch on scalars.
//
// Order of operations: a row's products are added in ascending column order
// (slices in order, and ascending columns inside a slice), starting from 0,
// with separately rounded multiplies and adds -- the reference host loop's
// order (test/dspmv_test_smt.cu:45-51).  A padding element adds -0.0, which
// leaves every sum unchanged.  With NX (FP32 / BF16 values, x staged from its
// FP32 copy) products are rounded to FP32 and added in FP64.
//
// BF16 values come packed (VT = uint32_t): each element's value in the high
// half of a word and its column offset in the low half, so a lane issues one
// 4-byte load per element instead of two 2-byte loads (banded n = 2e6, 8
// chunks: 0.299 against 0.314 ms).
#pragma once

#include <type_traits>

#include "internal.hpp"
#include "panel_device.hpp"

namespace sblas {
namespace sorted_panel {

// LDS of one row panel body, in doubles: the slice's x window, its row order
// (256*RPT 16-bit entries, staged in whole rounds of 256 words), its group
// offsets and the rows' running sums
template <int RPT, bool NX>
__host__ __device__ constexpr int lds_doubles() {
  return (NX ? kPanelCols / 2 : kPanelCols) + ((RPT + 1) / 2) * 128 + 2 * RPT + 1 + 256 * RPT;
}
// Elements of each of its rows a lane loads per round: the rounds of a slice
// follow one another (each waits for its loads), so a lane keeps about 16
// loads in flight whatever its rows per thread (rows of many elements per
// slice need them: at one row per lane and 4 per round, 46 elements a row
// per slice ran 0.16 ms where rows in order took 0.10), and at 6 or 8 rows
// three per row keep its registers within a quarter of the file.
__host__ __device__ constexpr int unroll_for(int rpt) {
  return rpt <= 1 ? 16 : rpt == 2 ? 8 : rpt <= 4 ? 4 : 3;
}

// a packed BF16 element (see above): its value
template <typename VT>
constexpr bool kPackedBf16 = std::is_same<VT, uint32_t>::value;

template <typename VT>
__device__ __forceinline__ double widen1(VT v) {
  if constexpr (kPackedBf16<VT>)
    return static_cast<double>(__uint_as_float(static_cast<uint32_t>(v) & 0xFFFF0000u));
  else if constexpr (sizeof(VT) == 2)
    return static_cast<double>(__uint_as_float(static_cast<uint32_t>(v) << 16));
  else
    return static_cast<double>(v);
}
template <typename VT>
__device__ __forceinline__ float widen1_f(VT v) {
  if constexpr (kPackedBf16<VT>)
    return __uint_as_float(static_cast<uint32_t>(v) & 0xFFFF0000u);
  else if constexpr (sizeof(VT) == 2)
    return __uint_as_float(static_cast<uint32_t>(v) << 16);
  else
    return static_cast<float>(v);
}

// Row panel p in the sorted-group layout, by one workgroup of 256 threads.
// lds holds lds_doubles<RPT, NX>() doubles.  R (a multiple of 64, at most
// 256*RPT) rows per panel; slots as in panel::row_panel.  Each non-empty
// slice: its x window, row order and group offsets are loaded into LDS (a
// barrier), its rows summed, and a barrier before the next slice's loads.
// KU: elements of each row a lane loads per round (unroll_for); packed BF16
// elements take one register each, so at 6 or 8 rows four a row (banded
// n = 2e6: 0.2946 against 0.2985 ms with three, 0.304 with six).
template <typename VT, int RPT, bool NX,
          int KU = (kPackedBf16<VT> && RPT >= 6) ? 4 : unroll_for(RPT)>
__device__ __forceinline__ void panel(int p, int n, int R, const PanelInfo &pi,
                                      const int *__restrict__ slice_base,
                                      const int64_t *__restrict__ slice_off,
                                      const uint16_t *__restrict__ perm,
                                      const uint32_t *__restrict__ goff,
                                      const VT *__restrict__ pval,
                                      const uint16_t *__restrict__ pcol,
                                      const double *__restrict__ x, const float *__restrict__ x32,
                                      double *__restrict__ y, double alpha, double beta,
                                      double *__restrict__ slots, double *lds) {
  constexpr int W = kPanelCols;
  constexpr int XD = NX ? W / 2 : W;
  constexpr int PD = ((RPT + 1) / 2) * 128;
  const int G = R >> 6;
  double *ysh = lds + XD + PD + 2 * RPT + 1;
  const int t = threadIdx.x;
  const int w = t >> 6, l = t & 63, wv = t & ~63;
  const int s0 = slice_base[p], s1 = slice_base[p + 1];
  const bool x16 = (reinterpret_cast<uintptr_t>(x) & 15) == 0 && n >= 2;
  double xfix = 0.0;
  int fix_at = -1;
  uint32_t gnext = 0;  // this thread's group offset of the slice being staged

#pragma unroll
  for (int j = 0; j < RPT; j++) ysh[t + 256 * j] = 0.0;

  // the first non-empty slice at or after s
  auto next_slice = [&](int s) {
    while (s < s1 && slice_off[s + 1] == slice_off[s]) s++;
    return s;
  };
  // slice s's x window, row order and (into a register) group offsets; pieces
  // of x past n load a valid address (their columns never occur); an odd n's
  // last x is loaded alone and written once it landed
  auto stage = [&](int s) {
    double *buf = lds;
    const int c0 = (pi.q_lo + (s - s0)) * W;
    if constexpr (NX) {
      float *xsf = reinterpret_cast<float *>(buf);
#pragma unroll
      for (int i = 0; i < W / 1024; i++)
        panel::load16_to_lds(x32 + c0 + 4 * (i * 256 + t), xsf + 4 * (i * 256 + wv));
      fix_at = -1;
    } else if (x16) {
#pragma unroll
      for (int i = 0; i < W / 512; i++) {
        const int c = c0 + 2 * (i * 256 + t);
        panel::load16_to_lds(c + 1 < n ? x + c : x, buf + 2 * (i * 256 + wv));
      }
      const int d = n - 1 - c0;
      const bool fix = (n & 1) && d >= 0 && d < W && ((d >> 1) & 255) == t;
      if (fix) xfix = x[n - 1];
      fix_at = fix ? d : -1;
    } else {
      const uint32_t *xw = reinterpret_cast<const uint32_t *>(x);
      const int64_t g0 = 2 * static_cast<int64_t>(c0), gmax = 2 * static_cast<int64_t>(n) - 1;
#pragma unroll 8
      for (int i = 0; i < 2 * W / 256; i++)
        panel::load4_to_lds(xw + min(g0 + i * 256 + t, gmax),
                            reinterpret_cast<uint32_t *>(buf) + i * 256 + wv);
      fix_at = -1;
    }
    // R 16-bit entries = R/2 words, in rounds of 256 words
    const uint32_t *ps = reinterpret_cast<const uint32_t *>(perm + static_cast<int64_t>(s) * R);
    uint32_t *pd = reinterpret_cast<uint32_t *>(buf + XD);
    const int words = R >> 1;
#pragma unroll
    for (int r = 0; r < (RPT + 1) / 2; r++)
      if (r * 256 < words) panel::load4_to_lds(ps + min(r * 256 + t, words - 1), pd + r * 256 + wv);
    gnext = t <= G ? goff[static_cast<int64_t>(s) * (G + 1) + t] : 0u;
  };
  // after this thread's staging loads have landed: the odd-n fix-up and the
  // group offsets
  auto land = [&]() {
    panel::wait_vector_memory();
    double *buf = lds;
    if (fix_at >= 0) buf[fix_at] = xfix;
    if (t <= G) reinterpret_cast<uint32_t *>(buf + XD + PD)[t] = gnext;
  };
  // the rows of slice s owned by this lane: group w + 4j, lane l
  auto sum_slice = [&](int s) {
    const double *buf = lds;
    const uint16_t *pm = reinterpret_cast<const uint16_t *>(buf + XD);
    const uint32_t *go = reinterpret_cast<const uint32_t *>(buf + XD + PD);
    const int64_t sb = slice_off[s];
    int len[RPT], gb[RPT], row[RPT];
    double acc[RPT];
#pragma unroll
    for (int j = 0; j < RPT; j++) {
      const int g = w + 4 * j;
      int b0 = 0, b1 = 0;
      if (g < G) {
        b0 = static_cast<int>(go[g]);
        b1 = static_cast<int>(go[g + 1]);
      }
      // the same in every lane: scalars
      gb[j] = __builtin_amdgcn_readfirstlane(b0);
      len[j] = __builtin_amdgcn_readfirstlane((b1 - b0) >> 6);
      row[j] = len[j] > 0 ? pm[64 * g + l] : 0;
      acc[j] = len[j] > 0 ? ysh[row[j]] : 0.0;
    }
    const VT *vs = pval + sb + l;
    const uint16_t *cs = pcol + sb + l;
    // groups are sorted longest first: group w is this wave's longest
    for (int k = 0; k < len[0]; k += KU) {
      VT v[RPT][KU];
      uint16_t c[RPT][KU];
#pragma unroll
      for (int j = 0; j < RPT; j++)
#pragma unroll
        for (int kk = 0; kk < KU; kk++)
          if (k + kk < len[j]) {
            const int e = gb[j] + 64 * (k + kk);
            v[j][kk] = __builtin_nontemporal_load(vs + e);
            if constexpr (!kPackedBf16<VT>) c[j][kk] = __builtin_nontemporal_load(cs + e);
          }
#pragma unroll
      for (int j = 0; j < RPT; j++)
#pragma unroll
        for (int kk = 0; kk < KU; kk++)
          if (k + kk < len[j]) {
            // the column of a packed element is split off only here, where
            // the product needs it
            const uint16_t cv = kPackedBf16<VT>
                                    ? static_cast<uint16_t>(static_cast<uint32_t>(v[j][kk]))
                                    : c[j][kk];
            const int ci = cv & (W - 1);
            const bool pad = cv == kSortedPad;
            if constexpr (NX) {
              const float pr = panel::mul_rn_f(widen1_f(v[j][kk]),
                                               reinterpret_cast<const float *>(buf)[ci]);
              acc[j] = panel::add_rn(acc[j], pad ? -0.0 : static_cast<double>(pr));
            } else {
              const double pr = panel::mul_rn(widen1(v[j][kk]), buf[ci]);
              acc[j] = panel::add_rn(acc[j], pad ? -0.0 : pr);
            }
          }
    }
#pragma unroll
    for (int j = 0; j < RPT; j++)
      if (len[j] > 0) ysh[row[j]] = acc[j];
  };

  int s = next_slice(s0);
  if (s < s1) {
    stage(s);
    land();
  }
  __syncthreads();  // running sums zeroed; the first slice staged
  while (s < s1) {
    const int sn = next_slice(s + 1);
    sum_slice(s);
    __syncthreads();  // sums written; the window no longer read
    if (sn < s1) {
      stage(sn);
      land();
      __syncthreads();
    }
    s = sn;
  }
#pragma unroll
  for (int j = 0; j < RPT; j++) {
    const int rl = t + 256 * j;
    if (rl < pi.nrow) {
      const int r = pi.row0 + rl;
      const double a = ysh[rl];
      if (rl == 0 && (pi.flags & 1)) {
        slots[0] = a;
      } else if (rl == pi.nrow - 1 && (pi.flags & 2)) {
        slots[1] = a;
      } else {
        y[r] = (beta == 0.0) ? panel::mul_rn(alpha, a)
                             : panel::add_rn(panel::mul_rn(alpha, a), panel::mul_rn(beta, y[r]));
      }
    }
  }
  __syncthreads();  // LDS may be reused right after
}

}  // namespace sorted_panel
}  // namespace sblas
