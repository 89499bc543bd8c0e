// panel_device.hpp -- device body of the column-panel form of kernel 1, shared
// by the single-matrix launch (sblas_panel.hip) and the batched persistent
// launch (sblas_batch.hip).  Layout and order of operations are described in
// sblas_panel.hip.
#pragma once

#include "internal.hpp"

namespace sblas {
namespace panel {

// Separately rounded multiply and add.  Contraction is switched off inside
// these bodies, so a product never fuses with the add that consumes it (the
// header intrinsics are plain * and + and would fuse under -ffp-contract=fast).
__device__ __forceinline__ double mul_rn(double a, double b) {
#pragma clang fp contract(off)
  return a * b;
}
__device__ __forceinline__ double add_rn(double a, double b) {
#pragma clang fp contract(off)
  return a + b;
}

constexpr int kItems = 8;               // elements per thread per chunk (4 pairs)
constexpr int kChunk = 256 * kItems;    // elements per chunk of a slice
constexpr int kLdsDoubles = kPanelCols + kChunk;  // x panel + products
static_assert(kPanelPad % 2 == 0, "pairs never straddle a slice end");

typedef double f64x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));

// Two consecutive values (an aligned pair) as doubles, one non-temporal load.
__device__ __forceinline__ void load2(const double *p, double &a, double &b) {
  const f64x2 q = __builtin_nontemporal_load(reinterpret_cast<const f64x2 *>(p));
  a = q.x;
  b = q.y;
}
__device__ __forceinline__ void load2(const float *p, double &a, double &b) {
  const f32x2 q = __builtin_nontemporal_load(reinterpret_cast<const f32x2 *>(p));
  a = q.x;
  b = q.y;
}
__device__ __forceinline__ void load2(const uint16_t *p, double &a, double &b) {
  const uint32_t w = __builtin_nontemporal_load(reinterpret_cast<const uint32_t *>(p));
  a = static_cast<double>(__uint_as_float(w << 16));
  b = static_cast<double>(__uint_as_float(w & 0xffff0000u));
}
// two consecutive 16-bit column offsets
__device__ __forceinline__ void load2_cols(const uint16_t *p, int &a, int &b) {
  const uint32_t w = __builtin_nontemporal_load(reinterpret_cast<const uint32_t *>(p));
  a = static_cast<int>(w & 0xffffu);
  b = static_cast<int>(w >> 16);
}

// Current slice of a row panel and the first element of the chunk worked on;
// len is the slice's length padded to a multiple of kPanelPad.
struct Cursor {
  int s, k0, len;
  int64_t base;
};

// Moves c to the first chunk of the first non-empty slice at or after s.
__device__ __forceinline__ void seek(Cursor &c, int s, int s1,
                                     const int64_t *__restrict__ slice_off) {
  c.k0 = 0;
  c.s = s;
  c.len = 0;
  c.base = 0;
  for (; c.s < s1; c.s++) {
    c.base = slice_off[c.s];
    c.len = static_cast<int>(slice_off[c.s + 1] - c.base);
    if (c.len > 0) return;
  }
}

// Row panel p of one matrix, by one workgroup of 256 threads.  lds holds
// kLdsDoubles doubles; slots[0] / slots[1] receive the partial sums of a
// first / last row that the panel's element window cuts (PanelInfo flags).
// Software-pipelined: the next chunk's values and column offsets are loaded
// while the current chunk's rows are summed and the next slice's x is staged.
// The workgroup leaves through a barrier, so LDS may be reused right after.
template <typename VT, int RPT, typename OT>
__device__ __forceinline__ void row_panel(
    int p, int n, int R, const PanelInfo &pi, const int *__restrict__ slice_base,
    const int64_t *__restrict__ slice_off, const OT *__restrict__ srp,
    const VT *__restrict__ pval, const uint16_t *__restrict__ pcol,
    const double *__restrict__ x, double *__restrict__ y, double alpha, double beta,
    double *__restrict__ slots, double *lds) {
  constexpr int W = kPanelCols;
  double *xs = lds;
  double *prod = lds + W;
  const int t = threadIdx.x;
  const int s0 = slice_base[p], s1 = slice_base[p + 1];
  const int qb = pi.q_lo;
  double acc[RPT];
  int b[RPT], e[RPT];
  double v[kItems];
  int l[kItems];
#pragma unroll
  for (int j = 0; j < RPT; j++) acc[j] = 0.0;

  // x[q*W, (q+1)*W) of slice s into LDS, two columns per thread and step
  // (one 16-byte LDS store); columns past n read as 0
  auto stage_x = [&](int s) {
    const int c0 = (qb + (s - s0)) * W;
#pragma unroll
    for (int i = 0; i < W / 512; i++) {
      const int c = c0 + 2 * t + i * 512;
      const double x0 = x[min(c, n - 1)], x1 = x[min(c + 1, n - 1)];  // no load under a branch
      f64x2 q;
      q.x = (c < n) ? x0 : 0.0;
      q.y = (c + 1 < n) ? x1 : 0.0;
      *reinterpret_cast<f64x2 *>(xs + 2 * t + i * 512) = q;
    }
  };
  // element ranges [b, e) of this thread's rows inside slice s
  auto row_ranges = [&](int s) {
    const OT *rps = srp + static_cast<int64_t>(s) * (R + 1);
#pragma unroll
    for (int j = 0; j < RPT; j++) {
      const int rl = t + 256 * j;
      const bool in = rl < R;
      const int lo = rps[in ? rl : 0], hi = rps[in ? rl + 1 : 0];  // unconditional loads
      b[j] = in ? lo : 0;
      e[j] = in ? hi : 0;
    }
  };
  // values and column offsets of a chunk: pair i of thread t is elements
  // 512*i + 2t, +1 (so one load instruction of a wave is a contiguous run of
  // whole lines), all issued back to back; threads past the slice's end
  // re-read its last pair (never summed)
  auto load_chunk = [&](const Cursor &c) {
#pragma unroll
    for (int i = 0; i < kItems / 2; i++) {
      const int64_t k = c.base + min(c.k0 + 512 * i + 2 * t, c.len - 2);
      load2(pval + k, v[2 * i], v[2 * i + 1]);
      load2_cols(pcol + k, l[2 * i], l[2 * i + 1]);
    }
  };

  Cursor cur;
  seek(cur, s0, s1, slice_off);
  if (cur.s < s1) {
    stage_x(cur.s);
    row_ranges(cur.s);
    load_chunk(cur);
  }
  __syncthreads();
  while (cur.s < s1) {
    // products of the current chunk (x of its slice is in LDS)
#pragma unroll
    for (int i = 0; i < kItems / 2; i++) {
      f64x2 pr;
      pr.x = mul_rn(v[2 * i], xs[l[2 * i]]);
      pr.y = mul_rn(v[2 * i + 1], xs[l[2 * i + 1]]);
      *reinterpret_cast<f64x2 *>(prod + 512 * i + 2 * t) = pr;  // 16-byte LDS stores
    }
    Cursor nxt = cur;
    nxt.k0 += kChunk;
    if (nxt.k0 >= cur.len) seek(nxt, cur.s + 1, s1, slice_off);
    // the next chunk's loads are in flight during the row sums below
    if (nxt.s < s1) load_chunk(nxt);
    __syncthreads();  // products visible; x of the current slice no longer read
    const bool new_slice = nxt.s != cur.s && nxt.s < s1;
    if (new_slice) stage_x(nxt.s);
    // each row adds its products of this chunk in element order, the next
    // product's LDS read issued ahead of each add
#pragma unroll
    for (int j = 0; j < RPT; j++) {
      const int lo = max(b[j], cur.k0) - cur.k0, hi = min(e[j], cur.k0 + kChunk) - cur.k0;
      double a = acc[j];
      if (lo < hi) {
        double q = prod[lo];
        for (int k = lo + 1; k < hi; k++) {
          const double nq = prod[k];
          a = add_rn(a, q);
          q = nq;
        }
        a = add_rn(a, q);
      }
      acc[j] = a;
    }
    if (new_slice) row_ranges(nxt.s);
    __syncthreads();  // products consumed; x of the next slice visible
    cur = nxt;
  }
#pragma unroll
  for (int j = 0; j < RPT; j++) {
    const int rl = t + 256 * j;
    if (rl < pi.nrow) {
      const int r = pi.row0 + rl;
      if (rl == 0 && (pi.flags & 1)) {
        slots[0] = acc[j];
      } else if (rl == pi.nrow - 1 && (pi.flags & 2)) {
        slots[1] = acc[j];
      } else {
        y[r] = (beta == 0.0) ? mul_rn(alpha, acc[j])
                             : add_rn(mul_rn(alpha, acc[j]), mul_rn(beta, y[r]));
      }
    }
  }
}

}  // namespace panel
}  // namespace sblas
