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

// An aligned pair of consecutive values as loaded (one non-temporal load),
// widened to doubles only where the products are formed.  Loads keep their
// raw registers until then: a conversion right behind a load would make the
// wave wait for it where the load is issued, ahead of the row sums that the
// load is meant to overlap.
template <typename VT> struct RawPair;
template <> struct RawPair<double> { typedef f64x2 type; };
template <> struct RawPair<float> { typedef f32x2 type; };
template <> struct RawPair<uint16_t> { typedef uint32_t type; };  // two bfloat16

template <typename VT>
__device__ __forceinline__ typename RawPair<VT>::type load_pair(const VT *p) {
  typedef typename RawPair<VT>::type R;
  return __builtin_nontemporal_load(reinterpret_cast<const R *>(p));
}
__device__ __forceinline__ void widen(f64x2 q, double &a, double &b) {
  a = q.x;
  b = q.y;
}
__device__ __forceinline__ void widen(f32x2 q, double &a, double &b) {
  a = q.x;
  b = q.y;
}
__device__ __forceinline__ void widen(uint32_t w, double &a, double &b) {
  a = static_cast<double>(__uint_as_float(w << 16));
  b = static_cast<double>(__uint_as_float(w & 0xffff0000u));
}
// two consecutive 16-bit column offsets, one 32-bit load, split at use
__device__ __forceinline__ uint32_t load_cols(const uint16_t *p) {
  return __builtin_nontemporal_load(reinterpret_cast<const uint32_t *>(p));
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
// while the current chunk's rows are summed.  Before a slice change the next
// slice's x entries and row ranges are requested ahead of the chunk's loads:
// vector loads complete in issue order, so x can then be stored to LDS while
// the chunk is still on its way.  The workgroup leaves through a barrier, so
// LDS may be reused right after.
template <typename VT, int RPT, typename OT>
__device__ __forceinline__ void row_panel(
    int p, int n, int R, const PanelInfo &pi, const int *__restrict__ slice_base,
    const int64_t *__restrict__ slice_off, const OT *__restrict__ srp,
    const VT *__restrict__ pval, const uint16_t *__restrict__ pcol,
    const double *__restrict__ x, double *__restrict__ y, double alpha, double beta,
    double *__restrict__ slots, double *lds) {
  constexpr int W = kPanelCols;
  constexpr int NX = W / 512;  // x pairs per thread and slice
  typedef typename RawPair<VT>::type RV;
  double *xs = lds;
  double *prod = lds + W;
  const int t = threadIdx.x;
  const int s0 = slice_base[p], s1 = slice_base[p + 1];
  const int qb = pi.q_lo;
  double acc[RPT];
  int b[RPT], e[RPT];    // element range [b, e) of row t + 256 j in the current slice
  int nb[RPT], ne[RPT];  // the same in the next slice
  RV v[kItems / 2];      // the next chunk's value pairs
  uint32_t cw[kItems / 2];  // and column offset pairs
  double x0[NX], x1[NX];    // the next slice's x entries
  int64_t end2 = 0;  // slice_off at the end of the slice after the current one
#pragma unroll
  for (int j = 0; j < RPT; j++) acc[j] = 0.0;

  // the end of the slice after s, read one slice ahead.  Odd threads read
  // the entry after it, which makes the address per-thread: a vector load,
  // which the chunk's waits cover, where a scalar load would be waited for at
  // the next barrier.  The value is taken from a wave's first (even) thread.
  auto load_end2 = [&](int s) { end2 = slice_off[min(s + 2 + (t & 1), s1)]; };
  // x[q*W, (q+1)*W) of slice s, two columns per thread and step; the
  // addresses are clamped (no load under a branch), columns past n are zeroed
  // when stored (one 16-byte LDS store per step)
  auto load_x = [&](int s) {
    const int c0 = (qb + (s - s0)) * W;
#pragma unroll
    for (int i = 0; i < NX; i++) {
      const int c = c0 + 2 * t + i * 512;
      x0[i] = x[min(c, n - 1)];
      x1[i] = x[min(c + 1, n - 1)];
    }
  };
  auto store_x = [&](int s) {
    const int c0 = (qb + (s - s0)) * W;
#pragma unroll
    for (int i = 0; i < NX; i++) {
      const int c = c0 + 2 * t + i * 512;
      f64x2 q;
      q.x = (c < n) ? x0[i] : 0.0;
      q.y = (c + 1 < n) ? x1[i] : 0.0;
      *reinterpret_cast<f64x2 *>(xs + 2 * t + i * 512) = q;
    }
  };
  // element ranges of this thread's rows inside slice s; a row past R reads
  // the slice's end twice (an empty range), with no select behind the load
  auto load_ranges = [&](int s) {
    const OT *rps = srp + static_cast<int64_t>(s) * (R + 1);
#pragma unroll
    for (int j = 0; j < RPT; j++) {
      const int rl = t + 256 * j;
      nb[j] = static_cast<int>(rps[min(rl, R)]);
      ne[j] = static_cast<int>(rps[min(rl + 1, R)]);
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
      v[i] = load_pair(pval + k);
      cw[i] = load_cols(pcol + k);
    }
  };

  Cursor cur;
  seek(cur, s0, s1, slice_off);
  if (cur.s < s1) {
    load_end2(cur.s);
    load_x(cur.s);
    load_ranges(cur.s);
    load_chunk(cur);
    store_x(cur.s);
#pragma unroll
    for (int j = 0; j < RPT; j++) {
      b[j] = nb[j];
      e[j] = ne[j];
    }
  }
  __syncthreads();
  while (cur.s < s1) {
    // products of the current chunk (x of its slice is in LDS)
#pragma unroll
    for (int i = 0; i < kItems / 2; i++) {
      double a0, a1;
      widen(v[i], a0, a1);
      f64x2 pr;
      pr.x = mul_rn(a0, xs[cw[i] & 0xffffu]);
      pr.y = mul_rn(a1, xs[cw[i] >> 16]);
      *reinterpret_cast<f64x2 *>(prod + 512 * i + 2 * t) = pr;  // 16-byte LDS stores
    }
    Cursor nxt = cur;
    nxt.k0 += kChunk;
    if (nxt.k0 >= cur.len) {  // the next slice starts where this one ends
      nxt.s = cur.s + 1;
      nxt.k0 = 0;
      nxt.base = cur.base + cur.len;
      if (nxt.s < s1) {
        const int lo = __builtin_amdgcn_readfirstlane(static_cast<int>(end2));
        const int hi = __builtin_amdgcn_readfirstlane(static_cast<int>(end2 >> 32));
        const int64_t e2 = static_cast<int64_t>((static_cast<uint64_t>(static_cast<uint32_t>(hi)) << 32) |
                                                static_cast<uint32_t>(lo));
        nxt.len = static_cast<int>(e2 - nxt.base);
        if (nxt.len <= 0) seek(nxt, nxt.s + 1, s1, slice_off);  // empty slices
      }
    }
    const bool new_slice = nxt.s != cur.s && nxt.s < s1;
    // the next chunk's loads are in flight during the row sums below
    if (new_slice) {
      load_end2(nxt.s);
      load_x(nxt.s);
      load_ranges(nxt.s);
      load_chunk(nxt);
      __syncthreads();  // products visible; x of the current slice no longer read
      store_x(nxt.s);
    } else {
      if (nxt.s < s1) load_chunk(nxt);
      __syncthreads();
    }
    // each row adds its products of this chunk in element order, four LDS
    // reads issued ahead of four adds; a place past the row's end adds -0.0,
    // which leaves every sum (signed zeros included) unchanged
#pragma unroll
    for (int j = 0; j < RPT; j++) {
      const int lo = max(b[j], cur.k0) - cur.k0, hi = min(e[j], cur.k0 + kChunk) - cur.k0;
      double a = acc[j];
      for (int k = lo; k < hi; k += 4) {
        double q[4];
#pragma unroll
        for (int i = 0; i < 4; i++) q[i] = prod[min(k + i, kChunk - 1)];
#pragma unroll
        for (int i = 0; i < 4; i++) a = add_rn(a, (k + i < hi) ? q[i] : -0.0);
      }
      acc[j] = a;
    }
    if (new_slice) {
#pragma unroll
      for (int j = 0; j < RPT; j++) {
        b[j] = nb[j];
        e[j] = ne[j];
      }
    }
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
