// panel_device.hpp -- device body of the column-panel form of kernel 1, shared
// by the single-matrix launch (sblas_panel.hip) and the batched persistent
// launch (sblas_batch.hip).  Layout and order of operations are described in
// sblas_panel.hip.
#pragma once

#include <type_traits>

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

// ---- segmented form (order-relaxed) ----------------------------------------

// Each chunk of kSegChunk = 256*E elements of a slice (E = kSegItems; elements
// in row-major order, the last chunk shorter, padded to E) is stored
// transposed in pairs: thread t of the chunk owns elements E*t .. E*t+E-1, its
// pair i (elements E*t+2i, E*t+2i+1) stored at chunk_base + 2*(i*T + t)
// (T = the chunk's length / E), so every load instruction reads T consecutive
// pairs.  One flag word of E bits per
// thread marks which of its elements start a row.  A thread
// sums its products in registers; the value of a row that crosses threads is
// combined by a segmented scan over the wave on DPP lane moves, across the
// four waves through summaries in LDS, and across chunks in registers, so each
// row of a slice is closed exactly once and added to its running sum in LDS.
// The order of additions inside a row is fixed (results repeat bit for bit
// from run to run) but differs from the host loop's sequential order.

static_assert(kSegItems == 8 || kSegItems == 16, "a thread's flags are one byte or one 16-bit word");
static_assert(kPanelPad % 2 == 0 && kSegItems % kPanelPad == 0, "segmented slices hold whole groups");
struct SegWave {                   // one wave's part of a chunk
  double open;                     // value of the row open at the wave's end (if has)
  double tot;                      // sum carried through a wave without row starts
  int cnt;                         // row starts in the wave
  int has;
};
constexpr int kSegWaves = kSegThreads / 64;
constexpr int kSegLdsBytes = (kPanelCols + kSegRowsMax) * 8 + 2 * kSegWaves * sizeof(SegWave) +
                             2 * kSegRowsMax * 2;
static_assert(kPanelCols % (2 * kSegThreads) == 0, "x panel staged in whole pairs");

template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_d(double v) {
  return __builtin_amdgcn_update_dpp(0.0, v, CTRL, ROWS, 0xf, false);
}
template <int CTRL, int ROWS>
__device__ __forceinline__ int dpp_i(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, ROWS, 0xf, false);
}
// inclusive prefix sum over the 64 lanes (row_shr inside 16-lane rows, then
// row_bcast15 / row_bcast31 across them; missing sources read 0)
__device__ __forceinline__ int seg_wave_incl(int v) {
  v += dpp_i<0x111, 0xf>(v);
  v += dpp_i<0x112, 0xf>(v);
  v += dpp_i<0x114, 0xf>(v);
  v += dpp_i<0x118, 0xf>(v);
  v += dpp_i<0x142, 0xa>(v);
  v += dpp_i<0x143, 0xc>(v);
  return v;
}
template <int CTRL, int ROWS>
__device__ __forceinline__ void seg_step(double &v, int &head) {
  const double vu = dpp_d<CTRL, ROWS>(v);
  const int hu = dpp_i<CTRL, ROWS>(head);
  if (!head) {
    v += vu;
    head = hu;
  }
}
// inclusive segmented sum over the 64 lanes; head = 1 opens a segment
__device__ __forceinline__ double seg_wave_sum(double v, int head) {
  seg_step<0x111, 0xf>(v, head);
  seg_step<0x112, 0xf>(v, head);
  seg_step<0x114, 0xf>(v, head);
  seg_step<0x118, 0xf>(v, head);
  seg_step<0x142, 0xa>(v, head);
  seg_step<0x143, 0xc>(v, head);
  return v;
}

// Adds a closed row's value of one slice to the row's sum in LDS.  A row is
// closed once per slice and slices are separated by barriers, so the adds to
// one row come in slice order; the LDS atomic (no value returned) keeps the
// read of the old sum off the thread's dependency chain.
__device__ __forceinline__ void seg_close(double *s, double v) {
  __hip_atomic_fetch_add(s, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Row panel p in the segmented form by one workgroup of kSegThreads threads.  lds
// holds kSegLdsBytes; R <= kSegRowsMax.  Same slot / y conventions as
// row_panel.  The workgroup leaves through a barrier.
template <typename VT>
__device__ __forceinline__ void row_panel_seg(
    int p, int n, int R, const PanelInfo &pi, const int *__restrict__ slice_base,
    const int64_t *__restrict__ slice_off, const uint8_t *__restrict__ flags,
    const uint16_t *__restrict__ rownz, const VT *__restrict__ pval,
    const uint16_t *__restrict__ pcol, const double *__restrict__ x, double *__restrict__ y,
    double alpha, double beta, double *__restrict__ slots, char *lds) {
  constexpr int W = kPanelCols;
  double *xs = reinterpret_cast<double *>(lds);
  double *sum = xs + W;
  SegWave *ws = reinterpret_cast<SegWave *>(sum + kSegRowsMax);  // two sets, one per wave
  uint16_t *rzb = reinterpret_cast<uint16_t *>(ws + 2 * kSegWaves);  // two slices' row ids
  constexpr int NT = kSegThreads;
  const int t = threadIdx.x;
  const int wave = t >> 6, lane = t & 63;
  const int s0 = slice_base[p], s1 = slice_base[p + 1];
  const int qb = pi.q_lo;
  for (int r = t; r < R; r += NT) sum[r] = 0.0;

  constexpr int E = kSegItems;
  typedef typename std::conditional<E == 16, uint16_t, uint8_t>::type FlagT;
  const FlagT *__restrict__ fw = reinterpret_cast<const FlagT *>(flags);
  double v[E];
  int c[E], f = 0;
  auto stage = [&](int s, uint16_t *rz) {  // x and the slice's rows with elements
    const int c0 = (qb + (s - s0)) * W;
#pragma unroll
    for (int i = 0; i < W / (2 * NT); i++) {
      const int cc = c0 + 2 * t + i * 2 * NT;
      const double x0 = x[min(cc, n - 1)], x1 = x[min(cc + 1, n - 1)];  // no load under a branch
      f64x2 q;
      q.x = (cc < n) ? x0 : 0.0;
      q.y = (cc + 1 < n) ? x1 : 0.0;
      *reinterpret_cast<f64x2 *>(xs + 2 * t + i * 2 * NT) = q;
    }
    for (int i = t; i < R; i += NT) rz[i] = rownz[static_cast<int64_t>(s) * R + i];
  };
  auto load_chunk = [&](const Cursor &cr) {
    const int T = min(kSegChunk, cr.len - cr.k0) / E;
    const int tt = min(t, T - 1);  // threads past the chunk re-read its last group
    const int64_t cb = cr.base + cr.k0;
    const VT *__restrict__ vb = pval + cb;
    const uint16_t *__restrict__ cbp = pcol + cb;
#pragma unroll
    for (int i = 0; i < E / 2; i++) {  // pair i of the thread: one load each
      const int k = 2 * (i * T + tt);
      load2(vb + k, v[2 * i], v[2 * i + 1]);
      load2_cols(cbp + k, c[2 * i], c[2 * i + 1]);
    }
    f = fw[cb / E + tt];
    if (t >= T) f = 0;
  };
  auto advance = [&](Cursor cr) {
    cr.k0 += kSegChunk;
    if (cr.k0 >= cr.len) seek(cr, cr.s + 1, s1, slice_off);
    return cr;
  };

  Cursor cur;
  seek(cur, s0, s1, slice_off);
  int rb = 0;  // row-id buffer of the current slice
  if (cur.s < s1) {
    stage(cur.s, rzb);
    load_chunk(cur);
  }
  __syncthreads();
  double carry = 0.0;  // value of the row open when the chunk begins
  int rank0 = 0;       // row starts in the slice before the chunk
  int buf = 0;
  while (cur.s < s1) {
    const int T = min(kSegChunk, cur.len - cur.k0) / E;
    const bool act = t < T;
    double pr[E];
#pragma unroll
    for (int i = 0; i < E; i++) {  // every gather issued, then the products selected
      const double xv = xs[c[i] & (W - 1)];  // padding (column 0xffff) reads an entry in range
      const double pm = mul_rn(v[i], xv);
      pr[i] = (act && c[i] != 0xffff) ? pm : 0.0;
    }
    const int fl = f;
    // in-thread walk: the part before the first start (a) and after the last
    double cur_s = 0.0, first = 0.0;
    int seen = 0;
#pragma unroll
    for (int i = 0; i < E; i++) {
      if ((fl >> i) & 1) {
        if (seen == 0) first = cur_s;
        seen++;
        cur_s = 0.0;
      }
      cur_s += pr[i];
    }
    const bool hs = seen > 0;
    const double a = hs ? first : cur_s;
    const double last = hs ? cur_s : 0.0;
    // wave scans: row-start counts, and the open row's value across lanes
    const int incl = seg_wave_incl(seen);
    const unsigned long long smask = __ballot(hs);
    const int head = (lane == 0) ? 1 : static_cast<int>((smask >> (lane - 1)) & 1ull);
    const double sv = seg_wave_sum(a, head);
    const unsigned long long below = smask & ((1ull << lane) - 1ull);
    const int prev = below ? 63 - __clzll(static_cast<long long>(below)) : 0;
    const double prev_last = __shfl(last, prev, 64);
    const int llast = smask ? 63 - __clzll(static_cast<long long>(smask)) : 0;
    const double open_last = __shfl(last, llast, 64);
    const double sv63 = __shfl(sv, 63, 64);
    const int cnt63 = __shfl(incl, 63, 64);
    if (lane == 0) {
      SegWave w;
      w.cnt = cnt63;
      w.has = smask != 0ull;
      w.open = open_last + (llast < 63 ? sv63 : 0.0);
      w.tot = sv63;
      ws[buf * kSegWaves + wave] = w;
    }
    const Cursor nxt = advance(cur);
    if (nxt.s < s1) load_chunk(nxt);  // in flight during the closes below
    const bool slice_end = nxt.s != cur.s;
    __syncthreads();  // wave summaries visible; x of this chunk read
    const uint16_t *rz = rzb + rb * kSegRowsMax;
    // the next slice's x and row ids (row ids into the other buffer: this
    // slice's are still read by the closes)
    if (slice_end && nxt.s < s1) stage(nxt.s, rzb + (rb ^ 1) * kSegRowsMax);
    double cin = carry, my_cin = 0.0;
    int rk = rank0, my_rank = 0;
#pragma unroll
    for (int w2 = 0; w2 < kSegWaves; w2++) {
      const SegWave w = ws[buf * kSegWaves + w2];
      if (w2 == wave) {
        my_cin = cin;
        my_rank = rk;
      }
      cin = w.has ? w.open : cin + w.tot;
      rk += w.cnt;
    }
    // the row closed by this thread's first start, then rows inside the thread
    if (hs) {
      const int rank = my_rank + incl - seen - 1;  // rank of the row open before the start
      const double val = sv + (below ? prev_last : my_cin);
      // a start at element i closes the row of rank (rank + starts before i);
      // every row id read first (rank + j < the slice's row count <= R), then
      // the adds
      int rid[E];
      int j = 0;
#pragma unroll
      for (int i = 0; i < E; i++) {
        rid[i] = rz[max(rank + j, 0)];
        j += (fl >> i) & 1;
      }
      double cs = 0.0;
      j = 0;
#pragma unroll
      for (int i = 0; i < E; i++) {
        if ((fl >> i) & 1) {
          if (j > 0 || rank >= 0) seg_close(sum + rid[i], j > 0 ? cs : val);
          j++;
          cs = 0.0;
        }
        cs += pr[i];
      }
    }
    carry = cin;
    rank0 = rk;
    buf ^= 1;
    if (slice_end) {
      if (t == 0 && rank0 > 0) seg_close(sum + rz[rank0 - 1], carry);  // the slice's last row
      carry = 0.0;
      rank0 = 0;
      rb ^= 1;
      __syncthreads();  // x and row ids of the next slice visible
    }
    cur = nxt;
  }
  __syncthreads();
  for (int rl = t; rl < pi.nrow; rl += NT) {
    const double s_ = sum[rl];
    if (rl == 0 && (pi.flags & 1)) {
      slots[0] = s_;
    } else if (rl == pi.nrow - 1 && (pi.flags & 2)) {
      slots[1] = s_;
    } else {
      const int r = pi.row0 + rl;
      y[r] = (beta == 0.0) ? alpha * s_ : alpha * s_ + beta * y[r];
    }
  }
  __syncthreads();
}

}  // namespace panel
}  // namespace sblas
