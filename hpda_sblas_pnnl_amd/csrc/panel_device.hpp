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
// Elements per thread per chunk of the narrow-x body: FP32 products let a
// chunk of twice the elements fit the LDS of an FP64 chunk, so as many bytes
// are in flight per workgroup as with FP64 values.
constexpr int kItemsNarrow = 16;
template <bool NX>
__host__ __device__ constexpr int chunk_elems() {
  return 256 * (NX ? kItemsNarrow : kItems);
}
static_assert(kPanelPad % 2 == 0, "pairs never straddle a slice end");

// A slice's row offsets are staged in LDS as 32-bit words, in rounds of 256
// (one per thread): 16-bit offsets as the words covering the slice's R + 1
// entries (one word more when they start mid-word), 32-bit offsets one word
// per entry.
template <typename OT, int RPT>
__host__ __device__ constexpr int range_rounds() {
  return sizeof(OT) == 2 ? (128 * RPT + 1 + 255) / 256 : (256 * RPT + 1 + 255) / 256;
}
// LDS of one row panel body, in doubles: x panel, products, row offsets.
// 16-bit offsets with up to 8 rows per thread keep three workgroups per CU.
// With NX (narrow x: FP32 x window and FP32 products) the first two halve.
template <typename OT, int RPT, bool NX = false>
__host__ __device__ constexpr int lds_doubles() {
  return (NX ? (kPanelCols + chunk_elems<true>()) / 2 : kPanelCols + kChunk) +
         range_rounds<OT, RPT>() * 128;
}

// s_waitcnt vmcnt(0) alone.  gfx9-family encoding of the immediate: vmcnt in
// bits 3:0 and 15:14, expcnt in 6:4, lgkmcnt in 11:8; expcnt (7) and lgkmcnt
// (15) are left at their maxima, so only vector memory is waited for.  The
// compiler's own wait insertion takes this wait into account.
__device__ __forceinline__ void wait_vector_memory() { __builtin_amdgcn_s_waitcnt(0x0F70); }

// One load of 16 (4) bytes per thread from global memory straight into LDS:
// the wave's 64 pieces land contiguously at lds_wave, which must be the same
// for every thread of the wave.  Tracked like a vector load (vmcnt).
typedef const __attribute__((address_space(1))) void *global_ptr;
typedef __attribute__((address_space(3))) void *lds_ptr;
__device__ __forceinline__ void load16_to_lds(const void *src, void *lds_wave) {
  __builtin_amdgcn_global_load_lds((global_ptr)src, (lds_ptr)lds_wave, 16, 0, 0);
}
__device__ __forceinline__ void load4_to_lds(const void *src, void *lds_wave) {
  __builtin_amdgcn_global_load_lds((global_ptr)src, (lds_ptr)lds_wave, 4, 0, 0);
}

typedef double f64x2 __attribute__((ext_vector_type(2)));
typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

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
// the same pairs as FP32 (exact for FP32 and BF16 values)
__device__ __forceinline__ void widen_f(f32x2 q, float &a, float &b) {
  a = q.x;
  b = q.y;
}
__device__ __forceinline__ void widen_f(uint32_t w, float &a, float &b) {
  a = __uint_as_float(w << 16);
  b = __uint_as_float(w & 0xffff0000u);
}
__device__ __forceinline__ void widen_f(f64x2 q, float &a, float &b) {
  a = static_cast<float>(q.x);
  b = static_cast<float>(q.y);
}
__device__ __forceinline__ float mul_rn_f(float a, float b) {
#pragma clang fp contract(off)
  return a * b;
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
// lds_doubles<OT, RPT>() doubles; slots[0] / slots[1] receive the partial sums
// of a first / last row that the panel's element window cuts (PanelInfo
// flags).  Software-pipelined: the next chunk's values and column offsets are
// loaded while the current chunk's rows are summed.  At a slice change the
// next slice's x window is loaded straight into LDS (no registers, no stores)
// once the current slice's products are formed, and lands while the rows are
// summed; its row ranges were loaded into registers at the slice change
// before (through LDS with x for the first slice and after an empty one).
// The workgroup leaves through a barrier, so LDS may be reused right after.
//
// NX (narrow x, for FP32 / BF16 values): x is staged from an FP32 copy
// (x32, padded with zeros to whole column panels), products are rounded to
// FP32 and kept as FP32 in LDS, and each row's products of a chunk are added
// in FP64 into two accumulators (even and odd places) that join the row's sum
// once per chunk.  Half the LDS traffic per element and half the dependent
// adds, at an error of a few FP32 roundings per product (x and the product)
// instead of bit-exact sums.
template <typename VT, int RPT, typename OT, bool NX = false>
__device__ __forceinline__ void row_panel(
    int p, int n, int R, const PanelInfo &pi, const int *__restrict__ slice_base,
    const int64_t *__restrict__ slice_off, const OT *__restrict__ srp,
    const VT *__restrict__ pval, const uint16_t *__restrict__ pcol,
    const double *__restrict__ x, double *__restrict__ y, double alpha, double beta,
    double *__restrict__ slots, double *lds, const float *__restrict__ x32 = nullptr) {
  constexpr int W = kPanelCols;
  constexpr int NR = range_rounds<OT, RPT>();
  constexpr int IT = NX ? kItemsNarrow : kItems;  // elements per thread per chunk
  constexpr int CH = chunk_elems<NX>();           // elements per chunk
  typedef typename RawPair<VT>::type RV;
  double *xs = lds;
  double *prod = lds + W;
  float *xsf = reinterpret_cast<float *>(lds);
  float *prodf = reinterpret_cast<float *>(lds + W / 2);
  uint32_t *rs = reinterpret_cast<uint32_t *>(lds + (NX ? (W + CH) / 2 : W + CH));
  const int t = threadIdx.x;
  const int wv = t & ~63;  // first thread of this thread's wave
  const int s0 = slice_base[p], s1 = slice_base[p + 1];
  const int qb = pi.q_lo;
  // x in 16-byte pairs when aligned (else in 4-byte pieces)
  const bool x16 = (reinterpret_cast<uintptr_t>(x) & 15) == 0 && n >= 2;
  double acc[RPT];
  int b[RPT], e[RPT];     // element range [b, e) of row t + 256 j in the current slice
  int nb[RPT], ne[RPT];   // the same in slice r_next, loaded a whole slice ahead
  int r_next = -1;
  RV v[IT / 2];           // the next chunk's value pairs
  uint32_t cw[IT / 2];    // and column offset pairs
  uint32_t end2 = 0;      // low word of slice_off at the end of the slice after the current one
  int r_shift = 0;        // first 16-bit offset of the staged slice inside its first word
  double xfix = 0.0;      // x[n - 1] when n is odd and this thread's pair holds it
  int fix_at = -1;
#pragma unroll
  for (int j = 0; j < RPT; j++) acc[j] = 0.0;

  // the end of the slice after s, read one slice ahead.  Odd threads read
  // the entry after it, which makes the address per-thread: a vector load,
  // which the chunk's waits cover, where a scalar load would be waited for at
  // the next barrier.  The value is taken from a wave's first (even) thread.
  auto load_end2 = [&](int s) {
    end2 = reinterpret_cast<const uint32_t *>(slice_off + min(s + 2 + (t & 1), s1))[0];
  };
  // slice s's x window x[q*W, (q+1)*W) and row offsets into LDS.  Pieces of x
  // past n load a valid address instead (their columns never occur); the pair
  // (n - 1, n) of an odd n would read past x, so its thread loads x[n - 1]
  // alone and writes it once the loads have landed.
  auto request_slice = [&](int s, bool ranges) {
    const int c0 = (qb + (s - s0)) * W;
    if constexpr (NX) {
      // four FP32 x per 16-byte piece; x32 is padded to whole panels
#pragma unroll
      for (int i = 0; i < W / 1024; i++)
        load16_to_lds(x32 + c0 + 4 * (i * 256 + t), xsf + 4 * (i * 256 + wv));
      fix_at = -1;
    } else if (x16) {
#pragma unroll
      for (int i = 0; i < W / 512; i++) {
        const int c = c0 + 2 * (i * 256 + t);
        load16_to_lds(c + 1 < n ? x + c : x, xs + 2 * (i * 256 + wv));
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
        load4_to_lds(xw + min(g0 + i * 256 + t, gmax),
                     reinterpret_cast<uint32_t *>(xs) + i * 256 + wv);
      fix_at = -1;
    }
    const int64_t base = static_cast<int64_t>(s) * (R + 1);  // first offset of slice s
    const uint32_t *src;
    int cnt;
    if constexpr (sizeof(OT) == 2) {
      src = reinterpret_cast<const uint32_t *>(srp) + (base >> 1);
      cnt = static_cast<int>(((base + R) >> 1) - (base >> 1)) + 1;
      r_shift = static_cast<int>(base & 1);
    } else {
      src = reinterpret_cast<const uint32_t *>(srp) + base;
      cnt = R + 1;
      r_shift = 0;
    }
    if (ranges) {
#pragma unroll
      for (int r = 0; r < NR; r++)
        if (r * 256 < cnt) load4_to_lds(src + min(r * 256 + t, cnt - 1), rs + r * 256 + wv);
    }
  };
  // row ranges of slice s into registers, a whole slice before they are used
  auto load_next_ranges = [&](int s) {
    const OT *rps = srp + static_cast<int64_t>(s) * (R + 1);
#pragma unroll
    for (int j = 0; j < RPT; j++) {
      const int rl = t + 256 * j;
      nb[j] = static_cast<int>(rps[min(rl, R)]);
      ne[j] = static_cast<int>(rps[min(rl + 1, R)]);
    }
    r_next = s;
  };
  // after the requests have landed (this thread's): the odd-n fix-up; after
  // the barrier that follows: each thread's row ranges, a row past R reading
  // the slice's end twice (an empty range)
  auto finish_slice = [&]() {
    wait_vector_memory();
    if (fix_at >= 0) xs[fix_at] = xfix;
  };
  auto read_ranges = [&]() {
#pragma unroll
    for (int j = 0; j < RPT; j++) {
      const int rl = t + 256 * j;
      if constexpr (sizeof(OT) == 2) {
        const uint16_t *r16 = reinterpret_cast<const uint16_t *>(rs) + r_shift;
        b[j] = r16[min(rl, R)];
        e[j] = r16[min(rl + 1, R)];
      } else {
        b[j] = static_cast<int>(rs[min(rl, R)]);
        e[j] = static_cast<int>(rs[min(rl + 1, R)]);
      }
    }
  };
  // values and column offsets of a chunk: pair i of thread t is elements
  // 512*i + 2t, +1 (so one load instruction of a wave is a contiguous run of
  // whole lines), all issued back to back; threads past the slice's end
  // re-read its last pair (never summed)
  auto load_chunk = [&](const Cursor &c) {
#pragma unroll
    for (int i = 0; i < IT / 2; i++) {
      const int64_t k = c.base + min(c.k0 + 512 * i + 2 * t, c.len - 2);
      v[i] = load_pair(pval + k);
      cw[i] = load_cols(pcol + k);
    }
  };

  Cursor cur;
  seek(cur, s0, s1, slice_off);
  if (cur.s < s1) {
    load_end2(cur.s);
    request_slice(cur.s, true);
    load_chunk(cur);
    finish_slice();
  }
  __syncthreads();
  if (cur.s < s1) {
    read_ranges();
    if (cur.s + 1 < s1) load_next_ranges(cur.s + 1);
  }
  while (cur.s < s1) {
    // products of the current chunk (x of its slice is in LDS)
#pragma unroll
    for (int i = 0; i < IT / 2; i++) {
      if constexpr (NX) {
        float a0, a1;
        widen_f(v[i], a0, a1);
        f32x2 pr;
        pr.x = mul_rn_f(a0, xsf[cw[i] & 0xffffu]);
        pr.y = mul_rn_f(a1, xsf[cw[i] >> 16]);
        *reinterpret_cast<f32x2 *>(prodf + 512 * i + 2 * t) = pr;  // 8-byte LDS stores
      } else {
        double a0, a1;
        widen(v[i], a0, a1);
        f64x2 pr;
        pr.x = mul_rn(a0, xs[cw[i] & 0xffffu]);
        pr.y = mul_rn(a1, xs[cw[i] >> 16]);
        *reinterpret_cast<f64x2 *>(prod + 512 * i + 2 * t) = pr;  // 16-byte LDS stores
      }
    }
    Cursor nxt = cur;
    nxt.k0 += CH;
    if (nxt.k0 >= cur.len) {  // the next slice starts where this one ends
      nxt.s = cur.s + 1;
      nxt.k0 = 0;
      nxt.base = cur.base + cur.len;
      if (nxt.s < s1) {
        // a slice's length fits an int: the low words' difference is exact
        const uint32_t e2 = static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(end2)));
        nxt.len = static_cast<int>(e2 - static_cast<uint32_t>(nxt.base));
        if (nxt.len <= 0) seek(nxt, nxt.s + 1, s1, slice_off);  // empty slices
      }
    }
    const bool new_slice = nxt.s != cur.s && nxt.s < s1;
    // the next chunk's loads are in flight during the row sums below
    if (nxt.s < s1) load_chunk(nxt);
    if (new_slice) load_end2(nxt.s);
    __syncthreads();  // products visible; x and row offsets of this slice no longer read
    if (new_slice) request_slice(nxt.s, r_next != nxt.s);
    // each row adds its products of this chunk in element order, four at a
    // time from two 16-byte LDS reads of aligned pairs (the run starts at the
    // pair holding its first product); a place before the row's first or
    // past its last product adds -0.0, which leaves every sum (signed zeros
    // included) unchanged
#pragma unroll
    for (int j = 0; j < RPT; j++) {
      const int lo = max(b[j], cur.k0) - cur.k0, hi = min(e[j], cur.k0 + CH) - cur.k0;
      if constexpr (NX) {
        // aligned runs of four FP32 products (one 16-byte LDS read), even
        // and odd places into two FP64 sums
        double a0 = 0.0, a1 = 0.0;
        for (int k = lo & ~3; k < hi; k += 4) {
          const f32x4 q = *reinterpret_cast<const f32x4 *>(prodf + k);
          a0 += static_cast<double>((k >= lo) ? q.x : -0.0f);
          a1 += static_cast<double>((k + 1 >= lo && k + 1 < hi) ? q.y : -0.0f);
          a0 += static_cast<double>((k + 2 >= lo && k + 2 < hi) ? q.z : -0.0f);
          a1 += static_cast<double>((k + 3 < hi) ? q.w : -0.0f);
        }
        acc[j] += a0 + a1;
      } else {
        double a = acc[j];
        for (int k = lo & ~1; k < hi; k += 4) {
          const f64x2 p0 = *reinterpret_cast<const f64x2 *>(prod + k);
          const f64x2 p1 = *reinterpret_cast<const f64x2 *>(prod + min(k + 2, kChunk - 2));
          a = add_rn(a, (k >= lo) ? p0.x : -0.0);
          a = add_rn(a, (k + 1 < hi) ? p0.y : -0.0);
          a = add_rn(a, (k + 2 < hi) ? p1.x : -0.0);
          a = add_rn(a, (k + 3 < hi) ? p1.y : -0.0);
        }
        acc[j] = a;
      }
    }
    if (new_slice) finish_slice();
    __syncthreads();  // products consumed; the next slice's x and row offsets in place
    if (new_slice) {
      if (r_next == nxt.s) {
#pragma unroll
        for (int j = 0; j < RPT; j++) {
          b[j] = nb[j];
          e[j] = ne[j];
        }
      } else {
        read_ranges();  // an empty slice lay between: staged with x instead
      }
      if (nxt.s + 1 < s1) load_next_ranges(nxt.s + 1);
    }
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
