// sblas_panel.hip -- column-panel form of kernel 1 for matrices whose rows
// draw x from a wide window (banded FEM-like matrices).
//
// When every element's x read is a separate L2 request (x window of a few
// hundred KB, far sparser than one element per cache line), the L2 request
// rate, not HBM, bounds a row-parallel kernel.  This form cuts the matrix
// into row panels of R rows and column panels of W columns.  A workgroup owns
// one row panel and walks its column panels in increasing order: it stages
// x[q*W, (q+1)*W) in LDS with coalesced loads (one pass over the panel's x
// lines), then streams that slice's elements in chunks of 2048 with
// consecutive threads on consecutive elements, multiplies each by its x from
// LDS and leaves the products in LDS.  The plan stores each (row panel,
// column panel) slice contiguously, rows in order, with 16-bit column offsets
// inside the panel and row offsets inside the slice (16-bit when every slice
// holds at most 65535 elements, else 32-bit).
//
// Each row is then summed by one thread from the staged products, element by
// element in ascending column order, with separately rounded multiplies and
// adds: the same operations in the same order as the reference's host loop
// (test/dspmv_test_smt.cu:45-51), so results equal the host oracle bit for bit.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <climits>

#include "internal.hpp"

namespace sblas {
namespace {

__global__ void k_fill_u32(int64_t count, uint32_t *__restrict__ a) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < count) a[i] = 0u;
}

__device__ __forceinline__ double pv(const double *p, int64_t j) {
  return __builtin_nontemporal_load(p + j);
}
__device__ __forceinline__ double pv(const float *p, int64_t j) {
  return static_cast<double>(__builtin_nontemporal_load(p + j));
}
__device__ __forceinline__ double pv(const uint16_t *p, int64_t j) {
  const uint32_t h = __builtin_nontemporal_load(p + j);
  return static_cast<double>(__uint_as_float(h << 16));
}

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

// Slice s of the current row panel and the first element of the chunk being
// worked on; advance() moves to the next chunk, skipping empty slices.
struct PanelCursor {
  int s, k0, len;
  int64_t base;
};

__device__ __forceinline__ void cursor_seek(PanelCursor &c, int s, int s1,
                                            const int64_t *__restrict__ slice_off) {
  c.k0 = 0;
  c.s = s;
  c.len = 0;
  for (; c.s < s1; c.s++) {
    c.base = slice_off[c.s];
    c.len = static_cast<int>(slice_off[c.s + 1] - c.base);
    if (c.len > 0) return;
  }
}

template <typename VT, int RPT, typename OT>
__global__ __launch_bounds__(256) void k_panel_spmv(
    int m, int n, int R, int npanels, int xcd_remap, const int *__restrict__ q_lo,
    const int *__restrict__ slice_base, const int64_t *__restrict__ slice_off,
    const OT *__restrict__ srp, const VT *__restrict__ pval,
    const uint16_t *__restrict__ pcol, const double *__restrict__ x, double *__restrict__ y,
    double alpha, double beta) {
  constexpr int W = kPanelCols;
  constexpr int ITEMS = 8;
  constexpr int CH = 256 * ITEMS;  // elements per chunk of a slice
  __shared__ double xs[W];
  __shared__ double prod[CH];
  int p = blockIdx.x;
  if (xcd_remap) {
    const int per = gridDim.x >> 3;
    p = (p & 7) * per + (p >> 3);
  }
  if (p >= npanels) return;  // uniform per workgroup
  const int t = threadIdx.x;
  const int r0 = p * R;
  const int s0 = slice_base[p], s1 = slice_base[p + 1];
  const int qb = q_lo[p];
  double acc[RPT];
  int b[RPT], e[RPT];
  double v[ITEMS];
  int l[ITEMS];
#pragma unroll
  for (int j = 0; j < RPT; j++) acc[j] = 0.0;

  // x[q*W, (q+1)*W) of slice s into LDS
  auto stage_x = [&](int s) {
    const int c0 = (qb + (s - s0)) * W;
#pragma unroll
    for (int i = 0; i < W / 256; i++) {
      const int c = c0 + t + i * 256;
      const double xv = x[min(c, n - 1)];  // no load under a branch
      xs[t + i * 256] = (c < n) ? xv : 0.0;
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
  // values and column offsets of chunk c: consecutive threads on consecutive
  // elements (whole lines per load instruction), all issued back to back
  auto load_chunk = [&](const PanelCursor &c) {
#pragma unroll
    for (int i = 0; i < ITEMS; i++) {
      const int64_t k = c.base + min(c.k0 + t + i * 256, c.len - 1);
      v[i] = pv(pval, k);
      l[i] = __builtin_nontemporal_load(pcol + k);
    }
  };

  PanelCursor cur;
  cursor_seek(cur, s0, s1, slice_off);
  if (cur.s < s1) {
    stage_x(cur.s);
    row_ranges(cur.s);
    load_chunk(cur);
  }
  __syncthreads();
  while (cur.s < s1) {
    // products of the current chunk (x of its slice is in LDS)
#pragma unroll
    for (int i = 0; i < ITEMS; i++) prod[t + i * 256] = mul_rn(v[i], xs[l[i]]);
    PanelCursor nxt = cur;
    nxt.k0 += CH;
    if (nxt.k0 >= cur.len) cursor_seek(nxt, cur.s + 1, s1, slice_off);
    // the next chunk's loads are in flight during the row sums below
    if (nxt.s < s1) load_chunk(nxt);
    __syncthreads();  // products visible; x of the current slice no longer read
    const bool new_slice = nxt.s != cur.s && nxt.s < s1;
    if (new_slice) stage_x(nxt.s);
    // each row adds its products of this chunk in element order
#pragma unroll
    for (int j = 0; j < RPT; j++) {
      const int lo = max(b[j], cur.k0), hi = min(e[j], cur.k0 + CH);
      double a = acc[j];
      for (int k = lo; k < hi; k++) a = add_rn(a, prod[k - cur.k0]);
      acc[j] = a;
    }
    if (new_slice) row_ranges(nxt.s);
    __syncthreads();  // products consumed; x of the next slice visible
    cur = nxt;
  }
#pragma unroll
  for (int j = 0; j < RPT; j++) {
    const int rl = t + 256 * j;
    const int r = r0 + rl;
    if (rl < R && r < m) {
      y[r] = (beta == 0.0) ? mul_rn(alpha, acc[j])
                           : add_rn(mul_rn(alpha, acc[j]), mul_rn(beta, y[r]));
    }
  }
}

// ---- format build ---------------------------------------------------------------

// Column-panel range of each row panel: first and last column panel any of
// its rows touches (columns are sorted within rows).
__global__ void k_panel_range(int m, int R, int npanels, const int *__restrict__ rp,
                              const int *__restrict__ col, int *__restrict__ q_lo,
                              int *__restrict__ q_n) {
  const int p = blockIdx.x;
  if (p >= npanels) return;
  __shared__ int smin[256], smax[256];
  int lo = INT_MAX, hi = -1;
  for (int r = p * R + threadIdx.x; r < min(m, (p + 1) * R); r += 256) {
    const int b = rp[r], e = rp[r + 1];
    if (e > b) {
      lo = min(lo, col[b]);
      hi = max(hi, col[e - 1]);
    }
  }
  smin[threadIdx.x] = lo;
  smax[threadIdx.x] = hi;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) {
      smin[threadIdx.x] = min(smin[threadIdx.x], smin[threadIdx.x + w]);
      smax[threadIdx.x] = max(smax[threadIdx.x], smax[threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    if (smax[0] < 0) {  // no element in the panel
      q_lo[p] = 0;
      q_n[p] = 0;
    } else {
      q_lo[p] = smin[0] / kPanelCols;
      q_n[p] = smax[0] / kPanelCols - smin[0] / kPanelCols + 1;
    }
  }
}

// Per row, element counts in each column panel of its row panel:
// cnt[(slice_base[p] + q - q_lo[p]) * R + local_row].
__global__ void k_panel_count(int m, int R, const int *__restrict__ rp,
                              const int *__restrict__ col, const int *__restrict__ q_lo,
                              const int *__restrict__ slice_base, uint32_t *__restrict__ cnt) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= m) return;
  const int p = r / R, rl = r % R;
  const int qb = q_lo[p], sb = slice_base[p];
  int cur = -1, c = 0;
  for (int j = rp[r]; j < rp[r + 1]; j++) {
    const int q = col[j] / kPanelCols;
    if (q != cur) {
      if (cur >= 0) cnt[static_cast<int64_t>(sb + cur - qb) * R + rl] = c;
      cur = q;
      c = 0;
    }
    c++;
  }
  if (cur >= 0) cnt[static_cast<int64_t>(sb + cur - qb) * R + rl] = c;
}

// Exclusive scan of each slice's R counts -> row offsets (R+1 per slice) and
// the slice's element count; flags slices above 65535 elements.  One
// workgroup per slice; thread t scans rows t*RPT .. t*RPT+RPT-1 (R <= 256*RPT).
template <int RPT>
__global__ void k_panel_slice_scan(int nslices, int R, const uint32_t *__restrict__ cnt,
                                   uint32_t *__restrict__ srp, int64_t *__restrict__ snnz,
                                   int *__restrict__ overflow) {
  const int s = blockIdx.x;
  if (s >= nslices) return;
  typedef hipcub::BlockScan<uint32_t, 256> Scan;
  __shared__ typename Scan::TempStorage tmp;
  uint32_t v[RPT];
  uint32_t local = 0;
#pragma unroll
  for (int j = 0; j < RPT; j++) {
    const int rl = threadIdx.x * RPT + j;
    v[j] = rl < R ? cnt[static_cast<int64_t>(s) * R + rl] : 0u;
    local += v[j];
  }
  uint32_t excl = 0, total = 0;
  Scan(tmp).ExclusiveSum(local, excl, total);
  uint32_t run = excl;
#pragma unroll
  for (int j = 0; j < RPT; j++) {
    const int rl = threadIdx.x * RPT + j;
    if (rl < R) srp[static_cast<int64_t>(s) * (R + 1) + rl] = run;
    run += v[j];
  }
  if (threadIdx.x == 0) {
    srp[static_cast<int64_t>(s) * (R + 1) + R] = total;
    snnz[s] = total;
    if (total > 65535u) overflow[0] = 1;
  }
}

__global__ void k_narrow_u16(int64_t count, const uint32_t *__restrict__ a,
                             uint16_t *__restrict__ b) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < count) b[i] = static_cast<uint16_t>(a[i]);
}

template <typename VT>
__global__ void k_panel_scatter(int m, int R, const int *__restrict__ rp,
                                const int *__restrict__ col, const VT *__restrict__ val,
                                const int *__restrict__ q_lo, const int *__restrict__ slice_base,
                                const int64_t *__restrict__ slice_off,
                                const uint32_t *__restrict__ srp, VT *__restrict__ pval,
                                uint16_t *__restrict__ pcol) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= m) return;
  const int p = r / R, rl = r % R;
  const int qb = q_lo[p], sb = slice_base[p];
  int cur = -1;
  int64_t pos = 0;
  for (int j = rp[r]; j < rp[r + 1]; j++) {
    const int c = col[j];
    const int q = c / kPanelCols;
    if (q != cur) {
      cur = q;
      const int s = sb + q - qb;
      pos = slice_off[s] + srp[static_cast<int64_t>(s) * (R + 1) + rl];
    }
    pval[pos] = val[j];
    pcol[pos] = static_cast<uint16_t>(c - q * kPanelCols);
    pos++;
  }
}

}  // namespace

int panel_build(int val_type, int rows, int m, int n, int64_t nnz, const int *rp, const int *col,
                const void *val, bool force, int min_panels, PanelFormat &f, bool *chosen,
                hipStream_t s) {
  (void)n;
  *chosen = false;
  if (rows < 1 || rows > kPanelRowsMax) {
    set_error(SBLAS_ERR_ARGS, "panel format: rows per panel must be in [1, 2048]");
    return SBLAS_ERR_ARGS;
  }
  const int R = rows;
  f.rows = R;
  f.rpt = panel_rpt(R);
  f.npanels = (m + R - 1) / R;
  f.buffers.clear();
  auto alloc = [&](size_t bytes) -> void * {
    void *p = nullptr;
    if (hipMalloc(&p, bytes ? bytes : 8) != hipSuccess) return nullptr;
    f.buffers.push_back(p);
    return p;
  };
  const int np = std::max(f.npanels, 1);
  f.q_lo = static_cast<int *>(alloc(sizeof(int) * np));
  int *q_n = static_cast<int *>(alloc(sizeof(int) * np));
  f.slice_base = static_cast<int *>(alloc(sizeof(int) * (np + 1)));
  if (!f.q_lo || !q_n || !f.slice_base) {
    set_error(SBLAS_ERR_ALLOC, "panel format: allocation failed");
    return SBLAS_ERR_ALLOC;
  }
  if (f.npanels > 0)
    hipLaunchKernelGGL(k_panel_range, dim3(f.npanels), dim3(256), 0, s, m, R, f.npanels, rp, col,
                       f.q_lo, q_n);
  std::vector<int> hq(np), hb(np + 1, 0);
  SBLAS_HIP_TRY(hipMemcpyAsync(hq.data(), q_n, sizeof(int) * np, hipMemcpyDeviceToHost, s));
  SBLAS_HIP_TRY(hipStreamSynchronize(s));
  int64_t total = 0;
  for (int p = 0; p < f.npanels; p++) {
    hb[p] = static_cast<int>(total);
    total += hq[p];
  }
  hb[f.npanels] = static_cast<int>(total);
  if (total > (int64_t{1} << 30)) {
    set_error(SBLAS_ERR_UNSUPPORTED, "panel format: too many slices");
    return SBLAS_ERR_UNSUPPORTED;
  }
  f.nslices = static_cast<int>(total);
  f.avg_panels_per_row_panel = f.npanels > 0 ? static_cast<double>(total) / f.npanels : 0.0;
  f.avg_slice = total > 0 ? static_cast<double>(nnz) / total : 0.0;
  // worth it when there are enough row panels to fill the GPU, slices are
  // dense enough that one coalesced pass over a column panel of x costs at
  // most a quarter of an x line per element, and a row-parallel block
  // (2048 elements) would touch fewer than one element per x line: its
  // window is avg_panels * 4096 columns = avg_panels * 256 lines.
  const bool helps = f.npanels >= min_panels && f.avg_slice >= 1024.0 &&
                     f.avg_panels_per_row_panel > 8.0;
  if (!force && !helps) {
    for (void *p : f.buffers) (void)hipFree(p);
    f.buffers.clear();
    f.npanels = 0;
    return SBLAS_OK;
  }
  SBLAS_HIP_TRY(hipMemcpyAsync(f.slice_base, hb.data(), sizeof(int) * (np + 1),
                               hipMemcpyHostToDevice, s));
  const int64_t ns = std::max<int64_t>(f.nslices, 1);
  uint32_t *cnt = static_cast<uint32_t *>(alloc(sizeof(uint32_t) * ns * R));
  f.srp = nullptr;
  f.srp32 = static_cast<uint32_t *>(alloc(sizeof(uint32_t) * ns * (R + 1)));
  int64_t *snnz = static_cast<int64_t *>(alloc(sizeof(int64_t) * (ns + 1)));
  f.slice_off = static_cast<int64_t *>(alloc(sizeof(int64_t) * (ns + 1)));
  int *overflow = static_cast<int *>(alloc(sizeof(int)));
  const size_t vbytes = val_type == SBLAS_VAL_F64 ? 8 : (val_type == SBLAS_VAL_F32 ? 4 : 2);
  f.pval = alloc(vbytes * std::max<int64_t>(nnz, 1));
  f.pcol = static_cast<uint16_t *>(alloc(sizeof(uint16_t) * std::max<int64_t>(nnz, 1)));
  if (!cnt || !f.srp32 || !snnz || !f.slice_off || !overflow || !f.pval || !f.pcol) {
    set_error(SBLAS_ERR_ALLOC, "panel format: allocation failed");
    return SBLAS_ERR_ALLOC;
  }
  // counts start at zero: written by a kernel (no memset nodes)
  const int64_t cnt_n = ns * R;
  hipLaunchKernelGGL(k_fill_u32, dim3(static_cast<unsigned>((cnt_n + 255) / 256)), dim3(256), 0, s,
                     cnt_n, cnt);
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(256), 0, s, int64_t{1},
                     reinterpret_cast<uint32_t *>(overflow));
  if (m > 0)
    hipLaunchKernelGGL(k_panel_count, dim3((m + 255) / 256), dim3(256), 0, s, m, R, rp, col,
                       f.q_lo, f.slice_base, cnt);
  if (f.nslices > 0) {
#define SBLAS_SCAN_CASE(RPT)                                                                  \
  case RPT:                                                                                   \
    hipLaunchKernelGGL(k_panel_slice_scan<RPT>, dim3(f.nslices), dim3(256), 0, s, f.nslices, R, \
                       cnt, f.srp32, snnz, overflow);                                         \
    break;
    switch (f.rpt) {
      SBLAS_SCAN_CASE(1)
      SBLAS_SCAN_CASE(2)
      SBLAS_SCAN_CASE(4)
      SBLAS_SCAN_CASE(6)
      SBLAS_SCAN_CASE(8)
    }
#undef SBLAS_SCAN_CASE
  }
  int h_over = 0;
  SBLAS_HIP_TRY(hipMemcpyAsync(&h_over, overflow, sizeof(int), hipMemcpyDeviceToHost, s));
  SBLAS_HIP_TRY(hipStreamSynchronize(s));
  if (!h_over) {
    // every slice fits 16-bit offsets: keep the narrow copy only
    const int64_t cnt_o = ns * (R + 1);
    f.srp = static_cast<uint16_t *>(alloc(sizeof(uint16_t) * cnt_o));
    if (!f.srp) {
      set_error(SBLAS_ERR_ALLOC, "panel format: allocation failed");
      return SBLAS_ERR_ALLOC;
    }
    hipLaunchKernelGGL(k_narrow_u16, dim3(static_cast<unsigned>((cnt_o + 255) / 256)), dim3(256),
                       0, s, cnt_o, f.srp32, f.srp);
  }
  // slice offsets: exclusive scan of slice sizes
  if (f.nslices > 0) {
    size_t tb = 0;
    SBLAS_HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, snnz, f.slice_off, f.nslices + 1, s));
    void *tmp = alloc(tb);
    if (!tmp) {
      set_error(SBLAS_ERR_ALLOC, "panel format: scan scratch");
      return SBLAS_ERR_ALLOC;
    }
    // snnz[nslices] must be 0 for the trailing total
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(256), 0, s, int64_t{2},
                       reinterpret_cast<uint32_t *>(snnz + f.nslices));
    SBLAS_HIP_TRY(hipcub::DeviceScan::ExclusiveSum(tmp, tb, snnz, f.slice_off, f.nslices + 1, s));
  }
  if (m > 0) {
    const dim3 g((m + 255) / 256), b(256);
    switch (val_type) {
      case SBLAS_VAL_F64:
        hipLaunchKernelGGL(k_panel_scatter<double>, g, b, 0, s, m, R, rp, col,
                           static_cast<const double *>(val), f.q_lo, f.slice_base, f.slice_off,
                           f.srp32, static_cast<double *>(f.pval), f.pcol);
        break;
      case SBLAS_VAL_F32:
        hipLaunchKernelGGL(k_panel_scatter<float>, g, b, 0, s, m, R, rp, col,
                           static_cast<const float *>(val), f.q_lo, f.slice_base, f.slice_off,
                           f.srp32, static_cast<float *>(f.pval), f.pcol);
        break;
      default:
        hipLaunchKernelGGL(k_panel_scatter<uint16_t>, g, b, 0, s, m, R, rp, col,
                           static_cast<const uint16_t *>(val), f.q_lo, f.slice_base, f.slice_off,
                           f.srp32, static_cast<uint16_t *>(f.pval), f.pcol);
        break;
    }
  }
  SBLAS_HIP_TRY(hipGetLastError());
  SBLAS_HIP_TRY(hipStreamSynchronize(s));
  // scratch of the build is no longer needed
  auto release = [&](void *p) {
    (void)hipFree(p);
    f.buffers.erase(std::find(f.buffers.begin(), f.buffers.end(), p));
  };
  release(cnt);
  if (f.srp) {
    release(f.srp32);
    f.srp32 = nullptr;
  }
  *chosen = true;
  return SBLAS_OK;
}

namespace {

template <typename VT, typename OT>
int panel_launch_typed(const PanelFormat &f, const OT *srp, int m, int n, int xcd_remap,
                       const double *x, double *y, double alpha, double beta, hipStream_t s) {
  const int grid = xcd_remap ? ((f.npanels + 7) / 8) * 8 : f.npanels;
#define SBLAS_PANEL_CASE(RPT)                                                                 \
  case RPT:                                                                                   \
    hipLaunchKernelGGL((k_panel_spmv<VT, RPT, OT>), dim3(grid), dim3(256), 0, s, m, n, f.rows, \
                       f.npanels, xcd_remap, f.q_lo, f.slice_base, f.slice_off, srp,          \
                       static_cast<const VT *>(f.pval), f.pcol, x, y, alpha, beta);          \
    break;
  switch (f.rpt) {
    SBLAS_PANEL_CASE(1)
    SBLAS_PANEL_CASE(2)
    SBLAS_PANEL_CASE(4)
    SBLAS_PANEL_CASE(6)
    SBLAS_PANEL_CASE(8)
    default:
      set_error(SBLAS_ERR_ARGS, "panel kernel: bad rows per thread");
      return SBLAS_ERR_ARGS;
  }
#undef SBLAS_PANEL_CASE
  SBLAS_HIP_TRY(hipGetLastError());
  return SBLAS_OK;
}

template <typename VT>
int panel_launch_vt(const PanelFormat &f, int m, int n, int xcd_remap, const double *x, double *y,
                    double alpha, double beta, hipStream_t s) {
  return f.srp ? panel_launch_typed<VT>(f, f.srp, m, n, xcd_remap, x, y, alpha, beta, s)
               : panel_launch_typed<VT>(f, f.srp32, m, n, xcd_remap, x, y, alpha, beta, s);
}

}  // namespace

int panel_rpt(int rows) {
  const int need = (rows + 255) / 256;
  return need <= 1 ? 1 : need <= 2 ? 2 : need <= 4 ? 4 : need <= 6 ? 6 : 8;
}

int panel_rows_for(int val_type, int m, int device) {
  int cus = 256, per_cu = 0;
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
  const void *fn = val_type == SBLAS_VAL_F64
                       ? reinterpret_cast<const void *>(&k_panel_spmv<double, 8, uint16_t>)
                   : val_type == SBLAS_VAL_F32
                       ? reinterpret_cast<const void *>(&k_panel_spmv<float, 8, uint16_t>)
                       : reinterpret_cast<const void *>(&k_panel_spmv<uint16_t, 8, uint16_t>);
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 256, 0) != hipSuccess ||
      per_cu < 1)
    per_cu = 2;
  const int64_t slots = static_cast<int64_t>(std::max(cus, 1)) * per_cu;
  // whole rounds of resident workgroups (at least two, so a slow workgroup is
  // not the whole tail), the tallest panels that fit
  for (int64_t k = 2;; k++) {
    int64_t R = (m + k * slots - 1) / (k * slots);
    R = (R + 63) / 64 * 64;
    if (R <= kPanelRowsMax) return static_cast<int>(std::max<int64_t>(R, 64));
  }
}

int launch_panel_spmv(int val_type, const PanelFormat &f, int m, int n, int xcd_remap,
                      const double *x, double *y, double alpha, double beta, hipStream_t s) {
  if (f.npanels <= 0) return SBLAS_OK;
  switch (val_type) {
    case SBLAS_VAL_F64: return panel_launch_vt<double>(f, m, n, xcd_remap, x, y, alpha, beta, s);
    case SBLAS_VAL_F32: return panel_launch_vt<float>(f, m, n, xcd_remap, x, y, alpha, beta, s);
    default: return panel_launch_vt<uint16_t>(f, m, n, xcd_remap, x, y, alpha, beta, s);
  }
}

}  // namespace sblas
