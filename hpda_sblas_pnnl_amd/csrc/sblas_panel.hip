// sblas_panel.hip -- column-panel form of kernel 1 for matrices whose rows
// draw x from a wide window (banded FEM-like matrices).
//
// When every element's x read is a separate L2 request (x window of a few
// hundred KB, far sparser than one element per cache line), the L2 request
// rate, not HBM, bounds a row-parallel kernel.  This form cuts each execution
// chunk into row panels of up to R rows and the columns into panels of W
// columns.  A workgroup owns one row panel and walks its column panels in
// increasing order: it loads x[q*W, (q+1)*W) straight into LDS with coalesced
// loads (one pass over the panel's x lines), then streams that slice's elements in
// chunks of 2048 (each thread four pairs of consecutive elements, so every
// load instruction of a wave reads whole lines; slices are padded to 8
// elements to keep pairs aligned), multiplies each by its x from LDS and
// leaves the products in LDS.  The plan stores each
// (row panel, column panel) slice contiguously, rows in order, with 16-bit
// column offsets inside the panel and row offsets inside the slice (16-bit
// when every slice holds at most 65535 elements, else 32-bit).
//
// Each row is then summed by one thread from the staged products, element by
// element in ascending column order, with separately rounded multiplies and
// adds: the same operations in the same order as the reference's host loop
// (test/dspmv_test_smt.cu:45-51), so a row that no chunk edge cuts equals the
// host oracle bit for bit.  A row cut by a chunk edge leaves one partial per
// chunk (slot 2p for a panel's first row, 2p+1 for its last), added in chunk
// order by the device merge.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <climits>
#include <type_traits>

#include "internal.hpp"
#include "panel_device.hpp"
#include "sorted_panel_device.hpp"

namespace sblas {
namespace {

__global__ void k_fill_u32(int64_t count, uint32_t *__restrict__ a, uint32_t v) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < count) a[i] = v;
}

}  // namespace

namespace {

template <typename VT, int RPT, typename OT, bool NX>
__global__ __launch_bounds__(256) void k_panel_spmv(
    int n, int R, int npanels, int xcd_remap, const PanelInfo *__restrict__ info,
    const int *__restrict__ slice_base, const int64_t *__restrict__ slice_off,
    const OT *__restrict__ srp, const VT *__restrict__ pval,
    const uint16_t *__restrict__ pcol, const double *__restrict__ x, double *__restrict__ y,
    double alpha, double beta, double *__restrict__ partials, const float *__restrict__ x32) {
  __shared__ double lds[panel::lds_doubles<OT, RPT, NX>()];
  int p = blockIdx.x;
  if (xcd_remap) {  // consecutive row panels on one XCD share their x window in its L2
    const int per = gridDim.x >> 3;
    p = (p & 7) * per + (p >> 3);
  }
  if (p >= npanels) return;  // uniform per workgroup
  const PanelInfo pi = info[p];
  panel::row_panel<VT, RPT, OT, NX>(p, n, R, pi, slice_base, slice_off, srp, pval, pcol, x, y,
                                    alpha, beta, partials + 2 * p, lds, x32);
}

// The narrow-x instances, held to four workgroups per CU (their 36 KB of LDS
// allow four; registers decide).
template <typename VT, int RPT, typename OT>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_panel_spmv_nx(
    int n, int R, int npanels, int xcd_remap, const PanelInfo *__restrict__ info,
    const int *__restrict__ slice_base, const int64_t *__restrict__ slice_off,
    const OT *__restrict__ srp, const VT *__restrict__ pval,
    const uint16_t *__restrict__ pcol, const double *__restrict__ x, double *__restrict__ y,
    double alpha, double beta, double *__restrict__ partials, const float *__restrict__ x32) {
  __shared__ double lds[panel::lds_doubles<OT, RPT, true>()];
  int p = blockIdx.x;
  if (xcd_remap) {
    const int per = gridDim.x >> 3;
    p = (p & 7) * per + (p >> 3);
  }
  if (p >= npanels) return;  // uniform per workgroup
  const PanelInfo pi = info[p];
  panel::row_panel<VT, RPT, OT, true>(p, n, R, pi, slice_base, slice_off, srp, pval, pcol, x, y,
                                      alpha, beta, partials + 2 * p, lds, x32);
}

// The sorted-group layout: one lane per row of a 64-row group.
// At most 128 registers a lane: four workgroups per CU.
template <typename VT, int RPT, bool NX>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4))) void k_panel_sorted(
    int n, int R, int npanels, int xcd_remap, const PanelInfo *__restrict__ info,
    const int *__restrict__ slice_base, const int64_t *__restrict__ slice_off,
    const uint16_t *__restrict__ perm, const uint32_t *__restrict__ goff,
    const VT *__restrict__ pval, const uint16_t *__restrict__ pcol, const double *__restrict__ x,
    const float *__restrict__ x32, double *__restrict__ y, double alpha, double beta,
    double *__restrict__ partials) {
  __shared__ double lds[sorted_panel::lds_doubles<RPT, NX>()];
  int p = blockIdx.x;
  if (xcd_remap) {
    const int per = gridDim.x >> 3;
    p = (p & 7) * per + (p >> 3);
  }
  if (p >= npanels) return;  // uniform per workgroup
  const PanelInfo pi = info[p];
  sorted_panel::panel<VT, RPT, NX>(p, n, R, pi, slice_base, slice_off, perm, goff, pval, pcol, x,
                                   x32, y, alpha, beta, partials + 2 * p, lds);
}

// BF16 elements of the sorted-group layout packed: value << 16 | column offset
__global__ void k_pack_bf16(int64_t count, const uint16_t *__restrict__ v,
                            const uint16_t *__restrict__ c, uint32_t *__restrict__ out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < count) out[i] = (static_cast<uint32_t>(v[i]) << 16) | c[i];
}

// x32[i] = x[i] rounded to FP32 (the narrow-x column-panel form's x copy)
__global__ void k_x_to_f32(int n, const double *__restrict__ x, float *__restrict__ x32) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x32[i] = static_cast<float>(x[i]);
}

}  // namespace

// ---- format build ---------------------------------------------------------------

namespace {

// Column-panel range of each row panel: first and last column panel any of
// its elements touches (columns are sorted within rows).
__global__ void k_panel_range(int npanels, PanelInfo *__restrict__ info,
                              const int2 *__restrict__ win, const int *__restrict__ rp,
                              const int *__restrict__ col, int *__restrict__ q_n) {
  const int p = blockIdx.x;
  if (p >= npanels) return;
  __shared__ int smin[256], smax[256];
  const PanelInfo pi = info[p];
  const int2 w = win[p];
  int lo = INT_MAX, hi = -1;
  for (int rl = threadIdx.x; rl < pi.nrow; rl += 256) {
    const int r = pi.row0 + rl;
    const int b = max(rp[r], w.x), e = min(rp[r + 1], w.y);
    if (e > b) {
      lo = min(lo, col[b]);
      hi = max(hi, col[e - 1]);
    }
  }
  smin[threadIdx.x] = lo;
  smax[threadIdx.x] = hi;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) {
      smin[threadIdx.x] = min(smin[threadIdx.x], smin[threadIdx.x + k]);
      smax[threadIdx.x] = max(smax[threadIdx.x], smax[threadIdx.x + k]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    if (smax[0] < 0) {  // no element in the panel
      info[p].q_lo = 0;
      q_n[p] = 0;
    } else {
      info[p].q_lo = smin[0] / kPanelCols;
      q_n[p] = smax[0] / kPanelCols - smin[0] / kPanelCols + 1;
    }
  }
}

// Element-parallel passes over one row panel's element window: one
// workgroup per row panel, 1024 elements per step (four consecutive elements
// per thread).  An element's row comes from a binary search of the panel's
// row pointers (clamped to the window) in LDS.  A run is a row's elements in
// one column panel -- contiguous, since columns ascend within a row -- and
// starts where the row starts or the column panel changes; an element's place
// in its run is its distance from the run's start, found by a workgroup
// max-scan of run starts carried from step to step.
//   count (SCATTER false): at a run's last element,
//       cnt[slice * R + row] = run length
//   scatter (SCATTER true): the element goes to
//       slice_off[slice] + srp[slice * (R + 1) + row] + place
//   or, in the sorted-group layout (SORTED), with i = rank[slice * R + row]
//   (the row's place in the slice's order) and G = R / 64 groups,
//       slice_off[slice] + goff[slice * (G + 1) + i / 64] + 64 * place + i % 64
//   with its value and its column's offset inside the column panel.
template <typename VT, bool SCATTER, bool SORTED = false>
__global__ __launch_bounds__(256) void k_panel_runs(
    int npanels, int R, const PanelInfo *__restrict__ info, const int2 *__restrict__ win,
    const int *__restrict__ rp, const int *__restrict__ col, const VT *__restrict__ val,
    const int *__restrict__ slice_base, const int64_t *__restrict__ slice_off,
    const uint32_t *__restrict__ srp, uint32_t *__restrict__ cnt, VT *__restrict__ pval,
    uint16_t *__restrict__ pcol, const uint16_t *__restrict__ rank = nullptr,
    const uint32_t *__restrict__ goff = nullptr) {
  constexpr int IT = 4;
  typedef hipcub::BlockScan<int, 256> Scan;
  __shared__ typename Scan::TempStorage tmp;
  __shared__ int lrp[kPanelRowsMax + 1];
  const int p = blockIdx.x;
  if (p >= npanels) return;  // uniform per workgroup
  const PanelInfo pi = info[p];
  const int2 w = win[p];
  const int sb = slice_base[p];
  const int t = threadIdx.x;
  for (int i = t; i <= pi.nrow; i += 256) lrp[i] = min(max(rp[pi.row0 + i], w.x), w.y);
  __syncthreads();
  int carry = w.x;  // latest run start before this step (the same in every thread)
  for (int base = w.x; base < w.y; base += 256 * IT) {
    int rl[IT], q[IT], c[IT], start[IT];
#pragma unroll
    for (int k = 0; k < IT; k++) {
      const int j = base + t * IT + k;
      rl[k] = 0;
      q[k] = 0;
      c[k] = 0;
      start[k] = -1;
      if (j < w.y) {
        int lo = 0, hi = pi.nrow - 1;  // the last row whose (clamped) start is <= j
        while (lo < hi) {
          const int mid = (lo + hi + 1) >> 1;
          if (lrp[mid] <= j) lo = mid; else hi = mid - 1;
        }
        rl[k] = lo;
        c[k] = col[j];
        q[k] = c[k] / kPanelCols;
        const bool first = j == lrp[lo] || col[j - 1] / kPanelCols != q[k];
        start[k] = first ? j : -1;
      }
    }
    int agg = -1;
    Scan(tmp).InclusiveScan(start, start, hipcub::Max(), agg);
#pragma unroll
    for (int k = 0; k < IT; k++) {
      const int j = base + t * IT + k;
      if (j < w.y) {
        const int s = sb + q[k] - pi.q_lo;
        const int place = j - max(start[k], carry);
        if constexpr (SCATTER) {
          int64_t pos;
          if constexpr (SORTED) {
            const int i = rank[static_cast<int64_t>(s) * R + rl[k]];
            pos = slice_off[s] + goff[static_cast<int64_t>(s) * ((R >> 6) + 1) + (i >> 6)] +
                  64 * place + (i & 63);
          } else {
            pos = slice_off[s] + srp[static_cast<int64_t>(s) * (R + 1) + rl[k]] + place;
          }
          pval[pos] = val[j];
          pcol[pos] = static_cast<uint16_t>(c[k] - q[k] * kPanelCols);
        } else {
          const bool last = j + 1 == lrp[rl[k] + 1] || col[j + 1] / kPanelCols != q[k];
          if (last) cnt[static_cast<int64_t>(s) * R + rl[k]] = static_cast<uint32_t>(place + 1);
        }
      }
    }
    carry = max(carry, agg);
    __syncthreads();  // the scan's storage is reused by the next step
  }
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
    snnz[s] = (total + kPanelPad - 1) / kPanelPad * kPanelPad;  // slices start aligned
    if (total > 65535u) overflow[0] = 1;
  }
}

// Sorted-group layout: each slice's R rows ordered by their count, longest
// first (ties in row order), one workgroup per slice; thread t holds places
// t*RPT .. t*RPT+RPT-1 of the order.  perm[slice * R + i] = the row at place
// i, rank[slice * R + row] = its place, goff[slice * (G + 1) + g] = 64 times
// the summed lengths of groups before g (a group's length is its first
// row's count), snnz[slice] = the slice's stored elements.
template <int RPT>
__global__ __launch_bounds__(256) void k_sorted_slice_order(int nslices, int R,
                                                            const uint32_t *__restrict__ cnt,
                                                            uint16_t *__restrict__ perm,
                                                            uint16_t *__restrict__ rank,
                                                            uint32_t *__restrict__ goff,
                                                            int64_t *__restrict__ snnz) {
  typedef hipcub::BlockRadixSort<uint32_t, 256, RPT, uint16_t> Sort;
  __shared__ typename Sort::TempStorage tmp;
  __shared__ uint32_t glen[kPanelRowsMax / 64];
  const int s = blockIdx.x;
  if (s >= nslices) return;  // uniform per workgroup
  const int t = threadIdx.x, G = R >> 6;
  uint32_t key[RPT];
  uint16_t row[RPT];
#pragma unroll
  for (int j = 0; j < RPT; j++) {
    const int i = t * RPT + j;
    // places past R sort after every row (count 0, stable order)
    key[j] = i < R ? cnt[static_cast<int64_t>(s) * R + i] : 0u;
    row[j] = static_cast<uint16_t>(i);
  }
  // counts are at most kPanelCols (13 bits)
  Sort(tmp).SortDescending(key, row, 0, 13);
#pragma unroll
  for (int j = 0; j < RPT; j++) {
    const int i = t * RPT + j;
    if (i < R) {
      perm[static_cast<int64_t>(s) * R + i] = row[j];
      rank[static_cast<int64_t>(s) * R + row[j]] = static_cast<uint16_t>(i);
      if ((i & 63) == 0) glen[i >> 6] = key[j];
    }
  }
  __syncthreads();
  if (t == 0) {
    uint32_t off = 0;
    for (int g = 0; g < G; g++) {
      goff[static_cast<int64_t>(s) * (G + 1) + g] = off;
      off += 64u * glen[g];
    }
    goff[static_cast<int64_t>(s) * (G + 1) + G] = off;
    snnz[s] = off;  // a multiple of 64: slices start aligned
  }
}

__global__ void k_narrow_u16(int64_t count, const uint32_t *__restrict__ a,
                             uint16_t *__restrict__ b) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < count) b[i] = static_cast<uint16_t>(a[i]);
}

}  // namespace

int panel_build(int val_type, int rows, int m, int n, int64_t nnz, const int *rp, const int *col,
                const void *val, const std::vector<ExecChunk> &ex, const int *h_rp,
                bool force, int min_panels, PanelFormat &f, bool *chosen, hipStream_t s,
                int layout) {
  (void)n;
  (void)m;
  *chosen = false;
  if (rows < 1 || rows > kPanelRowsMax) {
    set_error(SBLAS_ERR_ARGS, "panel format: rows per panel must be in [1, 2048]");
    return SBLAS_ERR_ARGS;
  }
  const bool sorted = layout == kPanelSorted;
  const int R = sorted ? (rows + 63) / 64 * 64 : rows;  // whole groups of 64 rows
  f.layout = sorted ? kPanelSorted : kPanelRowOrder;
  f.groups = R / 64;
  f.rows = R;
  f.rpt = panel_rpt(R);
  f.buffers.clear();
  // row panels: each execution chunk's rows in runs of R, its element window
  // clamped to the chunk; flags mark a first / last row that the window cuts
  f.hpanels.clear();
  std::vector<PanelInfo> hinfo;
  std::vector<int2> hwin;
  for (const ExecChunk &ch : ex) {
    for (int r = ch.row_begin; r < ch.row_end; r += R) {
      const int re = std::min(r + R, ch.row_end);
      const int lo = static_cast<int>(std::max<int64_t>(h_rp[r], ch.lo));
      const int hi = static_cast<int>(std::max<int64_t>(std::min<int64_t>(h_rp[re], ch.hi), lo));
      auto split = [&](int row) { return h_rp[row] < lo || h_rp[row + 1] > hi; };
      int flags = split(r) ? 1 : 0;
      if (re - 1 != r && split(re - 1)) flags |= 2;
      hinfo.push_back(PanelInfo{r, re - r, 0, flags});
      hwin.push_back(make_int2(lo, hi));
      f.hpanels.push_back(ExecChunk{lo, hi, r, re});
    }
  }
  f.npanels = static_cast<int>(hinfo.size());
  auto alloc = [&](size_t bytes) -> void * {
    void *p = nullptr;
    if (hipMalloc(&p, bytes ? bytes : 8) != hipSuccess) return nullptr;
    f.buffers.push_back(p);
    return p;
  };
  auto release = [&](void *p) {
    (void)hipFree(p);
    f.buffers.erase(std::find(f.buffers.begin(), f.buffers.end(), p));
  };
  const int np = std::max(f.npanels, 1);
  f.info = static_cast<PanelInfo *>(alloc(sizeof(PanelInfo) * np));
  int2 *win = static_cast<int2 *>(alloc(sizeof(int2) * np));
  int *q_n = static_cast<int *>(alloc(sizeof(int) * np));
  f.slice_base = static_cast<int *>(alloc(sizeof(int) * (np + 1)));
  if (!f.info || !win || !q_n || !f.slice_base) {
    set_error(SBLAS_ERR_ALLOC, "panel format: allocation failed");
    return SBLAS_ERR_ALLOC;
  }
  if (f.npanels > 0) {
    SBLAS_HIP_TRY(hipMemcpyAsync(f.info, hinfo.data(), sizeof(PanelInfo) * f.npanels,
                                 hipMemcpyHostToDevice, s));
    SBLAS_HIP_TRY(hipMemcpyAsync(win, hwin.data(), sizeof(int2) * f.npanels,
                                 hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_panel_range, dim3(f.npanels), dim3(256), 0, s, f.npanels, f.info, win, rp,
                       col, q_n);
  }
  std::vector<int> hq(np, 0), hb(np + 1, 0);
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
  // most a quarter of an x line per element, and a row panel's x window
  // spans more than three column panels (avg_panels * 256 x lines for a
  // row-parallel block of 2048 elements).  Measured on banded matrices: row
  // panels spanning 3.7, 5.2 and 7.7 column panels run 1.2x, 1.4x and 2.1x
  // faster than row blocks, 2.1 column panels 0.7x.
  // One workgroup sums a row panel, so a panel holding far more than its
  // share keeps its workgroup running long after the others (the first row
  // panels of an R-MAT matrix hold most of its elements; in the 64-matrix
  // batch ten such matrices held the persistent launch for 1.96 ms): the
  // densest panel may hold at most 4x the mean, or 65536 elements.
  int64_t densest = 0;
  for (const int2 &w : hwin) densest = std::max<int64_t>(densest, w.y - w.x);
  const bool balanced =
      densest <= std::max<int64_t>(65536, 4 * nnz / std::max(f.npanels, 1));
  const bool helps = f.npanels >= min_panels && f.avg_slice >= 1024.0 &&
                     f.avg_panels_per_row_panel > 3.0 && balanced;
  if (!force && !helps) {
    for (void *p : f.buffers) (void)hipFree(p);
    f.buffers.clear();
    f.npanels = 0;
    f.hpanels.clear();
    return SBLAS_OK;
  }
  SBLAS_HIP_TRY(hipMemcpyAsync(f.slice_base, hb.data(), sizeof(int) * (np + 1),
                               hipMemcpyHostToDevice, s));
  const int64_t ns = std::max<int64_t>(f.nslices, 1);
  uint32_t *cnt = static_cast<uint32_t *>(alloc(sizeof(uint32_t) * ns * R));
  f.srp = nullptr;
  f.srp32 = nullptr;
  uint16_t *rank = nullptr;
  if (sorted) {
    f.perm = static_cast<uint16_t *>(alloc(sizeof(uint16_t) * ns * R));
    rank = static_cast<uint16_t *>(alloc(sizeof(uint16_t) * ns * R));
    f.goff = static_cast<uint32_t *>(alloc(sizeof(uint32_t) * ns * (f.groups + 1)));
  } else {
    f.srp32 = static_cast<uint32_t *>(alloc(sizeof(uint32_t) * ns * (R + 1)));
  }
  int64_t *snnz = static_cast<int64_t *>(alloc(sizeof(int64_t) * (ns + 1)));
  f.slice_off = static_cast<int64_t *>(alloc(sizeof(int64_t) * (ns + 1)));
  int *overflow = static_cast<int *>(alloc(sizeof(int)));
  f.partials = static_cast<double *>(alloc(sizeof(double) * 2 * np));
  if (!cnt || (sorted ? (!f.perm || !rank || !f.goff) : !f.srp32) || !snnz || !f.slice_off ||
      !overflow || !f.partials) {
    set_error(SBLAS_ERR_ALLOC, "panel format: allocation failed");
    return SBLAS_ERR_ALLOC;
  }
  // counts start at zero: written by a kernel (no memset nodes)
  const int64_t cnt_n = ns * R;
  hipLaunchKernelGGL(k_fill_u32, dim3(static_cast<unsigned>((cnt_n + 255) / 256)), dim3(256), 0, s,
                     cnt_n, cnt, 0u);
  hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(256), 0, s, int64_t{1},
                     reinterpret_cast<uint32_t *>(overflow), 0u);
  if (f.npanels > 0)
    hipLaunchKernelGGL((k_panel_runs<double, false>), dim3(f.npanels), dim3(256), 0, s,
                       f.npanels, R, f.info, win, rp, col, nullptr, f.slice_base, nullptr,
                       nullptr, cnt, nullptr, nullptr);
  if (f.nslices > 0 && sorted) {
#define SBLAS_ORDER_CASE(RPT)                                                                 \
  case RPT:                                                                                   \
    hipLaunchKernelGGL(k_sorted_slice_order<RPT>, dim3(f.nslices), dim3(256), 0, s, f.nslices, \
                       R, cnt, f.perm, rank, f.goff, snnz);                                   \
    break;
    switch (f.rpt) {
      SBLAS_ORDER_CASE(1)
      SBLAS_ORDER_CASE(2)
      SBLAS_ORDER_CASE(4)
      SBLAS_ORDER_CASE(6)
      SBLAS_ORDER_CASE(8)
    }
#undef SBLAS_ORDER_CASE
  } else if (f.nslices > 0) {
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
  if (!h_over && !sorted) {
    // every slice fits 16-bit offsets: keep the narrow copy only
    // two entries of slack: the kernel stages a slice's offsets as whole
    // 32-bit words, and the last slice's last word may end past them
    const int64_t cnt_o = ns * (R + 1);
    f.srp = static_cast<uint16_t *>(alloc(sizeof(uint16_t) * (cnt_o + 2)));
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
                       reinterpret_cast<uint32_t *>(snnz + f.nslices), 0u);
    SBLAS_HIP_TRY(hipcub::DeviceScan::ExclusiveSum(tmp, tb, snnz, f.slice_off, f.nslices + 1, s));
  }
  // element arrays: slices padded to kPanelPad elements, padding zeroed
  int64_t padded = 0;
  if (f.nslices > 0) {
    SBLAS_HIP_TRY(hipMemcpyAsync(&padded, f.slice_off + f.nslices, sizeof(int64_t),
                                 hipMemcpyDeviceToHost, s));
    SBLAS_HIP_TRY(hipStreamSynchronize(s));
  } else {
    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(256), 0, s, int64_t{2},
                       reinterpret_cast<uint32_t *>(f.slice_off), 0u);
  }
  f.padded = padded;
  const size_t vbytes = val_type == SBLAS_VAL_F64 ? 8 : (val_type == SBLAS_VAL_F32 ? 4 : 2);
  const int64_t cap = std::max<int64_t>(padded, kPanelPad);
  f.pval = alloc(vbytes * cap);
  f.pcol = static_cast<uint16_t *>(alloc(sizeof(uint16_t) * cap));
  if (!f.pval || !f.pcol) {
    set_error(SBLAS_ERR_ALLOC, "panel format: allocation failed");
    return SBLAS_ERR_ALLOC;
  }
  {
    // padding: value 0, column 0 (row order) or kSortedPad (sorted groups)
    const int64_t wv = static_cast<int64_t>(vbytes) * cap / 4, wc = cap * 2 / 4;
    const uint32_t cpad = sorted ? (uint32_t{kSortedPad} << 16) | kSortedPad : 0u;
    hipLaunchKernelGGL(k_fill_u32, dim3(static_cast<unsigned>((wv + 255) / 256)), dim3(256), 0, s,
                       wv, static_cast<uint32_t *>(f.pval), 0u);
    hipLaunchKernelGGL(k_fill_u32, dim3(static_cast<unsigned>((wc + 255) / 256)), dim3(256), 0, s,
                       wc, reinterpret_cast<uint32_t *>(f.pcol), cpad);
  }
  if (f.npanels > 0 && sorted) {
    const dim3 g(f.npanels), b(256);
    switch (val_type) {
      case SBLAS_VAL_F64:
        hipLaunchKernelGGL((k_panel_runs<double, true, true>), g, b, 0, s, f.npanels, R, f.info,
                           win, rp, col, static_cast<const double *>(val), f.slice_base,
                           f.slice_off, nullptr, nullptr, static_cast<double *>(f.pval), f.pcol,
                           rank, f.goff);
        break;
      case SBLAS_VAL_F32:
        hipLaunchKernelGGL((k_panel_runs<float, true, true>), g, b, 0, s, f.npanels, R, f.info,
                           win, rp, col, static_cast<const float *>(val), f.slice_base,
                           f.slice_off, nullptr, nullptr, static_cast<float *>(f.pval), f.pcol,
                           rank, f.goff);
        break;
      default:
        hipLaunchKernelGGL((k_panel_runs<uint16_t, true, true>), g, b, 0, s, f.npanels, R, f.info,
                           win, rp, col, static_cast<const uint16_t *>(val), f.slice_base,
                           f.slice_off, nullptr, nullptr, static_cast<uint16_t *>(f.pval), f.pcol,
                           rank, f.goff);
        break;
    }
  } else if (f.npanels > 0) {
    const dim3 g(f.npanels), b(256);
    switch (val_type) {
      case SBLAS_VAL_F64:
        hipLaunchKernelGGL((k_panel_runs<double, true>), g, b, 0, s, f.npanels, R, f.info, win, rp,
                           col, static_cast<const double *>(val), f.slice_base, f.slice_off,
                           f.srp32, nullptr, static_cast<double *>(f.pval), f.pcol);
        break;
      case SBLAS_VAL_F32:
        hipLaunchKernelGGL((k_panel_runs<float, true>), g, b, 0, s, f.npanels, R, f.info, win, rp,
                           col, static_cast<const float *>(val), f.slice_base, f.slice_off,
                           f.srp32, nullptr, static_cast<float *>(f.pval), f.pcol);
        break;
      default:
        hipLaunchKernelGGL((k_panel_runs<uint16_t, true>), g, b, 0, s, f.npanels, R, f.info, win,
                           rp, col, static_cast<const uint16_t *>(val), f.slice_base, f.slice_off,
                           f.srp32, nullptr, static_cast<uint16_t *>(f.pval), f.pcol);
        break;
    }
  }
  SBLAS_HIP_TRY(hipGetLastError());
  SBLAS_HIP_TRY(hipStreamSynchronize(s));
  // scratch of the build is no longer needed
  release(cnt);
  release(win);
  release(q_n);
  if (rank) release(rank);
  if (f.srp) {
    release(f.srp32);
    f.srp32 = nullptr;
  }
  if (sorted && val_type == SBLAS_VAL_BF16) {
    // each element's value and column offset into one word
    f.packed = static_cast<uint32_t *>(alloc(sizeof(uint32_t) * std::max<int64_t>(f.padded, 1)));
    if (!f.packed) {
      set_error(SBLAS_ERR_ALLOC, "panel: packed BF16 elements");
      return SBLAS_ERR_ALLOC;
    }
    hipLaunchKernelGGL(k_pack_bf16, dim3(static_cast<unsigned>((f.padded + 255) / 256)), dim3(256),
                       0, s, f.padded, static_cast<const uint16_t *>(f.pval), f.pcol, f.packed);
    SBLAS_HIP_TRY(hipGetLastError());
    SBLAS_HIP_TRY(hipStreamSynchronize(s));
    release(f.pval);
    release(f.pcol);
    f.pval = nullptr;
    f.pcol = nullptr;
  }
  *chosen = true;
  return SBLAS_OK;
}

namespace {

template <typename VT, typename OT, bool NX>
int panel_launch_typed(const PanelFormat &f, const OT *srp, int m, int n, int xcd_remap,
                       const double *x, double *y, double alpha, double beta, hipStream_t s) {
  const int grid = xcd_remap ? ((f.npanels + 7) / 8) * 8 : f.npanels;
#define SBLAS_PANEL_CASE(RPT)                                                                 \
  case RPT:                                                                                   \
    if constexpr (NX)                                                                         \
      hipLaunchKernelGGL((k_panel_spmv_nx<VT, RPT, OT>), dim3(grid), dim3(256), 0, s, n,       \
                         f.rows, f.npanels, xcd_remap, f.info, f.slice_base, f.slice_off,     \
                         srp, static_cast<const VT *>(f.pval), f.pcol, x, y, alpha, beta,     \
                         f.partials, f.x32);                                                  \
    else                                                                                      \
      hipLaunchKernelGGL((k_panel_spmv<VT, RPT, OT, false>), dim3(grid), dim3(256), 0, s, n,  \
                         f.rows, f.npanels, xcd_remap, f.info, f.slice_base, f.slice_off,     \
                         srp, static_cast<const VT *>(f.pval), f.pcol, x, y, alpha, beta,     \
                         f.partials, nullptr);                                                \
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

template <typename VT, bool NX>
int panel_launch_nx(const PanelFormat &f, int m, int n, int xcd_remap, const double *x, double *y,
                    double alpha, double beta, hipStream_t s) {
  return f.srp ? panel_launch_typed<VT, uint16_t, NX>(f, f.srp, m, n, xcd_remap, x, y, alpha,
                                                      beta, s)
               : panel_launch_typed<VT, uint32_t, NX>(f, f.srp32, m, n, xcd_remap, x, y, alpha,
                                                      beta, s);
}

// VT = uint32_t: packed BF16 elements (f.packed)
template <typename VT, bool NX>
int sorted_launch(const PanelFormat &f, int n, int xcd_remap, const double *x, double *y,
                  double alpha, double beta, hipStream_t s) {
  const VT *vals = sorted_panel::kPackedBf16<VT> ? reinterpret_cast<const VT *>(f.packed)
                                                 : static_cast<const VT *>(f.pval);
  const int grid = xcd_remap ? ((f.npanels + 7) / 8) * 8 : f.npanels;
#define SBLAS_SORTED_CASE(RPT)                                                                \
  case RPT:                                                                                   \
    hipLaunchKernelGGL((k_panel_sorted<VT, RPT, NX>), dim3(grid), dim3(256), 0, s, n, f.rows,  \
                       f.npanels, xcd_remap, f.info, f.slice_base, f.slice_off, f.perm,       \
                       f.goff, vals, f.pcol, x, f.x32, y, alpha, beta, f.partials);           \
    break;
  switch (f.rpt) {
    SBLAS_SORTED_CASE(1)
    SBLAS_SORTED_CASE(2)
    SBLAS_SORTED_CASE(4)
    SBLAS_SORTED_CASE(6)
    SBLAS_SORTED_CASE(8)
    default:
      set_error(SBLAS_ERR_ARGS, "panel kernel: bad rows per thread");
      return SBLAS_ERR_ARGS;
  }
#undef SBLAS_SORTED_CASE
  SBLAS_HIP_TRY(hipGetLastError());
  return SBLAS_OK;
}

template <typename VT>
int panel_launch_vt(const PanelFormat &f, int m, int n, int xcd_remap, const double *x, double *y,
                    double alpha, double beta, hipStream_t s, bool x32_ready) {
  if constexpr (!std::is_same<VT, double>::value) {
    if (f.x32) {
      // the FP32 copy of x first, on the same stream
      if (!x32_ready) {
        hipLaunchKernelGGL(k_x_to_f32, dim3((n + 255) / 256), dim3(256), 0, s, n, x, f.x32);
        SBLAS_HIP_TRY(hipGetLastError());
      }
      if (f.layout == kPanelSorted && f.packed)
        return sorted_launch<uint32_t, true>(f, n, xcd_remap, x, y, alpha, beta, s);
      if (f.layout == kPanelSorted)
        return sorted_launch<VT, true>(f, n, xcd_remap, x, y, alpha, beta, s);
      return panel_launch_nx<VT, true>(f, m, n, xcd_remap, x, y, alpha, beta, s);
    }
  }
  if (f.layout == kPanelSorted && f.packed)
    return sorted_launch<uint32_t, false>(f, n, xcd_remap, x, y, alpha, beta, s);
  if (f.layout == kPanelSorted)
    return sorted_launch<VT, false>(f, n, xcd_remap, x, y, alpha, beta, s);
  return panel_launch_nx<VT, false>(f, m, n, xcd_remap, x, y, alpha, beta, s);
}

}  // namespace

int launch_x_to_f32(int n, const double *x, float *x32, hipStream_t s) {
  if (n <= 0) return SBLAS_OK;
  hipLaunchKernelGGL(k_x_to_f32, dim3((n + 255) / 256), dim3(256), 0, s, n, x, x32);
  SBLAS_HIP_TRY(hipGetLastError());
  return SBLAS_OK;
}

int panel_rpt(int rows) {
  const int need = (rows + 255) / 256;
  return need <= 1 ? 1 : need <= 2 ? 2 : need <= 4 ? 4 : need <= 6 ? 6 : 8;
}

namespace {
// resident workgroups per CU of the sorted-group kernel with RPT rows per thread
template <int RPT>
int sorted_occupancy_rpt(int val_type, bool narrow_x) {
  const void *fn =
      val_type == SBLAS_VAL_F64
          ? reinterpret_cast<const void *>(&k_panel_sorted<double, RPT, false>)
      : val_type == SBLAS_VAL_F32
          ? (narrow_x ? reinterpret_cast<const void *>(&k_panel_sorted<float, RPT, true>)
                      : reinterpret_cast<const void *>(&k_panel_sorted<float, RPT, false>))
          : (narrow_x ? reinterpret_cast<const void *>(&k_panel_sorted<uint32_t, RPT, true>)
                      : reinterpret_cast<const void *>(&k_panel_sorted<uint32_t, RPT, false>));
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 256, 0) != hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  return per_cu;
}
int sorted_occupancy(int val_type, bool narrow_x, int rpt) {
  switch (rpt) {
    case 1: return sorted_occupancy_rpt<1>(val_type, narrow_x);
    case 2: return sorted_occupancy_rpt<2>(val_type, narrow_x);
    case 4: return sorted_occupancy_rpt<4>(val_type, narrow_x);
    case 6: return sorted_occupancy_rpt<6>(val_type, narrow_x);
    default: return sorted_occupancy_rpt<8>(val_type, narrow_x);
  }
}
}  // namespace

int panel_rows_for(int val_type, int m, int device, bool narrow_x, int layout, int chunks) {
  int cus = 256, per_cu = 0;
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
  const void *fn =
      layout == kPanelSorted
          ? (val_type == SBLAS_VAL_F64
                 ? reinterpret_cast<const void *>(&k_panel_sorted<double, 8, false>)
             : val_type == SBLAS_VAL_F32
                 ? (narrow_x ? reinterpret_cast<const void *>(&k_panel_sorted<float, 8, true>)
                             : reinterpret_cast<const void *>(&k_panel_sorted<float, 8, false>))
                 : (narrow_x
                        ? reinterpret_cast<const void *>(&k_panel_sorted<uint32_t, 8, true>)
                        : reinterpret_cast<const void *>(&k_panel_sorted<uint32_t, 8, false>)))
      : val_type == SBLAS_VAL_F64
          ? reinterpret_cast<const void *>(&k_panel_spmv<double, 8, uint16_t, false>)
      : val_type == SBLAS_VAL_F32
          ? (narrow_x ? reinterpret_cast<const void *>(&k_panel_spmv_nx<float, 8, uint16_t>)
                      : reinterpret_cast<const void *>(&k_panel_spmv<float, 8, uint16_t, false>))
          : (narrow_x
                 ? reinterpret_cast<const void *>(&k_panel_spmv_nx<uint16_t, 8, uint16_t>)
                 : reinterpret_cast<const void *>(&k_panel_spmv<uint16_t, 8, uint16_t, false>));
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, 256, 0) != hipSuccess ||
      per_cu < 1)
    per_cu = 2;
  const int64_t slots = static_cast<int64_t>(std::max(cus, 1)) * per_cu;
  if (layout == kPanelSorted) {
    // one round of resident workgroups where the panels fit: a workgroup's
    // slices follow one another, each waiting for its loads, so every
    // panel in flight at once beats shorter panels in two rounds (FP32
    // banded n = 2e6, 8 chunks: 1984 rows, one round, 0.344 ms; 1024 rows,
    // 1.9 rounds, 0.414 ms).  The fewest rows per thread (the most
    // workgroups per CU) whose kernel holds every panel at once; each
    // chunk's last panel may be short, so one panel per chunk more.
    for (int64_t k = 1;; k++) {
      for (const int rpt : {1, 2, 4, 6, 8}) {
        const int occ = sorted_occupancy(val_type, narrow_x, rpt);
        const int64_t room = k * static_cast<int64_t>(std::max(cus, 1)) * occ -
                             std::max(chunks, 1);
        if (room < 1) continue;
        int64_t R = (m + room - 1) / room;
        R = (R + 63) / 64 * 64;
        if (R <= 256 * rpt) return static_cast<int>(std::max<int64_t>(R, 64));
      }
    }
  }
  // whole rounds of resident workgroups (at least two, so a slow workgroup is
  // not the whole tail), the tallest panels that fit
  for (int64_t k = 2;; k++) {
    int64_t R = (m + k * slots - 1) / (k * slots);
    R = (R + 63) / 64 * 64;
    if (R <= kPanelRowsMax) return static_cast<int>(std::max<int64_t>(R, 64));
  }
}

int launch_panel_spmv(int val_type, const PanelFormat &f, int m, int n, int xcd_remap,
                      const double *x, double *y, double alpha, double beta, hipStream_t s,
                      bool x32_ready) {
  if (f.npanels <= 0) return SBLAS_OK;
  switch (val_type) {
    case SBLAS_VAL_F64:
      return panel_launch_vt<double>(f, m, n, xcd_remap, x, y, alpha, beta, s, x32_ready);
    case SBLAS_VAL_F32:
      return panel_launch_vt<float>(f, m, n, xcd_remap, x, y, alpha, beta, s, x32_ready);
    default:
      return panel_launch_vt<uint16_t>(f, m, n, xcd_remap, x, y, alpha, beta, s, x32_ready);
  }
}

}  // namespace sblas
