// sblas_tiles.hip -- tile kernels whose work per launch block is fixed in
// advance whatever the row lengths: merge-path CSR (the reference's kernel
// 2, cusparseDcsrmv_mp at src/dspmv_mgpu_v1.cu:267 / src/dspmv_mgpu_v2.cu:457)
// and the COO segmented reduction that consumes row-sorted COO directly (the
// reference converts COO to CSR first, src/dspmv_mgpu_baseline_coo.cu:146-165,
// src/dspmv_mgpu_v1_numa_coo.cu:460-476).
//
// Both write y directly for rows that lie inside one tile and leave, per tile,
// the partial of a row that began in an earlier tile (head) and of a row that
// continues into the next tile (tail).  k_tile_fixup then adds those partials
// in tile order, so every row's sum has a fixed order: no atomics on y.
#include <hipcub/hipcub.hpp>

#include <climits>
#include <mutex>
#include <string>

#include "internal.hpp"

namespace sblas {
namespace {

__device__ __forceinline__ double tload_val(const double *p, int j) {
  return __builtin_nontemporal_load(p + j);
}
__device__ __forceinline__ double tload_val(const float *p, int j) {
  return static_cast<double>(__builtin_nontemporal_load(p + j));
}
__device__ __forceinline__ double tload_val(const uint16_t *p, int j) {
  const uint32_t h = __builtin_nontemporal_load(p + j);
  return static_cast<double>(__uint_as_float(h << 16));
}

template <int W>
__device__ __forceinline__ double lanes_sum(double v) {
#pragma unroll
  for (int off = W / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, W);
  return v;
}

__device__ __forceinline__ double axpby(double alpha, double s, double beta, double yv) {
  return (beta == 0.0) ? alpha * s : alpha * s + beta * yv;
}

// Products of elements [j0, j1) into prod[0, j1-j0), consecutive threads on
// consecutive elements, ITEMS loads of values and columns issued before the
// x gathers.  Returns nothing; caller synchronises.
template <typename VT, int ITEMS, int BLOCK>
__device__ __forceinline__ void stage_products(int j0, int j1, const int *__restrict__ col,
                                               const VT *__restrict__ val,
                                               const double *__restrict__ x, double *prod) {
  const int t = threadIdx.x;
  if (j1 <= j0) return;
  // lanes past the tile re-read its last element and store nothing, so no
  // load sits under a branch and all loads of a thread are in flight at once
  double v[ITEMS];
  int c[ITEMS];
#pragma unroll
  for (int k = 0; k < ITEMS; k++) {
    const int e = min(j0 + t + k * BLOCK, j1 - 1);
    v[k] = tload_val(val, e);
    c[k] = __builtin_nontemporal_load(col + e);
  }
  double xv[ITEMS];
#pragma unroll
  for (int k = 0; k < ITEMS; k++) xv[k] = x[c[k]];
#pragma unroll
  for (int k = 0; k < ITEMS; k++) {
    const int e = t + k * BLOCK;
    if (j0 + e < j1) prod[e] = v[k] * xv[k];
  }
}

// ---------------------------------------------------------------------------
// merge path (Merrill & Garland, SC'16): the m row ends and nnz elements are
// merged into one list of m + nnz items; tile b owns items [b*T, (b+1)*T).
// A row belongs to the tile holding its row-end item.
// ---------------------------------------------------------------------------

// Number of row ends among the first d merge items.
__device__ __forceinline__ int merge_search(int64_t d, int m, int nnz, const int *rp) {
  int lo = static_cast<int>(max<int64_t>(0, d - nnz));
  int hi = static_cast<int>(min<int64_t>(d, m));
  while (lo < hi) {
    const int piv = (lo + hi) >> 1;
    if (static_cast<int64_t>(rp[piv + 1]) <= d - piv - 1) lo = piv + 1; else hi = piv;
  }
  return lo;
}

template <typename VT, int ITEMS, int SUBW, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_csr_merge_path(
    int m, int nnz, int tiles, const int *__restrict__ rp, const int *__restrict__ col,
    const VT *__restrict__ val, const double *__restrict__ x, double *__restrict__ y,
    double alpha, double beta, TileCarry *__restrict__ carry) {
  constexpr int T = BLOCK * ITEMS;
  __shared__ double prod[T];
  __shared__ double ysh[T + 1];
  __shared__ int coord[4];
  __shared__ double wsum[BLOCK / 64];
  const int b = blockIdx.x;
  if (b >= tiles) return;
  const int t = threadIdx.x;
  const int64_t total = static_cast<int64_t>(m) + nnz;
  if (t < 2) {
    const int64_t d = min<int64_t>(static_cast<int64_t>(b + t) * T, total);
    const int i = merge_search(d, m, nnz, rp);
    coord[2 * t] = i;
    coord[2 * t + 1] = static_cast<int>(d - i);
  }
  __syncthreads();
  const int i0 = coord[0], j0 = coord[1], i1 = coord[2], j1 = coord[3];
  stage_products<VT, ITEMS, BLOCK>(j0, j1, col, val, x, prod);
  __syncthreads();

  // rows [i0, i1] meet this tile; row i1's end is not in it (i1 may be m)
  const int last = (i1 < m) ? i1 : i1 - 1;
  const int nrows = last - i0 + 1;
  const bool started_before = (i0 < m) && rp[i0] < j0;
  if (i0 == i1) {
    // no row ends here: the tile lies inside row i0 (or starts it)
    double s = 0.0;
    for (int e = t; e < j1 - j0; e += BLOCK) s += prod[e];
    s = lanes_sum<64>(s);
    if ((t & 63) == 0) wsum[t >> 6] = s;
    __syncthreads();
    if (t == 0) {
      double tot = 0.0;
#pragma unroll
      for (int w = 0; w < BLOCK / 64; w++) tot += wsum[w];
      TileCarry cy;
      cy.spanning = started_before ? 1 : 0;
      cy.head = started_before ? tot : 0.0;
      cy.head_row = started_before ? i0 : -1;
      cy.tail = started_before ? 0.0 : tot;
      cy.tail_row = started_before ? -1 : i0;
      cy.pad_ = 0;
      carry[b] = cy;
    }
    return;
  }
  constexpr int SUBS = BLOCK / SUBW;
  const int sub = t / SUBW, lane = t % SUBW;
  for (int k = sub; k < nrows; k += SUBS) {
    const int r = i0 + k;
    const int s0 = max(rp[r], j0) - j0, e0 = min(rp[r + 1], j1) - j0;
    double s = 0.0;
    for (int j = s0 + lane; j < e0; j += SUBW) s += prod[j];
    if constexpr (SUBW > 1) s = lanes_sum<SUBW>(s);
    if (lane == 0) ysh[k] = s;
  }
  __syncthreads();
  for (int k = t; k < i1 - i0; k += BLOCK) {
    const int r = i0 + k;
    if (k == 0 && started_before) continue;  // completed by the fix-up
    y[r] = axpby(alpha, ysh[k], beta, y[r]);
  }
  if (t == 0) {
    TileCarry cy;
    cy.spanning = 0;
    cy.head = started_before ? ysh[0] : 0.0;
    cy.head_row = started_before ? i0 : -1;
    const bool tail = (i1 < m) && max(rp[i1], j0) < j1;
    cy.tail = tail ? ysh[i1 - i0] : 0.0;
    cy.tail_row = tail ? i1 : -1;
    cy.pad_ = 0;
    carry[b] = cy;
  }
}

// ---------------------------------------------------------------------------
// COO segmented reduction: tile b owns elements [b*T, (b+1)*T) of the
// row-sorted arrays.  The tile's row segments (runs of equal row ids) are
// listed once by a workgroup prefix sum over row-start flags; sub-waves then
// sum one segment each, so the cost follows the number of non-empty rows, not
// the span of row ids.  Rows with no element are written (beta*y) by the
// segment just before them, or by the tile whose first element follows them;
// rows after the last element by the last tile.
// ---------------------------------------------------------------------------

template <typename VT, int ITEMS, int SUBW, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_coo_tiles(
    int m, int nnz, int tiles, const int *__restrict__ ri, const int *__restrict__ ci,
    const VT *__restrict__ val, const double *__restrict__ x, double *__restrict__ y,
    double alpha, double beta, TileCarry *__restrict__ carry) {
  constexpr int T = BLOCK * ITEMS;
  typedef hipcub::BlockScan<int, BLOCK> Scan;
  __shared__ double prod[T];
  __shared__ int rsh[T];
  __shared__ int seg[T + 1];  // first element of each segment, then ne
  __shared__ typename Scan::TempStorage scan_tmp;
  __shared__ double wsum[BLOCK / 64];
  const int b = blockIdx.x;
  if (b >= tiles) return;
  const int t = threadIdx.x;
  const int j0 = b * T;
  const int j1 = min(j0 + T, nnz);
  const int ne = j1 - j0;
  stage_products<VT, ITEMS, BLOCK>(j0, j1, ci, val, x, prod);
  for (int e = t; e < ne; e += BLOCK) rsh[e] = ri[j0 + e];
  __syncthreads();
  const int r_first = rsh[0], r_last = rsh[ne - 1];
  const bool head = j0 > 0 && ri[j0 - 1] == r_first;
  const bool tail = j1 < nnz && ri[j1] == r_last;

  // rows without elements before this tile's first row (and after the last
  // element for the last tile)
  const int gap_lo = (b == 0) ? 0 : ri[j0 - 1] + 1;
  for (int r = gap_lo + t; r < r_first; r += BLOCK) y[r] = axpby(alpha, 0.0, beta, y[r]);
  if (b == tiles - 1)
    for (int r = r_last + 1 + t; r < m; r += BLOCK) y[r] = axpby(alpha, 0.0, beta, y[r]);

  if (r_first == r_last && head && tail) {
    double s = 0.0;
    for (int e = t; e < ne; e += BLOCK) s += prod[e];
    s = lanes_sum<64>(s);
    if ((t & 63) == 0) wsum[t >> 6] = s;
    __syncthreads();
    if (t == 0) {
      double tot = 0.0;
#pragma unroll
      for (int w = 0; w < BLOCK / 64; w++) tot += wsum[w];
      TileCarry cy{tot, 0.0, r_first, -1, 1, 0};
      carry[b] = cy;
    }
    return;
  }
  // segment list: thread t flags its ITEMS consecutive elements
  int cnt = 0;
#pragma unroll
  for (int i = 0; i < ITEMS; i++) {
    const int e = t * ITEMS + i;
    cnt += (e < ne && (e == 0 || rsh[e] != rsh[e - 1])) ? 1 : 0;
  }
  int pos = 0, nseg = 0;
  Scan(scan_tmp).ExclusiveSum(cnt, pos, nseg);
#pragma unroll
  for (int i = 0; i < ITEMS; i++) {
    const int e = t * ITEMS + i;
    if (e < ne && (e == 0 || rsh[e] != rsh[e - 1])) seg[pos++] = e;
  }
  if (t == 0) seg[nseg] = ne;
  __syncthreads();
  constexpr int SUBS = BLOCK / SUBW;
  const int sub = t / SUBW, lane = t % SUBW;
  for (int k = sub; k < nseg; k += SUBS) {
    const int s0 = seg[k], e0 = seg[k + 1];
    const int r = rsh[s0];
    double s = 0.0;
    for (int j = s0 + lane; j < e0; j += SUBW) s += prod[j];
    if constexpr (SUBW > 1) s = lanes_sum<SUBW>(s);
    if (lane == 0) {
      if (k == 0 && head) {
        carry[b].head = s;
      } else if (k == nseg - 1 && tail) {
        carry[b].tail = s;
      } else {
        y[r] = axpby(alpha, s, beta, y[r]);
      }
    }
    if (k + 1 < nseg) {  // rows without elements up to the next segment's row
      const int rn = rsh[e0];
      for (int q = r + 1 + lane; q < rn; q += SUBW) y[q] = axpby(alpha, 0.0, beta, y[q]);
    }
  }
  if (t == 0) {
    carry[b].head_row = head ? r_first : -1;
    carry[b].tail_row = tail ? r_last : -1;
    carry[b].spanning = 0;
    if (!head) carry[b].head = 0.0;
    if (!tail) carry[b].tail = 0.0;
  }
}

// Rows cut by tile edges: one thread per tile whose last row continues; it
// adds the head partials of the following tiles in order (through tiles that
// lie wholly inside the row) and writes y once.
__global__ void k_tile_fixup(int tiles, const TileCarry *__restrict__ carry,
                             double *__restrict__ y, double alpha, double beta) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= tiles) return;
  const TileCarry c = carry[b];
  if (c.tail_row < 0 || c.spanning) return;
  const int r = c.tail_row;
  double s = c.tail;
  int n = b + 1;
  while (n < tiles && carry[n].spanning) s += carry[n++].head;
  if (n < tiles && carry[n].head_row == r) s += carry[n].head;
  y[r] = axpby(alpha, s, beta, y[r]);
}

// Host-side dispatch helpers ---------------------------------------------------

template <typename VT, int ITEMS>
int mp_launch(int sub, int m, int64_t nnz, int tiles, const int *rp, const int *col,
              const VT *val, const double *x, double *y, double alpha, double beta,
              TileCarry *carry, hipStream_t s) {
  constexpr int B = kTileBlock;
#define SBLAS_MP_CASE(W)                                                                   \
  case W:                                                                                  \
    hipLaunchKernelGGL((k_csr_merge_path<VT, ITEMS, W, B>), dim3(tiles), dim3(B), 0, s, m, \
                       static_cast<int>(nnz), tiles, rp, col, val, x, y, alpha, beta, carry); \
    break;
  switch (sub) {
    SBLAS_MP_CASE(1)
    SBLAS_MP_CASE(4)
    SBLAS_MP_CASE(16)
    SBLAS_MP_CASE(64)
    default:
      set_error(SBLAS_ERR_ARGS, "merge path: sub-wave must be 1, 4, 16 or 64");
      return SBLAS_ERR_ARGS;
  }
#undef SBLAS_MP_CASE
  SBLAS_HIP_TRY(hipGetLastError());
  return SBLAS_OK;
}

template <typename VT, int ITEMS>
int coo_launch(int sub, int m, int64_t nnz, int tiles, const int *ri, const int *ci,
               const VT *val, const double *x, double *y, double alpha, double beta,
               TileCarry *carry, hipStream_t s) {
  constexpr int B = kTileBlock;
#define SBLAS_COO_CASE(W)                                                                  \
  case W:                                                                                  \
    hipLaunchKernelGGL((k_coo_tiles<VT, ITEMS, W, B>), dim3(tiles), dim3(B), 0, s, m,      \
                       static_cast<int>(nnz), tiles, ri, ci, val, x, y, alpha, beta, carry); \
    break;
  switch (sub) {
    SBLAS_COO_CASE(1)
    SBLAS_COO_CASE(4)
    SBLAS_COO_CASE(16)
    SBLAS_COO_CASE(64)
    default:
      set_error(SBLAS_ERR_ARGS, "coo: sub-wave must be 1, 4, 16 or 64");
      return SBLAS_ERR_ARGS;
  }
#undef SBLAS_COO_CASE
  SBLAS_HIP_TRY(hipGetLastError());
  return SBLAS_OK;
}

}  // namespace

int tile_subwave(double mean_row) {
  return mean_row >= 48 ? 64 : (mean_row >= 12 ? 16 : (mean_row >= 3 ? 4 : 1));
}

int launch_csr_merge_path(int val_type, int sub, int m, int64_t nnz, int tiles, const int *rp,
                          const int *col, const void *val, const double *x, double *y,
                          double alpha, double beta, TileCarry *carry, hipStream_t s) {
  if (tiles <= 0) return SBLAS_OK;
  switch (val_type) {
    case SBLAS_VAL_F64:
      return mp_launch<double, kTileItems>(sub, m, nnz, tiles, rp, col,
                                           static_cast<const double *>(val), x, y, alpha, beta,
                                           carry, s);
    case SBLAS_VAL_F32:
      return mp_launch<float, kTileItems>(sub, m, nnz, tiles, rp, col,
                                          static_cast<const float *>(val), x, y, alpha, beta,
                                          carry, s);
    case SBLAS_VAL_BF16:
      return mp_launch<uint16_t, kTileItems>(sub, m, nnz, tiles, rp, col,
                                             static_cast<const uint16_t *>(val), x, y, alpha,
                                             beta, carry, s);
    default:
      set_error(SBLAS_ERR_ARGS, "unknown value type");
      return SBLAS_ERR_ARGS;
  }
}

int launch_coo_tiles(int val_type, int sub, int m, int64_t nnz, int tiles, const int *row_idx,
                     const int *col_idx, const void *val, const double *x, double *y,
                     double alpha, double beta, TileCarry *carry, hipStream_t s) {
  if (tiles <= 0) return SBLAS_OK;
  switch (val_type) {
    case SBLAS_VAL_F64:
      return coo_launch<double, kTileItems>(sub, m, nnz, tiles, row_idx, col_idx,
                                            static_cast<const double *>(val), x, y, alpha, beta,
                                            carry, s);
    case SBLAS_VAL_F32:
      return coo_launch<float, kTileItems>(sub, m, nnz, tiles, row_idx, col_idx,
                                           static_cast<const float *>(val), x, y, alpha, beta,
                                           carry, s);
    case SBLAS_VAL_BF16:
      return coo_launch<uint16_t, kTileItems>(sub, m, nnz, tiles, row_idx, col_idx,
                                              static_cast<const uint16_t *>(val), x, y, alpha,
                                              beta, carry, s);
    default:
      set_error(SBLAS_ERR_ARGS, "unknown value type");
      return SBLAS_ERR_ARGS;
  }
}

int launch_tile_fixup(int tiles, const TileCarry *carry, double *y, double alpha, double beta,
                      hipStream_t s) {
  if (tiles <= 0) return SBLAS_OK;
  hipLaunchKernelGGL(k_tile_fixup, dim3((tiles + 255) / 256), dim3(256), 0, s, tiles, carry, y,
                     alpha, beta);
  SBLAS_HIP_TRY(hipGetLastError());
  return SBLAS_OK;
}

}  // namespace sblas

// ---------------------------------------------------------------------------
// Merge-path SpMV that allocates nothing: the same two launches as a kernel-2
// plan (k_csr_merge_path, then k_tile_fixup), with the per-tile carries in a
// file-scope device array instead of a plan's workspace.  Every tile writes
// its carry on every call before the fix-up reads it, so nothing needs
// clearing between calls.  One array per GPU, shared by every call on that
// GPU: calls that overlap on one GPU race.
// ---------------------------------------------------------------------------
namespace sblas {
namespace {

constexpr int kNoAllocTiles = 262144;  // 8 MiB of carries: m + nnz <= 2^29 merge items
constexpr int kNoAllocDevices = 64;

static __device__ __align__(256) TileCarry scratch[kNoAllocTiles];

// Device address of `scratch` on each GPU, looked up once per device with
// that device current (each GPU holds its own copy of the array).
struct ScratchAddress {
  std::once_flag once;
  TileCarry *ptr = nullptr;
  hipError_t status = hipSuccess;
};
ScratchAddress g_scratch[kNoAllocDevices];

int scratch_for_current_device(TileCarry **out) {
  int dev = -1;
  SBLAS_HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kNoAllocDevices) {
    set_error(SBLAS_ERR_UNSUPPORTED, "spmv noalloc: device index beyond the scratch table");
    return SBLAS_ERR_UNSUPPORTED;
  }
  ScratchAddress &a = g_scratch[dev];
  std::call_once(a.once, [&a] {
    void *p = nullptr;
    a.status = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    a.ptr = static_cast<TileCarry *>(p);
  });
  if (a.status != hipSuccess || !a.ptr) {
    set_error(SBLAS_ERR_HIP, std::string("spmv noalloc: hipGetSymbolAddress(scratch): ") +
                                 hipGetErrorString(a.status));
    return SBLAS_ERR_HIP;
  }
  *out = a.ptr;
  return SBLAS_OK;
}

}  // namespace
}  // namespace sblas

extern "C" int sblas_spmv_csr_noalloc(int m, int n, int64_t nnz, const int *d_row_ptr,
                                      const int *d_col, const void *d_val, int val_type,
                                      double alpha, const double *d_x, double beta, double *d_y,
                                      void *stream) {
  using namespace sblas;
  clear_error();
  if (m < 0 || n < 0 || nnz < 0 || nnz > INT_MAX - 4096 || (m > 0 && (!d_row_ptr || !d_y)) ||
      (nnz > 0 && (!d_col || !d_val || !d_x)) || val_type < SBLAS_VAL_F64 ||
      val_type > SBLAS_VAL_BF16) {
    set_error(SBLAS_ERR_ARGS, "spmv noalloc: invalid sizes, pointers or value type");
    return SBLAS_ERR_ARGS;
  }
  if (m == 0) return SBLAS_OK;
  const int64_t items = static_cast<int64_t>(m) + nnz;
  const int64_t tiles = (items + kTileSize - 1) / kTileSize;
  if (tiles > kNoAllocTiles) {
    set_error(SBLAS_ERR_ARGS, "spmv noalloc: m + nnz exceeds 2^29 (the scratch holds 262144 "
                              "tile carries of 2048 merge items)");
    return SBLAS_ERR_ARGS;
  }
  TileCarry *carry = nullptr;
  int st = scratch_for_current_device(&carry);
  if (st != SBLAS_OK) return st;
  const hipStream_t s = static_cast<hipStream_t>(stream);
  const int sub = tile_subwave(static_cast<double>(nnz) / m);
  st = launch_csr_merge_path(val_type, sub, m, nnz, static_cast<int>(tiles), d_row_ptr, d_col,
                             d_val, d_x, d_y, alpha, beta, carry, s);
  if (st != SBLAS_OK) return st;
  return launch_tile_fixup(static_cast<int>(tiles), carry, d_y, alpha, beta, s);
}
