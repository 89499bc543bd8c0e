// sblas_csr5.hip -- CSR5-style tiled SpMV for gfx950 (the reference's kernel 3,
// csr5_kernel at src/csr5_kernel.cu:12-31 over the vendored CSR5 library,
// include/anonymouslib_cuda.h and include/detail/cuda/csr5_spmv_cuda.h).
//
// Written from the published method (Liu & Vinter, "CSR5: An Efficient
// Storage Format for Cross-Platform Sparse Matrix-Vector Multiplication",
// ICS'15): the element stream is cut into tiles of omega x sigma elements;
// lane l of a tile owns sigma consecutive elements and the tile is stored
// transposed so that the s-th load of all lanes is one contiguous run; a bit
// per element marks row starts.  Here omega = 64 (one wave64 per tile), and:
//   * the format keeps, per tile, the index of its first row among the
//     non-empty rows (bit 31: the tile begins exactly at a row start) and,
//     per lane, a 32-bit word of row-start flags; the rank of a row start in
//     the tile comes from a prefix sum of flag counts over lanes, and the
//     list of non-empty rows maps it to the row (empty rows are written by
//     a separate pass);
//   * each lane sums its segments sequentially in FP64; rows that end inside
//     one lane are written directly;
//   * rows crossing lanes are finished by a segmented inclusive scan across
//     the wave on DPP lane moves (row_shr inside 16-lane rows, row_bcast
//     across them); the four tiles of a workgroup stage their carries in LDS
//     and fold them in order, and rows crossing workgroups are finished by
//     the ordered carry fix-up shared with the merge-path and COO kernels.
//     alpha and beta are applied (the vendored CSR5 path ignores them, csr5_spmv_cuda.h:22,350)
//     and no floating-point atomics are used (it adds tile carries with
//     atomicAdd, :350,377), so results do not depend on scheduling.
#include "internal.hpp"

namespace sblas {
namespace {

__device__ __forceinline__ double c5_val(const double *p, int64_t j) {
  return __builtin_nontemporal_load(p + j);
}
__device__ __forceinline__ double c5_val(const float *p, int64_t j) {
  return static_cast<double>(__builtin_nontemporal_load(p + j));
}
__device__ __forceinline__ double c5_val(const uint16_t *p, int64_t j) {
  const uint32_t h = __builtin_nontemporal_load(p + j);
  return static_cast<double>(__uint_as_float(h << 16));
}

__device__ __forceinline__ double c5_axpby(double alpha, double s, double beta, double yv) {
  return (beta == 0.0) ? alpha * s : alpha * s + beta * yv;
}

// Wave-wide scans on DPP lane moves.  Lanes whose source lies outside the
// move (row_shr past a 16-lane row start, rows that a broadcast does not
// reach) read 0, the identity of both scans.
//   row_shr:d   lane l reads l-d inside its row of 16 (d = 1, 2, 4, 8)
//   row_bcast15 rows 1 and 3 read lane 15 of the row before (row mask 0xa)
//   row_bcast31 rows 2 and 3 read lane 31 (row mask 0xc)
// so after the four in-row steps each lane holds its row prefix, and the two
// broadcasts add the totals of the rows before it.
template <int CTRL, int ROWS>
__device__ __forceinline__ double dpp_mov(double v) {
  return __builtin_amdgcn_update_dpp(0.0, v, CTRL, ROWS, 0xf, false);
}
template <int CTRL, int ROWS>
__device__ __forceinline__ int dpp_mov(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, ROWS, 0xf, false);
}
constexpr int kRowShr1 = 0x111, kRowShr2 = 0x112, kRowShr4 = 0x114, kRowShr8 = 0x118;
constexpr int kRowBcast15 = 0x142, kRowBcast31 = 0x143;

// inclusive prefix sum over the 64 lanes
__device__ __forceinline__ int wave_incl_sum(int v) {
  v += dpp_mov<kRowShr1, 0xf>(v);
  v += dpp_mov<kRowShr2, 0xf>(v);
  v += dpp_mov<kRowShr4, 0xf>(v);
  v += dpp_mov<kRowShr8, 0xf>(v);
  v += dpp_mov<kRowBcast15, 0xa>(v);
  v += dpp_mov<kRowBcast31, 0xc>(v);
  return v;
}

// One step of the segmented scan: (v, head) absorbs the moved-in partner
// unless a segment already begins inside this lane's range.
template <int CTRL, int ROWS>
__device__ __forceinline__ void seg_step(double &v, int &head) {
  const double vu = dpp_mov<CTRL, ROWS>(v);
  const int hu = dpp_mov<CTRL, ROWS>(head);
  if (!head) {
    v += vu;
    head = hu;
  }
}

// inclusive segmented sum over the 64 lanes; head = 1 opens a segment
__device__ __forceinline__ double wave_seg_sum(double v, int head) {
  seg_step<kRowShr1, 0xf>(v, head);
  seg_step<kRowShr2, 0xf>(v, head);
  seg_step<kRowShr4, 0xf>(v, head);
  seg_step<kRowShr8, 0xf>(v, head);
  seg_step<kRowBcast15, 0xa>(v, head);
  seg_step<kRowBcast31, 0xc>(v, head);
  return v;
}

// One tile by one wave: rows closed inside the tile are written; the tile's
// head / tail / spanning partials go to *out (lane 0).
template <typename VT, int SIGMA>
__device__ __forceinline__ void csr5_tile(
    int p, int ntiles, int64_t nnz, const uint32_t *__restrict__ tile_nz,
    const uint32_t *__restrict__ flags, const int *__restrict__ rows_nz,
    const int *__restrict__ colT, const VT *__restrict__ valT, const double *__restrict__ x,
    double *__restrict__ y, double alpha, double beta, TileCarry *out) {
  constexpr int T = 64 * SIGMA;
  const int lane = threadIdx.x & 63;
  const int64_t base = static_cast<int64_t>(p) * T;
  const uint32_t tnz = tile_nz[p];
  const int first_nz = static_cast<int>(tnz & 0x7fffffffu);
  const bool starts_row = (tnz >> 31) != 0;
  const bool ends_row = (p + 1 == ntiles) || (tile_nz[p + 1] >> 31) != 0;
  const uint32_t f = flags[static_cast<int64_t>(p) * 64 + lane];

  // loads: the s-th element of every lane is one contiguous run of 64
  double v[SIGMA];
  int c[SIGMA];
#pragma unroll
  for (int s = 0; s < SIGMA; s++) {
    const int64_t o = base + s * 64 + lane;
    v[s] = c5_val(valT, o);
    c[s] = __builtin_nontemporal_load(colT + o);
  }
  double xv[SIGMA];
#pragma unroll
  for (int s = 0; s < SIGMA; s++) xv[s] = x[c[s]];

  // rank of this lane's first row start among the tile's row starts
  const int cnt = __popc(f);
  const int incl = wave_incl_sum(cnt);
  const int excl = incl - cnt;
  const int total_starts = __builtin_amdgcn_readlane(incl, 63);

  // in-lane segments
  const int64_t lane_first = base + static_cast<int64_t>(lane) * SIGMA;
  double first = 0.0, cur = 0.0;
  int k = 0;  // row starts seen in this lane
#pragma unroll
  for (int s = 0; s < SIGMA; s++) {
    const double prod = (lane_first + s < nnz) ? v[s] * xv[s] : 0.0;
    if ((f >> s) & 1u) {
      if (k == 0) {
        first = cur;
      } else {
        const int r = rows_nz[first_nz + excl + k];
        y[r] = c5_axpby(alpha, cur, beta, y[r]);
      }
      k++;
      cur = 0.0;
    }
    cur += prod;
  }
  const bool has_start = cnt > 0;
  const double a = has_start ? first : cur;      // part before this lane's first start
  const double last = has_start ? cur : 0.0;     // part after its last start

  // segmented inclusive scan of a; a segment begins after each lane with a
  // start (lane 0 always begins one)
  const unsigned long long starts_mask = __ballot(has_start);
  const int head = (lane == 0) ? 1 : static_cast<int>((starts_mask >> (lane - 1)) & 1ull);
  const double sv = wave_seg_sum(a, head);
  const unsigned long long below = starts_mask & ((1ull << lane) - 1ull);
  const int prev = below ? 63 - __clzll(static_cast<long long>(below)) : -1;
  const double prev_last = __shfl(last, prev < 0 ? 0 : prev, 64);

  // the row closed by this lane's first start
  double head_part = 0.0;
  int head_row = -1;
  if (has_start) {
    const double total = sv + (prev >= 0 ? prev_last : 0.0);
    const int r = rows_nz[first_nz + excl];
    if (excl == 0 && !starts_row) {
      head_part = total;  // began in an earlier tile
      head_row = r;
    } else {
      y[r] = c5_axpby(alpha, total, beta, y[r]);
    }
  }
  // the segment open at the tile's end
  const unsigned long long any = starts_mask;
  const int last_lane = any ? 63 - __clzll(static_cast<long long>(any)) : -1;
  const double tail_scan = __shfl(sv, 63, 64);  // uniform source lane
  const double open_last = __shfl(last, last_lane < 0 ? 0 : last_lane, 64);
  // the head partial (if any) sits in the lowest lane with a start
  const int first_lane = any ? __ffsll(static_cast<long long>(any)) - 1 : 0;
  const double head_part0 = __shfl(head_part, first_lane, 64);
  const int head_row0 = __shfl(head_row, first_lane, 64);
  if (lane == 0) {
    TileCarry cy{head_part0, 0.0, head_row0, -1, 0, 0};
    if (head_row0 < 0) cy.head = 0.0;
    const int open_row = rows_nz[first_nz + total_starts];
    double open_sum;
    if (last_lane < 0) {
      open_sum = tail_scan;  // no start in the tile: one segment from lane 0
    } else {
      open_sum = open_last + (last_lane < 63 ? tail_scan : 0.0);
    }
    const bool open_began_before = (last_lane < 0) && !starts_row;
    if (ends_row) {
      if (open_began_before) {
        cy.head = open_sum;
        cy.head_row = open_row;
      } else {
        y[open_row] = c5_axpby(alpha, open_sum, beta, y[open_row]);
      }
    } else {
      if (open_began_before) {
        cy.head = open_sum;
        cy.head_row = open_row;
        cy.spanning = 1;
      } else {
        cy.tail = open_sum;
        cy.tail_row = open_row;
      }
    }
    *out = cy;
  }
}

// Four tiles per workgroup.  Their carries are staged in LDS and folded in
// tile order by one thread: rows that begin and end inside the workgroup are
// written here, and one carry per workgroup (same meaning as a tile's) is
// left for the ordered fix-up over workgroups.
template <typename VT, int SIGMA>
__global__ __launch_bounds__(256) void k_csr5(
    int ntiles, int64_t nnz, const uint32_t *__restrict__ tile_nz,
    const uint32_t *__restrict__ flags, const int *__restrict__ rows_nz,
    const int *__restrict__ colT, const VT *__restrict__ valT, const double *__restrict__ x,
    double *__restrict__ y, double alpha, double beta, TileCarry *__restrict__ carry) {
  __shared__ TileCarry wc[4];
  const int wave = threadIdx.x >> 6;
  const int p = blockIdx.x * 4 + wave;
  if (p < ntiles)  // uniform per wave; every wave reaches the barrier below
    csr5_tile<VT, SIGMA>(p, ntiles, nnz, tile_nz, flags, rows_nz, colT, valT, x, y, alpha, beta,
                         &wc[wave]);
  __syncthreads();
  if (threadIdx.x != 0) return;
  const int count = min(4, ntiles - static_cast<int>(blockIdx.x) * 4);
  TileCarry out{0.0, 0.0, -1, -1, 0, 0};
  bool all_span = true;
  int open = -1;  // row opened by a tile of this workgroup and not yet closed
  double osum = 0.0;
  for (int i = 0; i < count; i++) {
    const TileCarry c = wc[i];
    if (c.spanning) {
      if (open >= 0) {
        osum += c.head;
      } else {
        out.head += c.head;
        out.head_row = c.head_row;
      }
      continue;
    }
    all_span = false;
    if (c.head_row >= 0) {
      if (open >= 0) {
        y[open] = c5_axpby(alpha, osum + c.head, beta, y[open]);
        open = -1;
      } else {
        out.head += c.head;
        out.head_row = c.head_row;
      }
    } else if (open >= 0) {
      y[open] = c5_axpby(alpha, osum, beta, y[open]);
      open = -1;
    }
    if (c.tail_row >= 0) {
      open = c.tail_row;
      osum = c.tail;
    }
  }
  if (all_span) {
    out.spanning = 1;
  } else if (open >= 0) {
    out.tail = osum;
    out.tail_row = open;
  }
  carry[blockIdx.x] = out;
}

// Transposed, zero-padded copy: element (lane l, s) of tile p moves to
// p*T + s*64 + l.
template <typename VT>
__global__ void k_csr5_transpose(int64_t padded, int64_t nnz, int sigma, const int *__restrict__ col,
                                 const VT *__restrict__ val, int *__restrict__ colT,
                                 VT *__restrict__ valT) {
  const int64_t o = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (o >= padded) return;
  const int64_t T = 64LL * sigma;
  const int64_t p = o / T, w = o % T;
  const int64_t s = w / 64, l = w % 64;
  const int64_t src = p * T + l * sigma + s;
  colT[o] = src < nnz ? col[src] : 0;
  valT[o] = src < nnz ? val[src] : VT(0);
}

__global__ void k_scale_rows(int count, const int *__restrict__ rows, double beta,
                             double *__restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const int r = rows[i];
  y[r] = (beta == 0.0) ? 0.0 : beta * y[r];
}

template <typename VT>
int csr5_launch_sigma(int sigma, int ntiles, int64_t nnz, const uint32_t *tile_nz,
                      const uint32_t *flags, const int *rows_nz, const int *colT,
                      const VT *valT, const double *x, double *y, double alpha, double beta,
                      TileCarry *carry, hipStream_t s) {
  const dim3 grid((ntiles + 3) / 4), block(256);
  switch (sigma) {
    case 8:
      hipLaunchKernelGGL((k_csr5<VT, 8>), grid, block, 0, s, ntiles, nnz, tile_nz, flags, rows_nz,
                         colT, valT, x, y, alpha, beta, carry);
      break;
    case 16:
      hipLaunchKernelGGL((k_csr5<VT, 16>), grid, block, 0, s, ntiles, nnz, tile_nz, flags,
                         rows_nz, colT, valT, x, y, alpha, beta, carry);
      break;
    case 32:
      hipLaunchKernelGGL((k_csr5<VT, 32>), grid, block, 0, s, ntiles, nnz, tile_nz, flags,
                         rows_nz, colT, valT, x, y, alpha, beta, carry);
      break;
    default:
      set_error(SBLAS_ERR_ARGS, "csr5: sigma must be 8, 16 or 32");
      return SBLAS_ERR_ARGS;
  }
  SBLAS_HIP_TRY(hipGetLastError());
  return SBLAS_OK;
}

}  // namespace

int launch_csr5(int val_type, int sigma, int ntiles, int64_t nnz, const uint32_t *tile_nz,
                const uint32_t *flags, const int *rows_nz, const int *colT, const void *valT,
                const double *x, double *y, double alpha, double beta, TileCarry *carry,
                hipStream_t s) {
  if (ntiles <= 0) return SBLAS_OK;
  switch (val_type) {
    case SBLAS_VAL_F64:
      return csr5_launch_sigma(sigma, ntiles, nnz, tile_nz, flags, rows_nz, colT,
                               static_cast<const double *>(valT), x, y, alpha, beta, carry, s);
    case SBLAS_VAL_F32:
      return csr5_launch_sigma(sigma, ntiles, nnz, tile_nz, flags, rows_nz, colT,
                               static_cast<const float *>(valT), x, y, alpha, beta, carry, s);
    case SBLAS_VAL_BF16:
      return csr5_launch_sigma(sigma, ntiles, nnz, tile_nz, flags, rows_nz, colT,
                               static_cast<const uint16_t *>(valT), x, y, alpha, beta, carry, s);
    default:
      set_error(SBLAS_ERR_ARGS, "unknown value type");
      return SBLAS_ERR_ARGS;
  }
}

int launch_csr5_transpose(int val_type, int sigma, int64_t padded, int64_t nnz, const int *col,
                          const void *val, int *colT, void *valT, hipStream_t s) {
  if (padded <= 0) return SBLAS_OK;
  const dim3 grid(static_cast<unsigned>((padded + 255) / 256)), block(256);
  switch (val_type) {
    case SBLAS_VAL_F64:
      hipLaunchKernelGGL(k_csr5_transpose<double>, grid, block, 0, s, padded, nnz, sigma, col,
                         static_cast<const double *>(val), colT, static_cast<double *>(valT));
      break;
    case SBLAS_VAL_F32:
      hipLaunchKernelGGL(k_csr5_transpose<float>, grid, block, 0, s, padded, nnz, sigma, col,
                         static_cast<const float *>(val), colT, static_cast<float *>(valT));
      break;
    default:
      hipLaunchKernelGGL(k_csr5_transpose<uint16_t>, grid, block, 0, s, padded, nnz, sigma, col,
                         static_cast<const uint16_t *>(val), colT,
                         static_cast<uint16_t *>(valT));
      break;
  }
  SBLAS_HIP_TRY(hipGetLastError());
  return SBLAS_OK;
}

int launch_scale_rows(int count, const int *rows, double beta, double *y, hipStream_t s) {
  if (count <= 0) return SBLAS_OK;
  hipLaunchKernelGGL(k_scale_rows, dim3((count + 255) / 256), dim3(256), 0, s, count, rows, beta,
                     y);
  SBLAS_HIP_TRY(hipGetLastError());
  return SBLAS_OK;
}

}  // namespace sblas
