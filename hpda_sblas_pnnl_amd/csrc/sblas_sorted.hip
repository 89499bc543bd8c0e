// sblas_sorted.hip -- column-sorted row blocks: a form of kernel 1 (the
// reference's cusparseDcsrmv, src/dspmv_mgpu_v1.cu:261) for matrices whose x
// reads are scattered over a large x, such as power-law and R-MAT graphs.
//
// On such matrices a row-parallel kernel is bound by the L2: nearly every
// element's x read is its own request for a whole cache line.  Here each row
// block (up to 8192 elements of consecutive rows, at most kStreamRowsMax
// rows) stores its elements sorted by column, with each element's offset in
// the block's window.  A workgroup of 512 threads loads the sorted elements
// (consecutive threads on consecutive elements), reads x in column order --
// lanes of one load instruction then mostly fall on neighbouring columns, so
// they share cache lines and the L2 serves fewer requests -- and stores each
// product into LDS at the element's own offset.  The rows are then summed
// from LDS in element order, SUBW lanes per row, as in the row-block kernel.
// A row longer than a block is cut into block-sized pieces whose partials
// the device merge adds in block order.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <climits>

#include "internal.hpp"
#include "rows_device.hpp"

namespace sblas {
namespace {

constexpr int B = kSortedBlock;
constexpr int IT = kSortedItems;

__device__ __forceinline__ double sload_val(const double *p, int j) {
  return __builtin_nontemporal_load(p + j);
}
__device__ __forceinline__ double sload_val(const float *p, int j) {
  return static_cast<double>(__builtin_nontemporal_load(p + j));
}
__device__ __forceinline__ double sload_val(const uint16_t *p, int j) {
  const uint32_t h = __builtin_nontemporal_load(p + j);
  return static_cast<double>(__uint_as_float(h << 16));
}

template <int SUBW>
__device__ __forceinline__ double lanes_sum(double v) {
#pragma unroll
  for (int off = SUBW / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, SUBW);
  return v;
}

template <typename VT, int SUBW>
__global__ __launch_bounds__(B) __attribute__((amdgpu_waves_per_eu(4))) void k_csr_sorted(
    const RowBlock *__restrict__ blocks, int nblocks, int xcd_remap,
    const int *__restrict__ rp, const int *__restrict__ scol, const VT *__restrict__ sval,
    const uint16_t *__restrict__ spos, const double *__restrict__ x, double *__restrict__ y,
    double alpha, double beta, double *__restrict__ partials) {
  constexpr int RPT = (kStreamRowsMax + B) / B;  // row pointers per thread
  constexpr int YPT = kStreamRowsMax / B;        // rows per thread in the y writes
  __shared__ double prod[kSortedCap];
  __shared__ double ysh[kStreamRowsMax];
  __shared__ int rps[kStreamRowsMax + 1];
  __shared__ double wsum[B / 64];
  int b = blockIdx.x;
  if (xcd_remap) {  // each XCD takes a contiguous eighth of the blocks
    const int per = gridDim.x >> 3;
    b = (b & 7) * per + (b >> 3);
  }
  if (b >= nblocks) return;  // uniform per workgroup
  const RowBlock rb = blocks[b];
  const int lo = rb.lo, hi = rb.hi;
  const int t = threadIdx.x;
  const int r0 = rb.row_begin, nrows = rb.row_end - rb.row_begin;

  // every independent load first: sorted elements, row pointers, y
  const int rp_first = rp[r0], rp_second = rp[r0 + 1];
  int rpv[RPT];
#pragma unroll
  for (int k = 0; k < RPT; k++) rpv[k] = rp[r0 + min(t + k * B, nrows)];
  double yv[YPT];
  if (beta != 0.0) {
#pragma unroll
    for (int k = 0; k < YPT; k++) yv[k] = y[r0 + min(t + k * B, nrows - 1)];
  } else {
#pragma unroll
    for (int k = 0; k < YPT; k++) yv[k] = 0.0;
  }
  double p[IT];
  int pos[IT];
  if (hi > lo) {  // uniform: a block of empty rows has no element
    double v[IT];
    int c[IT];
#pragma unroll
    for (int k = 0; k < IT; k++) {
      const int e = min(lo + t + k * B, hi - 1);
      v[k] = sload_val(sval, e);
      c[k] = __builtin_nontemporal_load(scol + e);
      pos[k] = __builtin_nontemporal_load(spos + e);
    }
    double xv[IT];
#pragma unroll
    for (int k = 0; k < IT; k++) xv[k] = x[c[k]];
#pragma unroll
    for (int k = 0; k < IT; k++) p[k] = (lo + t + k * B < hi) ? v[k] * xv[k] : 0.0;
  } else {
#pragma unroll
    for (int k = 0; k < IT; k++) {
      p[k] = 0.0;
      pos[k] = 0;
    }
  }

  if (nrows == 1 && (rp_first < lo || rp_second > hi)) {
    // piece of a long row: register sums in the stored (column) order, wave
    // reduction, then the waves in order
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < IT; k++) s += p[k];
    s = lanes_sum<64>(s);
    if ((t & 63) == 0) wsum[t >> 6] = s;
    __syncthreads();
    if (t == 0) {
      double tot = 0.0;
#pragma unroll
      for (int w = 0; w < B / 64; w++) tot += wsum[w];
      partials[2 * b] = tot;
    }
    return;
  }
  // each product to its element's place in the window
#pragma unroll
  for (int k = 0; k < IT; k++)
    if (lo + t + k * B < hi) prod[pos[k]] = p[k];
#pragma unroll
  for (int k = 0; k < RPT; k++)
    if (t + k * B <= nrows) rps[t + k * B] = rpv[k];
  __syncthreads();

  // row sums in element order: one sub-wave of SUBW lanes per row, whole
  // waves for long rows
  rows::block_row_sums<SUBW, B>(prod, rps, nrows, lo, hi, ysh);
  __syncthreads();

  // y for whole rows (consecutive threads on consecutive rows), partials for
  // the rows the window cuts
#pragma unroll
  for (int k = 0; k < YPT; k++) {
    const int i = t + k * B;
    if (i < nrows) {
      const double s = ysh[i];
      const bool edge = (i == 0 || i == nrows - 1);
      if (edge && (rps[i] < lo || rps[i + 1] > hi)) {
        partials[2 * b + (i == 0 ? 0 : 1)] = s;
      } else {
        y[r0 + i] = (beta == 0.0) ? alpha * s : alpha * s + beta * yv[k];
      }
    }
  }
}

// Format build: one workgroup per block sorts the block's columns (stable
// LSD radix sort over the bits n needs; elements with equal columns keep
// their order) with their offsets, then writes the sorted columns, offsets
// and values.  Items are held blocked (thread t: offsets 16t .. 16t+15), so
// the input order the sort keeps is the element order.
template <typename VT>
__global__ __launch_bounds__(B) void k_sorted_build(const RowBlock *__restrict__ blocks,
                                                    const int *__restrict__ col,
                                                    const VT *__restrict__ val, int bits,
                                                    int *__restrict__ scol,
                                                    VT *__restrict__ sval,
                                                    uint16_t *__restrict__ spos) {
  typedef hipcub::BlockRadixSort<unsigned int, B, IT, unsigned short> Sort;
  __shared__ typename Sort::TempStorage tmp;
  const RowBlock rb = blocks[blockIdx.x];
  const int lo = rb.lo, count = rb.hi - rb.lo;
  const int t = threadIdx.x;
  unsigned int keys[IT];
  unsigned short offs[IT];
#pragma unroll
  for (int k = 0; k < IT; k++) {
    const int o = t * IT + k;
    // past the window: keys with every sorted bit set sort after (or, equal,
    // stay after) every element
    keys[k] = o < count ? static_cast<unsigned int>(col[lo + o]) : 0xffffffffu;
    offs[k] = static_cast<unsigned short>(o);
  }
  Sort(tmp).Sort(keys, offs, 0, bits);
#pragma unroll
  for (int k = 0; k < IT; k++) {
    const int r = t * IT + k;
    if (r < count) {
      scol[lo + r] = static_cast<int>(keys[k]);
      spos[lo + r] = offs[k];
      sval[lo + r] = val[lo + offs[k]];
    }
  }
}

template <typename VT>
int sorted_launch_typed(const SortedFormat &f, const int *rp, const double *x, double *y,
                        double alpha, double beta, hipStream_t s) {
  const dim3 grid(f.grid), block(B);
  const VT *sv = static_cast<const VT *>(f.sval);
#define SBLAS_SORTED_CASE(W)                                                                  \
  case W:                                                                                     \
    hipLaunchKernelGGL((k_csr_sorted<VT, W>), grid, block, 0, s, f.blocks, f.nblocks,         \
                       f.xcd_remap, rp, f.scol, sv, f.spos, x, y, alpha, beta, f.partials);  \
    break;
  switch (f.subwave) {
    SBLAS_SORTED_CASE(1)
    SBLAS_SORTED_CASE(2)
    SBLAS_SORTED_CASE(4)
    SBLAS_SORTED_CASE(8)
    SBLAS_SORTED_CASE(16)
    SBLAS_SORTED_CASE(32)
    SBLAS_SORTED_CASE(64)
    default:
      set_error(SBLAS_ERR_ARGS, "sorted blocks: sub-wave must be a power of two in [1,64]");
      return SBLAS_ERR_ARGS;
  }
#undef SBLAS_SORTED_CASE
  SBLAS_HIP_TRY(hipGetLastError());
  return SBLAS_OK;
}

}  // namespace

int sorted_build(int val_type, int n, const int *col, const void *val,
                 const std::vector<RowBlock> &hblocks, int subwave, int xcd_remap,
                 SortedFormat &f) {
  f.hblocks = hblocks;
  f.nblocks = static_cast<int>(hblocks.size());
  f.subwave = subwave;
  f.xcd_remap = xcd_remap ? 1 : 0;
  f.grid = f.xcd_remap ? ((f.nblocks + 7) / 8) * 8 : f.nblocks;
  int64_t nnz = 0;
  for (const RowBlock &r : hblocks) {
    if (r.hi - r.lo > kSortedCap) {
      set_error(SBLAS_ERR_ARGS, "sorted blocks: a block holds more than 8192 elements");
      return SBLAS_ERR_ARGS;
    }
    nnz = std::max<int64_t>(nnz, r.hi);
  }
  const size_t vbytes = val_type == SBLAS_VAL_F64 ? 8 : (val_type == SBLAS_VAL_F32 ? 4 : 2);
  auto alloc = [&](size_t bytes) -> void * {
    void *p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(bytes, 16)) != hipSuccess) return nullptr;
    f.buffers.push_back(p);
    f.bytes += static_cast<int64_t>(std::max<size_t>(bytes, 16));
    return p;
  };
  f.blocks = static_cast<RowBlock *>(alloc(sizeof(RowBlock) * std::max(f.nblocks, 1)));
  f.scol = static_cast<int *>(alloc(sizeof(int) * nnz));
  f.sval = alloc(vbytes * nnz);
  f.spos = static_cast<uint16_t *>(alloc(sizeof(uint16_t) * nnz));
  f.partials = static_cast<double *>(alloc(sizeof(double) * 2 * std::max(f.nblocks, 1)));
  if (!f.blocks || !f.scol || !f.sval || !f.spos || !f.partials) {
    set_error(SBLAS_ERR_ALLOC, "sorted blocks: device allocation failed");
    return SBLAS_ERR_ALLOC;
  }
  if (f.nblocks == 0) return SBLAS_OK;
  SBLAS_HIP_TRY(hipMemcpy(f.blocks, hblocks.data(), sizeof(RowBlock) * f.nblocks,
                          hipMemcpyHostToDevice));
  int bits = 1;
  while (bits < 31 && (int64_t{1} << bits) < n) bits++;
  switch (val_type) {
    case SBLAS_VAL_F64:
      hipLaunchKernelGGL(k_sorted_build<double>, dim3(f.nblocks), dim3(B), 0, nullptr, f.blocks,
                         col, static_cast<const double *>(val), bits, f.scol,
                         static_cast<double *>(f.sval), f.spos);
      break;
    case SBLAS_VAL_F32:
      hipLaunchKernelGGL(k_sorted_build<float>, dim3(f.nblocks), dim3(B), 0, nullptr, f.blocks,
                         col, static_cast<const float *>(val), bits, f.scol,
                         static_cast<float *>(f.sval), f.spos);
      break;
    default:
      hipLaunchKernelGGL(k_sorted_build<uint16_t>, dim3(f.nblocks), dim3(B), 0, nullptr,
                         f.blocks, col, static_cast<const uint16_t *>(val), bits, f.scol,
                         static_cast<uint16_t *>(f.sval), f.spos);
      break;
  }
  SBLAS_HIP_TRY(hipGetLastError());
  SBLAS_HIP_TRY(hipStreamSynchronize(nullptr));
  return SBLAS_OK;
}

int launch_csr_sorted(int val_type, const SortedFormat &f, const int *rp, const double *x,
                      double *y, double alpha, double beta, hipStream_t s) {
  if (f.nblocks <= 0) return SBLAS_OK;
  switch (val_type) {
    case SBLAS_VAL_F64: return sorted_launch_typed<double>(f, rp, x, y, alpha, beta, s);
    case SBLAS_VAL_F32: return sorted_launch_typed<float>(f, rp, x, y, alpha, beta, s);
    default: return sorted_launch_typed<uint16_t>(f, rp, x, y, alpha, beta, s);
  }
}

}  // namespace sblas
