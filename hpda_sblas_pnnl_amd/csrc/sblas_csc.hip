// sblas_csc.hip -- deterministic CSC: a stable by-row transpose on the device
// followed by the CSR kernels.
//
// The reference runs CSC as cusparseDcsrmv(TRANSPOSE) per column block and
// adds the blocks' full-length partial y on the host
// (src/dspmv_mgpu_baseline_csc.cu:131,171-175, src/dspmv_mgpu_v1_numa_csc.cu:
// 382,427-435); the transposed product scatters into y, so its sums have no
// fixed order.  Here the plan builds, once, the row-major form of the matrix:
// row indices are sorted with a stable LSD radix sort keyed by row, so the
// elements of each row keep the CSC (ascending column) order, and every row
// is then summed by one CSR kernel in a fixed order.  Indices and the sort
// are integer work; no floating-point atomics are involved.
#include <hipcub/hipcub.hpp>

#include "internal.hpp"

namespace sblas {
namespace {

// column index of every CSC element (upper bound of its position in col_ptr)
__global__ void k_expand_cols(int n, int nnz, const int *__restrict__ cp, int *__restrict__ cols) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nnz) return;
  int lo = 0, hi = n;  // last c with cp[c] <= j
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (cp[mid] <= j) lo = mid; else hi = mid;
  }
  cols[j] = lo;
}

__global__ void k_iota(int nnz, int *__restrict__ out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < nnz) out[j] = j;
}

// row pointers from sorted row keys: rp[r] = first position with key >= r
__global__ void k_row_ptr_from_sorted(int m, int nnz, const int *__restrict__ keys,
                                      int *__restrict__ rp) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r > m) return;
  int lo = 0, hi = nnz;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (keys[mid] < r) lo = mid + 1; else hi = mid;
  }
  rp[r] = lo;
}

template <typename VT>
__global__ void k_gather_csr(int nnz, const int *__restrict__ perm, const int *__restrict__ cols,
                             const VT *__restrict__ val, int *__restrict__ out_col,
                             VT *__restrict__ out_val) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nnz) return;
  const int j = perm[k];
  out_col[k] = cols[j];
  out_val[k] = val[j];
}

int bits_for(int m) {
  int b = 1;
  while (b < 31 && (1 << b) < m) b++;
  return b;
}

}  // namespace

// Builds row_ptr[m+1], col[nnz], val_out[nnz] (device, caller-allocated) from
// CSC (cp[n+1], ri[nnz], val[nnz]).  Scratch is allocated and freed here.
int csc_to_csr_device(int m, int n, int nnz, const int *cp, const int *ri, const void *val,
                      int val_type, int *rp, int *col, void *val_out, hipStream_t s) {
  if (nnz == 0) {
    SBLAS_HIP_TRY(hipMemsetAsync(rp, 0, sizeof(int) * (static_cast<size_t>(m) + 1), s));
    return SBLAS_OK;
  }
  int *cols = nullptr, *keys_out = nullptr, *perm_in = nullptr, *perm_out = nullptr;
  void *temp = nullptr;
  size_t temp_bytes = 0;
  int st = SBLAS_OK;
  auto fail = [&](const char *what, hipError_t e) {
    set_error(SBLAS_ERR_HIP, std::string("csc transpose: ") + what + ": " + hipGetErrorString(e));
    st = SBLAS_ERR_HIP;
  };
  hipError_t e;
  const size_t bytes = sizeof(int) * static_cast<size_t>(nnz);
  if ((e = hipMalloc(&cols, bytes)) != hipSuccess) fail("alloc", e);
  if (st == SBLAS_OK && (e = hipMalloc(&keys_out, bytes)) != hipSuccess) fail("alloc", e);
  if (st == SBLAS_OK && (e = hipMalloc(&perm_in, bytes)) != hipSuccess) fail("alloc", e);
  if (st == SBLAS_OK && (e = hipMalloc(&perm_out, bytes)) != hipSuccess) fail("alloc", e);
  const int end_bit = bits_for(m);
  if (st == SBLAS_OK) {
    e = hipcub::DeviceRadixSort::SortPairs(nullptr, temp_bytes, ri, keys_out, perm_in, perm_out,
                                           nnz, 0, end_bit, s);
    if (e != hipSuccess) fail("sort size", e);
  }
  if (st == SBLAS_OK && (e = hipMalloc(&temp, temp_bytes)) != hipSuccess) fail("alloc", e);
  if (st == SBLAS_OK) {
    const dim3 blk(256), grd((nnz + 255) / 256);
    hipLaunchKernelGGL(k_expand_cols, grd, blk, 0, s, n, nnz, cp, cols);
    hipLaunchKernelGGL(k_iota, grd, blk, 0, s, nnz, perm_in);
    e = hipcub::DeviceRadixSort::SortPairs(temp, temp_bytes, ri, keys_out, perm_in, perm_out,
                                           nnz, 0, end_bit, s);
    if (e != hipSuccess) fail("sort", e);
  }
  if (st == SBLAS_OK) {
    hipLaunchKernelGGL(k_row_ptr_from_sorted, dim3((m + 1 + 255) / 256), dim3(256), 0, s, m, nnz,
                       keys_out, rp);
    const dim3 blk(256), grd((nnz + 255) / 256);
    switch (val_type) {
      case SBLAS_VAL_F64:
        hipLaunchKernelGGL(k_gather_csr<double>, grd, blk, 0, s, nnz, perm_out, cols,
                           static_cast<const double *>(val), col, static_cast<double *>(val_out));
        break;
      case SBLAS_VAL_F32:
        hipLaunchKernelGGL(k_gather_csr<float>, grd, blk, 0, s, nnz, perm_out, cols,
                           static_cast<const float *>(val), col, static_cast<float *>(val_out));
        break;
      default:
        hipLaunchKernelGGL(k_gather_csr<uint16_t>, grd, blk, 0, s, nnz, perm_out, cols,
                           static_cast<const uint16_t *>(val), col,
                           static_cast<uint16_t *>(val_out));
        break;
    }
    if ((e = hipGetLastError()) != hipSuccess) fail("launch", e);
    if (st == SBLAS_OK && (e = hipStreamSynchronize(s)) != hipSuccess) fail("sync", e);
  }
  if (temp) (void)hipFree(temp);
  if (cols) (void)hipFree(cols);
  if (keys_out) (void)hipFree(keys_out);
  if (perm_in) (void)hipFree(perm_in);
  if (perm_out) (void)hipFree(perm_out);
  return st;
}

}  // namespace sblas
