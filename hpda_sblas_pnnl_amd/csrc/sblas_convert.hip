// sblas_convert.hip -- matrix ingestion on the device: COO sorts and the
// compressed pointer arrays of COO -> CSR / CSC (src/spmv_helper.cu:277-642).
//
// The reference sorts with cusparseXcoosortByRow / ByColumn plus a gather of
// the values through the returned permutation (src/spmv_helper.cu:277-331,
// 333-387, 406-476) and builds the pointer array with cusparseXcoo2csr
// (:478-491).  Here one stable LSD radix sort (hipCUB) orders packed 64-bit
// keys -- the row in the high bits and the column in the low bits for a row
// sort, the other way round for a column sort -- with the values as the sort's
// payload.  Elements come out by row, then column (or by column, then row);
// elements with equal (row, column) keep their input order.  The radix sort
// covers only the bits the two indices need.  Pointers are the first position
// of every row in the sorted index array, one binary search per row.  All of
// it is integer work and data movement: values are copied, never combined.
//
// Also here: the sparse-x filter of spMspV (src/dspmv_mgpu_v1.cu:19-74,
// src/dspmv_mgpu_v2.cu:35-101), which the reference runs as a host loop
// before uploading: elements whose x[col] is zero are dropped, row by row in
// order.  On the device: a keep flag per element, an exclusive scan of the
// flags, a scatter of the kept elements, and new row pointers read off the
// scan at the old ones.
#include <hipcub/hipcub.hpp>

#include <climits>

#include "internal.hpp"

namespace sblas {
namespace {

// bits that hold every value in [0, count)
int index_bits(int count) {
  int b = 1;
  while (b < 31 && (1 << b) < count) b++;
  return b;
}

__global__ void k_pack_keys(int nnz, const int *__restrict__ hi, const int *__restrict__ lo,
                            int lo_bits, uint64_t *__restrict__ keys) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nnz) return;
  keys[k] = (static_cast<uint64_t>(static_cast<uint32_t>(hi[k])) << lo_bits) |
            static_cast<uint32_t>(lo[k]);
}

__global__ void k_unpack_keys(int nnz, const uint64_t *__restrict__ keys, int lo_bits,
                              int *__restrict__ hi, int *__restrict__ lo) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nnz) return;
  const uint64_t key = keys[k];
  hi[k] = static_cast<int>(key >> lo_bits);
  lo[k] = static_cast<int>(key & ((uint64_t{1} << lo_bits) - 1));
}

// ptr[r] = first position k with idx[k] >= r, for r in [0, count]
__global__ void k_ptr_from_sorted(int count, int nnz, const int *__restrict__ idx,
                                  int *__restrict__ ptr) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r > count) return;
  int lo = 0, hi = nnz;
  while (lo < hi) {
    const int mid = lo + ((hi - lo) >> 1);
    if (idx[mid] < r) lo = mid + 1; else hi = mid;
  }
  ptr[r] = lo;
}

__global__ void k_fill_zero_i32(int *__restrict__ p) { *p = 0; }

template <typename VT>
int sort_pairs(int nnz, int end_bit, uint64_t *keys_in, uint64_t *keys_out, VT *val,
               VT *val_out, hipStream_t s) {
  size_t temp_bytes = 0;
  hipError_t e = hipcub::DeviceRadixSort::SortPairs(nullptr, temp_bytes, keys_in, keys_out, val,
                                                    val_out, nnz, 0, end_bit, s);
  if (e != hipSuccess) {
    set_error(SBLAS_ERR_HIP, std::string("coo sort: ") + hipGetErrorString(e));
    return SBLAS_ERR_HIP;
  }
  void *temp = nullptr;
  if (hipMalloc(&temp, temp_bytes ? temp_bytes : 1) != hipSuccess) {
    set_error(SBLAS_ERR_ALLOC, "coo sort: scratch allocation failed");
    return SBLAS_ERR_ALLOC;
  }
  e = hipcub::DeviceRadixSort::SortPairs(temp, temp_bytes, keys_in, keys_out, val, val_out, nnz,
                                         0, end_bit, s);
  if (e == hipSuccess) e = hipMemcpyAsync(val, val_out, sizeof(VT) * static_cast<size_t>(nnz),
                                          hipMemcpyDeviceToDevice, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(temp);
  if (e != hipSuccess) {
    set_error(SBLAS_ERR_HIP, std::string("coo sort: ") + hipGetErrorString(e));
    return SBLAS_ERR_HIP;
  }
  return SBLAS_OK;
}

// element kept by the sparse-x filter: x at its column is not zero
__global__ void k_keep_flags(int nnz, const int *__restrict__ col, const double *__restrict__ x,
                             int *__restrict__ keep) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nnz) return;
  keep[j] = x[col[j]] != 0.0 ? 1 : 0;
}

__global__ void k_scatter_kept(int nnz, const int *__restrict__ keep, const int *__restrict__ pos,
                               const int *__restrict__ col, const double *__restrict__ val,
                               int *__restrict__ col_out, double *__restrict__ val_out) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= nnz || !keep[j]) return;
  col_out[pos[j]] = col[j];
  val_out[pos[j]] = val[j];
}

// rp_out[r] = kept elements before rp[r]; pos has nnz + 1 entries (the last
// one is the kept total)
__global__ void k_rebase_rows(int m, const int *__restrict__ rp, const int *__restrict__ pos,
                              int *__restrict__ rp_out) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r > m) return;
  rp_out[r] = pos[rp[r]];
}

// column indices outside [0, n) become 0 and raise *bad (never cleared here)
__global__ void k_sanitize_cols(int count, int *__restrict__ col, int n, int *__restrict__ bad) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= count) return;
  const int c = col[j];
  if (c < 0 || c >= n) {
    col[j] = 0;
    atomicOr(bad, 1);
  }
}

}  // namespace

int launch_sanitize_cols(int count, int *col, int n, int *bad, hipStream_t s) {
  if (count <= 0) return SBLAS_OK;
  hipLaunchKernelGGL(k_sanitize_cols, dim3((count + 255) / 256), dim3(256), 0, s, count, col, n,
                     bad);
  SBLAS_HIP_TRY(hipGetLastError());
  return SBLAS_OK;
}

int csr_drop_zero_x_device(int m, int n, int64_t nnz, const int *rp, const int *col,
                           const double *val, const double *x, int *rp_out, int *col_out,
                           double *val_out, int64_t *nnz_out, hipStream_t s) {
  if (m < 0 || n < 0 || nnz < 0 || nnz > INT_MAX - 1 || !rp || !rp_out || !nnz_out ||
      (nnz > 0 && (!col || !val || !x || !col_out || !val_out))) {
    set_error(SBLAS_ERR_ARGS, "sparse-x filter: bad sizes or pointers");
    return SBLAS_ERR_ARGS;
  }
  bool ok = false;
  int st = check_indices(rp, static_cast<int64_t>(m) + 1, 0, static_cast<int>(nnz) + 1, true, s,
                         &ok);
  if (st != SBLAS_OK) return st;
  if (ok && nnz > 0 && (st = check_indices(col, nnz, 0, n, false, s, &ok)) != SBLAS_OK) return st;
  if (!ok) {
    set_error(SBLAS_ERR_ARGS, "sparse-x filter: row pointers or column indices out of range");
    return SBLAS_ERR_ARGS;
  }
  const int count = static_cast<int>(nnz);
  int *keep = nullptr, *pos = nullptr;
  void *temp = nullptr;
  size_t temp_bytes = 0;
  auto release = [&] {
    if (keep) (void)hipFree(keep);
    if (pos) (void)hipFree(pos);
    if (temp) (void)hipFree(temp);
  };
  // keep[nnz] = 0 so the scan's last entry is the kept total
  if (hipMalloc(&keep, sizeof(int) * (count + 1)) != hipSuccess ||
      hipMalloc(&pos, sizeof(int) * (count + 1)) != hipSuccess) {
    release();
    set_error(SBLAS_ERR_ALLOC, "sparse-x filter: allocation failed");
    return SBLAS_ERR_ALLOC;
  }
  hipError_t e = hipcub::DeviceScan::ExclusiveSum(nullptr, temp_bytes, keep, pos, count + 1, s);
  if (e == hipSuccess && hipMalloc(&temp, temp_bytes ? temp_bytes : 1) != hipSuccess) {
    release();
    set_error(SBLAS_ERR_ALLOC, "sparse-x filter: allocation failed");
    return SBLAS_ERR_ALLOC;
  }
  const dim3 blk(256), grd((count + 256) / 256);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(k_fill_zero_i32, dim3(1), dim3(1), 0, s, keep + count);
    if (count > 0) hipLaunchKernelGGL(k_keep_flags, grd, blk, 0, s, count, col, x, keep);
    e = hipcub::DeviceScan::ExclusiveSum(temp, temp_bytes, keep, pos, count + 1, s);
  }
  if (e == hipSuccess) {
    if (count > 0)
      hipLaunchKernelGGL(k_scatter_kept, grd, blk, 0, s, count, keep, pos, col, val, col_out,
                         val_out);
    hipLaunchKernelGGL(k_rebase_rows, dim3((m + 1 + 255) / 256), blk, 0, s, m, rp, pos, rp_out);
    e = hipGetLastError();
  }
  int total = 0;
  if (e == hipSuccess) e = hipMemcpyAsync(&total, pos + count, sizeof(int), hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  release();
  if (e != hipSuccess) {
    set_error(SBLAS_ERR_HIP, std::string("sparse-x filter: ") + hipGetErrorString(e));
    return SBLAS_ERR_HIP;
  }
  *nnz_out = total;
  return SBLAS_OK;
}

int coo_sort_device(int m, int n, int64_t nnz, int *row, int *col, void *val, int val_type,
                    bool by_col, hipStream_t s) {
  if (m < 0 || n < 0 || nnz < 0 || nnz > INT_MAX || (nnz > 0 && (!row || !col || !val)) ||
      val_type < SBLAS_VAL_F64 || val_type > SBLAS_VAL_BF16) {
    set_error(SBLAS_ERR_ARGS, "coo sort: bad sizes, pointers or value type");
    return SBLAS_ERR_ARGS;
  }
  if (nnz == 0) return SBLAS_OK;
  bool ok = false;
  int st;
  if ((st = check_indices(row, nnz, 0, m, false, s, &ok)) != SBLAS_OK) return st;
  if (ok && (st = check_indices(col, nnz, 0, n, false, s, &ok)) != SBLAS_OK) return st;
  if (!ok) {
    set_error(SBLAS_ERR_ARGS, "coo sort: a row or column index is out of range");
    return SBLAS_ERR_ARGS;
  }
  const int count = static_cast<int>(nnz);
  int *hi = by_col ? col : row, *lo = by_col ? row : col;
  const int lo_bits = index_bits(by_col ? m : n);
  const int end_bit = lo_bits + index_bits(by_col ? n : m);
  const size_t vbytes = val_type == SBLAS_VAL_F64 ? 8 : val_type == SBLAS_VAL_F32 ? 4 : 2;
  uint64_t *keys = nullptr, *keys_out = nullptr;
  void *val_out = nullptr;
  auto release = [&] {
    if (keys) (void)hipFree(keys);
    if (keys_out) (void)hipFree(keys_out);
    if (val_out) (void)hipFree(val_out);
  };
  if (hipMalloc(&keys, sizeof(uint64_t) * count) != hipSuccess ||
      hipMalloc(&keys_out, sizeof(uint64_t) * count) != hipSuccess ||
      hipMalloc(&val_out, vbytes * count) != hipSuccess) {
    release();
    set_error(SBLAS_ERR_ALLOC, "coo sort: allocation failed");
    return SBLAS_ERR_ALLOC;
  }
  const dim3 blk(256), grd((count + 255) / 256);
  hipLaunchKernelGGL(k_pack_keys, grd, blk, 0, s, count, hi, lo, lo_bits, keys);
  switch (val_type) {
    case SBLAS_VAL_F64:
      st = sort_pairs(count, end_bit, keys, keys_out, static_cast<double *>(val),
                      static_cast<double *>(val_out), s);
      break;
    case SBLAS_VAL_F32:
      st = sort_pairs(count, end_bit, keys, keys_out, static_cast<float *>(val),
                      static_cast<float *>(val_out), s);
      break;
    default:
      st = sort_pairs(count, end_bit, keys, keys_out, static_cast<uint16_t *>(val),
                      static_cast<uint16_t *>(val_out), s);
      break;
  }
  if (st == SBLAS_OK) {
    hipLaunchKernelGGL(k_unpack_keys, grd, blk, 0, s, count, keys_out, lo_bits, hi, lo);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
      set_error(SBLAS_ERR_HIP, std::string("coo sort: ") + hipGetErrorString(e));
      st = SBLAS_ERR_HIP;
    }
  }
  release();
  return st;
}

int ptr_from_sorted_device(int count, int64_t nnz, const int *idx, int *ptr, hipStream_t s) {
  if (count < 0 || nnz < 0 || nnz > INT_MAX || !ptr || (nnz > 0 && !idx)) {
    set_error(SBLAS_ERR_ARGS, "pointer array: bad sizes or pointers");
    return SBLAS_ERR_ARGS;
  }
  if (nnz > 0) {
    bool ok = false;
    const int st = check_indices(idx, nnz, 0, count, true, s, &ok);
    if (st != SBLAS_OK) return st;
    if (!ok) {
      set_error(SBLAS_ERR_ARGS, "pointer array: indices unsorted or out of range");
      return SBLAS_ERR_ARGS;
    }
  }
  hipLaunchKernelGGL(k_ptr_from_sorted, dim3((count + 1 + 255) / 256), dim3(256), 0, s, count,
                     static_cast<int>(nnz), idx, ptr);
  SBLAS_HIP_TRY(hipGetLastError());
  SBLAS_HIP_TRY(hipStreamSynchronize(s));
  return SBLAS_OK;
}

}  // namespace sblas
