// ingest.cpp -- C entry points of matrix ingestion: the reference's host
// helpers sortCOORow / sortCOOCol / coo2csr / coo2csc (src/spmv_helper.cu:
// 277-642), their device-array forms, and spMspV's sparse-x filter on device
// arrays.
//
// Like the reference's, the host helpers upload the caller's arrays, work on
// the device and copy the result back, synchronously.  sortCOORow / sortCOOCol
// sort the caller's arrays in place; coo2csr / coo2csc leave the COO input
// untouched (the reference sorts device copies) and fill the caller's output
// arrays.  The reference returns nothing and exits on CUDA errors; these
// return nothing too and leave a status in sblas_last_error() (SBLAS_OK after
// success), with the outputs untouched on error.  Indices are validated on
// the device first (the reference passes bad indices to cuSPARSE unchecked).
#include <cstdlib>

#include "host_util.hpp"
#include "internal.hpp"
#include "plan.hpp"

namespace sblas {
namespace {

int device_from_env() {
  const char *e = std::getenv("SBLAS_DEVICE");
  return (e && *e) ? std::atoi(e) : 0;
}

bool host_args_ok(int m, int n, int nnz, const void *a, const void *b, const void *c) {
  if (m < 0 || n < 0 || nnz < 0 || (nnz > 0 && (!a || !b || !c))) {
    set_error(SBLAS_ERR_ARGS, "bad sizes or null arrays");
    return false;
  }
  return true;
}

#define SBLAS_COPY(dst, src, bytes, kind)                                      \
  do {                                                                         \
    if ((bytes) > 0) {                                                         \
      hipError_t e_ = hipMemcpyAsync(dst, src, bytes, kind, st.s);             \
      if (e_ == hipSuccess) e_ = hipStreamSynchronize(st.s);                   \
      if (e_ != hipSuccess) {                                                  \
        set_error(SBLAS_ERR_HIP, std::string("copy: ") + hipGetErrorString(e_)); \
        return SBLAS_ERR_HIP;                                                  \
      }                                                                        \
    }                                                                          \
  } while (0)

// Sorts host COO arrays in place through the device.
int host_sort(int m, int n, int nnz, double *val, int *row, int *col, bool by_col) {
  if (!host_args_ok(m, n, nnz, val, row, col)) return SBLAS_ERR_ARGS;
  if (nnz == 0) return SBLAS_OK;
  DeviceGuard guard(device_from_env());
  if (!guard.ok) {
    set_error(SBLAS_ERR_NO_DEVICE, "no HIP device");
    return SBLAS_ERR_NO_DEVICE;
  }
  OwnedStream st;
  if (!st.s) {
    set_error(SBLAS_ERR_HIP, "stream creation failed");
    return SBLAS_ERR_HIP;
  }
  DeviceArrays mem;
  double *d_val = mem.alloc<double>(nnz);
  int *d_row = mem.alloc<int>(nnz), *d_col = mem.alloc<int>(nnz);
  if (!d_val || !d_row || !d_col) {
    set_error(SBLAS_ERR_ALLOC, "device allocation failed");
    return SBLAS_ERR_ALLOC;
  }
  const size_t nv = sizeof(double) * nnz, ni = sizeof(int) * nnz;
  SBLAS_COPY(d_val, val, nv, hipMemcpyHostToDevice);
  SBLAS_COPY(d_row, row, ni, hipMemcpyHostToDevice);
  SBLAS_COPY(d_col, col, ni, hipMemcpyHostToDevice);
  const int s = coo_sort_device(m, n, nnz, d_row, d_col, d_val, SBLAS_VAL_F64, by_col, st.s);
  if (s != SBLAS_OK) return s;
  SBLAS_COPY(val, d_val, nv, hipMemcpyDeviceToHost);
  SBLAS_COPY(row, d_row, ni, hipMemcpyDeviceToHost);
  SBLAS_COPY(col, d_col, ni, hipMemcpyDeviceToHost);
  return SBLAS_OK;
}

// COO -> CSR (to_csc false: out_ptr[m+1], out_idx = columns) or CSC (out_ptr
// [n+1], out_idx = rows), values in the same order.
int host_convert(int m, int n, int nnz, const double *val, const int *row, const int *col,
                 double *out_val, int *out_ptr, int *out_idx, bool to_csc) {
  if (!host_args_ok(m, n, nnz, val, row, col)) return SBLAS_ERR_ARGS;
  if (!out_ptr || (nnz > 0 && (!out_val || !out_idx))) {
    set_error(SBLAS_ERR_ARGS, "null output arrays");
    return SBLAS_ERR_ARGS;
  }
  DeviceGuard guard(device_from_env());
  if (!guard.ok) {
    set_error(SBLAS_ERR_NO_DEVICE, "no HIP device");
    return SBLAS_ERR_NO_DEVICE;
  }
  OwnedStream st;
  if (!st.s) {
    set_error(SBLAS_ERR_HIP, "stream creation failed");
    return SBLAS_ERR_HIP;
  }
  const int major = to_csc ? n : m;
  DeviceArrays mem;
  double *d_val = mem.alloc<double>(nnz);
  int *d_row = mem.alloc<int>(nnz), *d_col = mem.alloc<int>(nnz);
  int *d_ptr = mem.alloc<int>(static_cast<size_t>(major) + 1);
  if (!d_val || !d_row || !d_col || !d_ptr) {
    set_error(SBLAS_ERR_ALLOC, "device allocation failed");
    return SBLAS_ERR_ALLOC;
  }
  const size_t nv = sizeof(double) * nnz, ni = sizeof(int) * nnz;
  SBLAS_COPY(d_val, val, nv, hipMemcpyHostToDevice);
  SBLAS_COPY(d_row, row, ni, hipMemcpyHostToDevice);
  SBLAS_COPY(d_col, col, ni, hipMemcpyHostToDevice);
  int s = coo_sort_device(m, n, nnz, d_row, d_col, d_val, SBLAS_VAL_F64, to_csc, st.s);
  if (s != SBLAS_OK) return s;
  const int *d_major = to_csc ? d_col : d_row, *d_minor = to_csc ? d_row : d_col;
  s = ptr_from_sorted_device(major, nnz, d_major, d_ptr, st.s);
  if (s != SBLAS_OK) return s;
  SBLAS_COPY(out_val, d_val, nv, hipMemcpyDeviceToHost);
  SBLAS_COPY(out_idx, d_minor, ni, hipMemcpyDeviceToHost);
  SBLAS_COPY(out_ptr, d_ptr, sizeof(int) * (static_cast<size_t>(major) + 1),
             hipMemcpyDeviceToHost);
  return SBLAS_OK;
}

#undef SBLAS_COPY

}  // namespace
}  // namespace sblas

extern "C" {

int sblas_coo_sort(int m, int n, int64_t nnz, int *d_row_idx, int *d_col_idx, void *d_val,
                   int val_type, int by_col, void *stream) {
  sblas::clear_error();
  return sblas::coo_sort_device(m, n, nnz, d_row_idx, d_col_idx, d_val, val_type, by_col != 0,
                                static_cast<hipStream_t>(stream));
}

int sblas_coo_ptr(int count, int64_t nnz, const int *d_sorted_idx, int *d_ptr, void *stream) {
  sblas::clear_error();
  return sblas::ptr_from_sorted_device(count, nnz, d_sorted_idx, d_ptr,
                                       static_cast<hipStream_t>(stream));
}

int sblas_csr_drop_zero_x(int m, int n, int64_t nnz, const int *d_row_ptr, const int *d_col,
                          const double *d_val, const double *d_x, int *d_row_ptr_out,
                          int *d_col_out, double *d_val_out, int64_t *nnz_out, void *stream) {
  sblas::clear_error();
  return sblas::csr_drop_zero_x_device(m, n, nnz, d_row_ptr, d_col, d_val, d_x, d_row_ptr_out,
                                       d_col_out, d_val_out, nnz_out,
                                       static_cast<hipStream_t>(stream));
}

void sortCOORow(int m, int n, int nnz, double *cooVal, int *cooRowIdx, int *cooColIdx) {
  sblas::clear_error();
  (void)sblas::host_sort(m, n, nnz, cooVal, cooRowIdx, cooColIdx, false);
}

void sortCOOCol(int m, int n, int nnz, double *cooVal, int *cooRowIdx, int *cooColIdx) {
  sblas::clear_error();
  (void)sblas::host_sort(m, n, nnz, cooVal, cooRowIdx, cooColIdx, true);
}

void coo2csr(int m, int n, int nnz, double *cooVal, int *cooRowIdx, int *cooColIdx,
             double *csrVal, int *csrRowPtr, int *csrColIdx) {
  sblas::clear_error();
  (void)sblas::host_convert(m, n, nnz, cooVal, cooRowIdx, cooColIdx, csrVal, csrRowPtr,
                            csrColIdx, false);
}

void coo2csc(int m, int n, int nnz, double *cooVal, int *cooRowIdx, int *cooColIdx,
             double *cscVal, int *cscColPtr, int *cscRowIdx) {
  sblas::clear_error();
  (void)sblas::host_convert(m, n, nnz, cooVal, cooRowIdx, cooColIdx, cscVal, cscColPtr,
                            cscRowIdx, true);
}

}  // extern "C"
