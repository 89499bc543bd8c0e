/*
 * spmv_kernel.h -- C++ compatibility header for code written against the
 * reference's include/spmv_kernel.h (the test harnesses
 * test/dspmv_test_smt.cu and test/dspmv_test_dgx1.cu).  It declares the same
 * entry points (from sblas.h, extern "C") and gives spmv_ret the helper
 * methods the reference defines at include/spmv_kernel.h:22-56 (init, add,
 * avg, print, to_string), so a recompiled harness links against libsblas.so
 * unchanged.  The methods are non-virtual: the struct keeps the 40-byte
 * layout that the C entry points return by value.
 */
#ifndef SBLAS_SPMV_KERNEL_COMPAT_H
#define SBLAS_SPMV_KERNEL_COMPAT_H

#ifdef __cplusplus
#include <cstdio>
#include <sstream>
#include <string>

#define SBLAS_SPMV_RET_DEFINED
struct spmv_ret {
  double numa_part_time;
  double part_time;
  double comp_time;
  double comm_time;
  double merg_time;

  void init() { numa_part_time = part_time = comp_time = comm_time = merg_time = 0.0; }
  // field-wise sum, as the harness accumulates repeated calls
  void add(spmv_ret o) {
    numa_part_time += o.numa_part_time;
    part_time += o.part_time;
    comp_time += o.comp_time;
    comm_time += o.comm_time;
    merg_time += o.merg_time;
  }
  void avg(int n) {
    const double d = static_cast<double>(n);
    numa_part_time /= d;
    part_time /= d;
    comp_time /= d;
    comm_time /= d;
    merg_time /= d;
  }
  // the reference's printed order: NUMA partition, partition, communication,
  // computation, merge
  void print() {
    std::printf("%f, %f, %f, %f, %f\n", numa_part_time, part_time, comm_time, comp_time,
                merg_time);
  }
  std::string to_string() {
    std::ostringstream out;
    out << numa_part_time << "," << part_time << "," << comm_time << "," << comp_time << ","
        << merg_time << "\n";
    return out.str();
  }
};
typedef struct spmv_ret spmv_ret;
#endif /* __cplusplus */

#include "sblas.h"

#endif /* SBLAS_SPMV_KERNEL_COMPAT_H */
