"""The C++ compatibility header include/spmv_kernel.h: a harness written
against the reference's header (spmv_ret with init/add/avg/print/to_string,
the drop-in entry points) compiles with g++ and links against libsblas.so;
the struct keeps its 40-byte layout.  Runs without a GPU (the call below is
refused before any device work)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

HARNESS = r"""
#include "spmv_kernel.h"
#include <cmath>
#include <cstddef>
#include <type_traits>
static_assert(sizeof(spmv_ret) == 40, "layout");
static_assert(std::is_standard_layout<spmv_ret>::value, "layout");
static_assert(offsetof(spmv_ret, merg_time) == 32, "layout");
int main() {
  spmv_ret acc;
  acc.init();
  spmv_ret a{1.0, 2.0, 3.0, 4.0, 5.0};
  acc.add(a);
  acc.add(a);
  acc.avg(2);
  acc.print();
  std::printf("%s", acc.to_string().c_str());
  // ngpu = 2 is refused with NaN timings and a status, before any device work
  int rp[2] = {0, 1}, col[1] = {0};
  double val[1] = {1.0}, x[1] = {1.0}, y[1] = {0.0}, alpha = 1.0, beta = 0.0;
  spmv_ret r = spMV_mgpu_v1(1, 1, 1, &alpha, val, rp, col, x, &beta, y, 2, 1);
  std::printf("%d %d\n", std::isnan(r.comp_time) ? 1 : 0, sblas_last_error());
  return 0;
}
"""


def test_harness_compiles_and_links(tmp_path):
    from hpda_sblas_pnnl_amd import _lib
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    src = tmp_path / "harness.cpp"
    src.write_text(HARNESS)
    exe = tmp_path / "harness"
    libdir = os.path.dirname(_lib.LIB_PATH)
    cmd = ["g++", "-std=c++17", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe),
           "-L", libdir, "-lsblas", f"-Wl,-rpath,{libdir}", "-L/opt/rocm/lib",
           "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.splitlines()
    assert lines[0] == "1.000000, 2.000000, 4.000000, 3.000000, 5.000000"  # comm before comp
    assert lines[1] == "1,2,4,3,5"
    assert lines[2] == "1 2"  # NaN timings, SBLAS_ERR_NGPU
