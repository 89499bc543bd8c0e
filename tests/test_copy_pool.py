"""The host copy pool behind the pinned staging of host arrays
(hpda_sblas_pnnl_amd/csrc/copy_pool.hpp), built with g++ alone: copies of
many sizes, on pools of 1 to 16 threads, reused for thousands of copies in a
row, equal the source byte for byte; pools start and stop cleanly.  Runs
without a GPU."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

HARNESS = r"""
#include "copy_pool.hpp"
#include <cstdio>
#include <cstdlib>
#include <vector>
int main() {
  const size_t sizes[] = {0, 1, 7, 4096, (size_t{4} << 20) - 1, size_t{4} << 20,
                          (size_t{4} << 20) + 3, (size_t{37} << 20) + 5};
  std::vector<unsigned char> src((size_t{37} << 20) + 64), dst(src.size());
  unsigned s = 12345;
  for (auto &c : src) { s = s * 1103515245u + 12345u; c = static_cast<unsigned char>(s >> 16); }
  for (unsigned threads : {1u, 2u, 5u, 16u}) {
    sblas::CopyPool pool(threads);
    if (pool.threads() != threads) return 10;
    for (int rep = 0; rep < 3; rep++)
      for (size_t n : sizes)
        for (size_t off : {size_t{0}, size_t{3}}) {
          std::fill(dst.begin(), dst.end(), 0xAB);
          pool.copy(dst.data() + off, src.data() + 2 * off, n, 4096);
          for (size_t i = 0; i < dst.size(); i++) {
            const bool inside = i >= off && i < off + n;
            const unsigned char want = inside ? src[i - off + 2 * off] : 0xAB;
            if (dst[i] != want) { std::printf("mismatch t=%u n=%zu off=%zu i=%zu\n", threads, n, off, i); return 1; }
          }
        }
    // many small parallel copies in a row: every worker sees every copy once
    std::vector<unsigned char> a(1 << 16), b(1 << 16);
    for (int k = 0; k < 5000; k++) {
      a[k % a.size()] = static_cast<unsigned char>(k);
      pool.copy(b.data(), a.data(), a.size(), 1);
      if (b != a) { std::printf("mismatch in run k=%d t=%u\n", k, threads); return 2; }
    }
  }
  std::printf("ok\n");
  return 0;
}
"""


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not available")
def test_copy_pool(tmp_path):
    src = tmp_path / "pool.cpp"
    exe = tmp_path / "pool"
    src.write_text(HARNESS)
    inc = os.path.join(ROOT, "hpda_sblas_pnnl_amd", "csrc")
    r = subprocess.run(["g++", "-std=c++17", "-O2", "-pthread", "-I", inc, str(src), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.strip() == "ok", (r.returncode, r.stdout[-500:])
