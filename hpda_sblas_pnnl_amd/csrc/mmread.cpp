// mmread.cpp -- Matrix Market coordinate files into COO arrays, host side.
//
// Follows the reference's reader: the banner and size line of include/mmio.h
// (mm_read_banner :254-330, mm_read_mtx_crd_size :339-367) and the entry loop
// of the harness's file path (test/dspmv_test_dgx1.cu:133-199).  Banner
// tokens after "%%MatrixMarket" are case-insensitive; comment lines start
// with '%'; the size line holds M N nz; each entry is "i j [value]" with
// 1-based indices.  Fields: real, integer (read as an integer, stored as a
// double) and pattern (value 1.0).  Entries are returned in file order with
// 0-based indices; symmetric, skew-symmetric and Hermitian files return the
// stored entries as they are, not expanded, as the harness does.  Complex
// fields and dense (array) files are reported unsupported (the harness has no
// path for them).  Unlike the reference, malformed files and out-of-range
// indices are errors instead of garbage arrays.
#include <cctype>
#include <charconv>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "internal.hpp"
#include "plan.hpp"

namespace sblas {
namespace {

int fail(int status, const std::string &what) {
  set_error(status, "matrix market: " + what);
  return status;
}

bool read_file(const char *path, std::string &out) {
  FILE *f = std::fopen(path, "rb");
  if (!f) return false;
  std::vector<char> buf(1 << 20);
  size_t got;
  while ((got = std::fread(buf.data(), 1, buf.size(), f)) > 0) out.append(buf.data(), got);
  const bool ok = !std::ferror(f);
  std::fclose(f);
  return ok;
}

struct Cursor {
  const char *p, *end;
  void skip_space() {
    while (p < end && (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r')) p++;
  }
  bool next_int(long long &v) {
    skip_space();
    if (p < end && *p == '+') p++;
    const auto r = std::from_chars(p, end, v);
    if (r.ec != std::errc()) return false;
    p = r.ptr;
    return true;
  }
  bool next_double(double &v) {
    skip_space();
    if (p < end && *p == '+') p++;
    const auto r = std::from_chars(p, end, v);
    if (r.ec != std::errc()) return false;
    p = r.ptr;
    return true;
  }
  // the rest of the current line (without the line break), advancing past it
  std::string line() {
    const char *q = p;
    while (p < end && *p != '\n') p++;
    std::string s(q, p);
    if (p < end) p++;
    if (!s.empty() && s.back() == '\r') s.pop_back();
    return s;
  }
};

std::string lower(std::string s) {
  for (char &c : s) c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
  return s;
}

}  // namespace

int mm_read_coo(const char *path, int *m, int *n, int *nnz, int **row_idx, int **col_idx,
                double **val, int *symmetry) {
  if (!path || !m || !n || !nnz || !row_idx || !col_idx || !val) {
    set_error(SBLAS_ERR_ARGS, "matrix market: null argument");
    return SBLAS_ERR_ARGS;
  }
  *row_idx = nullptr;
  *col_idx = nullptr;
  *val = nullptr;
  std::string text;
  if (!read_file(path, text)) return fail(SBLAS_ERR_ARGS, std::string("cannot read ") + path);
  Cursor c{text.data(), text.data() + text.size()};
  // banner: %%MatrixMarket object format field symmetry
  char tok[5][64];
  const std::string banner = c.line();
  if (std::sscanf(banner.c_str(), "%63s %63s %63s %63s %63s", tok[0], tok[1], tok[2], tok[3],
                  tok[4]) != 5)
    return fail(SBLAS_ERR_ARGS, "premature end of the banner");
  if (std::strncmp(tok[0], "%%MatrixMarket", 14) != 0) return fail(SBLAS_ERR_ARGS, "no banner");
  const std::string object = lower(tok[1]), format = lower(tok[2]), field = lower(tok[3]),
                    sym = lower(tok[4]);
  if (object != "matrix") return fail(SBLAS_ERR_UNSUPPORTED, "object " + object);
  if (format != "coordinate") return fail(SBLAS_ERR_UNSUPPORTED, "format " + format);
  int kind;  // 0 real, 1 integer, 2 pattern
  if (field == "real") kind = 0;
  else if (field == "integer") kind = 1;
  else if (field == "pattern") kind = 2;
  else return fail(SBLAS_ERR_UNSUPPORTED, "field " + field);
  int symm;
  if (sym == "general") symm = 0;
  else if (sym == "symmetric") symm = 1;
  else if (sym == "skew-symmetric") symm = 2;
  else if (sym == "hermitian") symm = 3;
  else return fail(SBLAS_ERR_UNSUPPORTED, "symmetry " + sym);
  // comments, then the size line
  while (c.p < c.end && *c.p == '%') c.line();
  long long M, N, NZ;
  if (!c.next_int(M) || !c.next_int(N) || !c.next_int(NZ))
    return fail(SBLAS_ERR_ARGS, "premature end of the size line");
  if (M < 0 || N < 0 || NZ < 0 || M > INT_MAX || N > INT_MAX || NZ > INT_MAX)
    return fail(SBLAS_ERR_ARGS, "sizes out of range");
  const size_t count = static_cast<size_t>(NZ);
  int *ri = static_cast<int *>(std::malloc(sizeof(int) * (count ? count : 1)));
  int *ci = static_cast<int *>(std::malloc(sizeof(int) * (count ? count : 1)));
  double *v = static_cast<double *>(std::malloc(sizeof(double) * (count ? count : 1)));
  auto release = [&] {
    std::free(ri);
    std::free(ci);
    std::free(v);
  };
  if (!ri || !ci || !v) {
    release();
    return fail(SBLAS_ERR_ALLOC, "host allocation failed");
  }
  for (size_t k = 0; k < count; k++) {
    long long i, j;
    if (!c.next_int(i) || !c.next_int(j)) {
      release();
      return fail(SBLAS_ERR_ARGS, "premature end at entry " + std::to_string(k));
    }
    if (i < 1 || i > M || j < 1 || j > N) {
      release();
      return fail(SBLAS_ERR_ARGS, "index out of range at entry " + std::to_string(k));
    }
    double x = 1.0;
    if (kind == 0) {
      if (!c.next_double(x)) {
        release();
        return fail(SBLAS_ERR_ARGS, "bad value at entry " + std::to_string(k));
      }
    } else if (kind == 1) {
      long long q;
      if (!c.next_int(q) || q < INT_MIN || q > INT_MAX) {
        release();
        return fail(SBLAS_ERR_ARGS, "bad value at entry " + std::to_string(k));
      }
      x = static_cast<double>(q);
    }
    ri[k] = static_cast<int>(i - 1);
    ci[k] = static_cast<int>(j - 1);
    v[k] = x;
  }
  *m = static_cast<int>(M);
  *n = static_cast<int>(N);
  *nnz = static_cast<int>(NZ);
  *row_idx = ri;
  *col_idx = ci;
  *val = v;
  if (symmetry) *symmetry = symm;
  return SBLAS_OK;
}

}  // namespace sblas

extern "C" {

int sblas_mm_read_coo(const char *path, int *m, int *n, int *nnz, int **row_idx, int **col_idx,
                      double **val, int *symmetry) {
  sblas::clear_error();
  return sblas::mm_read_coo(path, m, n, nnz, row_idx, col_idx, val, symmetry);
}

void sblas_free(void *p) { std::free(p); }

}  // extern "C"
