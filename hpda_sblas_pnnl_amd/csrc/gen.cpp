// gen.cpp -- seeded synthetic matrices and vectors for the engine's tests and
// benchmarks (libsblas_gen.so).
//
// Every random draw comes from a counter-based generator: the state of the
// stream for (seed, tag, index) is a splitmix64 hash of those three values, so
// a row's (or an edge's) numbers do not depend on the thread that makes them
// and results are identical for any OpenMP thread count.  Samplers are written
// here (no std:: distributions, whose outputs vary between libraries):
//   uniform [0,1): top 53 bits of a 64-bit draw;
//   normal: Box-Muller on two uniforms;
//   discrete power law p(L) ~ (L + shift)^-expo on [lmin, lmax]: inverse CDF
//     over a table;
//   k distinct sorted integers of a window: draw, sort, unique, refill; or
//     selection sampling when k is a large share of the window;
//   R-MAT: one 32-bit uniform per level, quadrants (a, b, c, d) in that order.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace {

constexpr uint64_t kGolden = 0x9e3779b97f4a7c15ULL;

inline uint64_t fmix(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

struct Rng {
  uint64_t s;
  Rng(uint64_t seed, uint64_t tag, uint64_t index)
      : s(fmix(fmix(seed * kGolden + tag) ^ (index * 0xd1b54a32d192ed03ULL + kGolden))) {}
  uint64_t next() {
    s += kGolden;
    return fmix(s);
  }
  double uniform() { return static_cast<double>(next() >> 11) * 0x1.0p-53; }
  double normal() {
    double u1 = uniform();
    const double u2 = uniform();
    if (u1 <= 0.0) u1 = 0x1.0p-53;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
  }
  // uniform integer in [0, n) by multiply-shift on 64-bit draws
  uint64_t below(uint64_t n) {
    return static_cast<uint64_t>((static_cast<unsigned __int128>(next()) * n) >> 64);
  }
};

enum : uint64_t {
  kTagRowLen = 0x11,
  kTagCols = 0x22,
  kTagVals = 0x33,
  kTagRmat = 0x44,
};

// k distinct sorted values of [lo, lo + w), appended to out.
void sample_distinct(Rng &g, int64_t lo, int64_t w, int k, int *out) {
  if (k <= 0) return;
  if (k >= w) {
    for (int i = 0; i < k; i++) out[i] = static_cast<int>(lo + i);
    return;
  }
  if (static_cast<int64_t>(k) * 4 >= w) {
    // selection sampling over the window
    int chosen = 0;
    for (int64_t t = 0; t < w && chosen < k; t++) {
      const double u = g.uniform();
      if (static_cast<double>(w - t) * u < static_cast<double>(k - chosen))
        out[chosen++] = static_cast<int>(lo + t);
    }
    return;
  }
  std::vector<int> v;
  v.reserve(static_cast<size_t>(k) * 2);
  while (static_cast<int>(v.size()) < k) {
    const int need = k - static_cast<int>(v.size());
    for (int i = 0; i < need; i++) v.push_back(static_cast<int>(lo + g.below(w)));
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
  }
  std::copy(v.begin(), v.end(), out);
}

}  // namespace

extern "C" {

int64_t sblas_gen_rmat_n(int scale, int64_t edges, double a, double b, double c, uint64_t seed,
                         int64_t n, int **rp_out, int **col_out);

// Row lengths from p(L) ~ (L + shift)^-expo on [lmin, lmax].
int sblas_gen_rowlen_powerlaw(int m, uint64_t seed, int lmin, int lmax, double shift,
                              double expo, int *rowlen) {
  if (m < 0 || lmin < 0 || lmax < lmin || !rowlen) return 1;
  std::vector<double> cdf(static_cast<size_t>(lmax - lmin + 1));
  double acc = 0;
  for (int L = lmin; L <= lmax; L++) {
    acc += std::pow(static_cast<double>(L) + shift, -expo);
    cdf[L - lmin] = acc;
  }
  for (double &c : cdf) c /= acc;
#pragma omp parallel for schedule(static)
  for (int r = 0; r < m; r++) {
    Rng g(seed, kTagRowLen, static_cast<uint64_t>(r));
    const double u = g.uniform();
    const size_t k = std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin();
    rowlen[r] = lmin + static_cast<int>(std::min(k, cdf.size() - 1));
  }
  return 0;
}

// Row lengths round(N(mu, sigma)) clipped to [lmin, lmax].
int sblas_gen_rowlen_normal(int m, uint64_t seed, double mu, double sigma, int lmin, int lmax,
                            int *rowlen) {
  if (m < 0 || lmax < lmin || !rowlen) return 1;
#pragma omp parallel for schedule(static)
  for (int r = 0; r < m; r++) {
    Rng g(seed, kTagRowLen, static_cast<uint64_t>(r));
    const double L = std::nearbyint(mu + sigma * g.normal());
    rowlen[r] = static_cast<int>(std::min<double>(lmax, std::max<double>(lmin, L)));
  }
  return 0;
}

// Distinct sorted columns uniform in [0, n) (halfwidth < 0) or in
// [max(0, r - halfwidth), min(n - 1, r + halfwidth)]; rp gives the row
// lengths, which must not exceed the window.
int sblas_gen_cols(int m, int n, int halfwidth, uint64_t seed, const int *rp, int *col) {
  if (m < 0 || n < 1 || !rp || (!col && rp[m] > 0)) return 1;
  int bad = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : bad)
  for (int r = 0; r < m; r++) {
    int64_t lo = 0, hi = n - 1;
    if (halfwidth >= 0) {
      lo = std::max<int64_t>(0, static_cast<int64_t>(r) - halfwidth);
      hi = std::min<int64_t>(n - 1, static_cast<int64_t>(r) + halfwidth);
    }
    const int k = rp[r + 1] - rp[r];
    if (k > hi - lo + 1) {
      bad++;
      continue;
    }
    Rng g(seed, kTagCols, static_cast<uint64_t>(r));
    sample_distinct(g, lo, hi - lo + 1, k, col + rp[r]);
  }
  return bad ? 2 : 0;
}

// count values: kind 0 uniform(-1, 1), 1 normal(0, 1), 2 all ones.
// Value i uses stream (stream_tag, i / 4096) so blocks fill in parallel.
int sblas_gen_values(int64_t count, uint64_t seed, uint64_t stream_tag, int kind, double *out) {
  if (count < 0 || (!out && count > 0)) return 1;
  const int64_t block = 4096;
  const int64_t nblk = (count + block - 1) / block;
#pragma omp parallel for schedule(static)
  for (int64_t b = 0; b < nblk; b++) {
    Rng g(seed, kTagVals ^ (stream_tag << 8), static_cast<uint64_t>(b));
    const int64_t end = std::min(count, (b + 1) * block);
    for (int64_t i = b * block; i < end; i++) {
      if (kind == 0) out[i] = 2.0 * g.uniform() - 1.0;
      else if (kind == 1) out[i] = g.normal();
      else out[i] = 1.0;
    }
  }
  return 0;
}

// R-MAT graph with 2^scale vertices and `edges` draws (quadrant probabilities
// a, b, c, 1-a-b-c; no vertex permutation); duplicates removed, self loops
// kept.  Returns nnz and malloc'ed CSR arrays (free with sblas_gen_free), or
// -1 on failure.
int64_t sblas_gen_rmat(int scale, int64_t edges, double a, double b, double c, uint64_t seed,
                       int **rp_out, int **col_out) {
  return sblas_gen_rmat_n(scale, edges, a, b, c, seed, int64_t{1} << scale, rp_out, col_out);
}

// As sblas_gen_rmat on the first n vertices: edges with an endpoint >= n are
// dropped (an R-MAT graph at the next power of two, cut to n).
int64_t sblas_gen_rmat_n(int scale, int64_t edges, double a, double b, double c, uint64_t seed,
                         int64_t n, int **rp_out, int **col_out) {
  if (scale < 1 || scale > 30 || edges < 0 || !rp_out || !col_out || n < 1 ||
      n > (int64_t{1} << scale))
    return -1;
  const int64_t m = n;
  std::vector<int> er(static_cast<size_t>(edges)), ec(static_cast<size_t>(edges));
  const double ab = a + b, abc = a + b + c;
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < edges; e++) {
    Rng g(seed, kTagRmat, static_cast<uint64_t>(e));
    uint32_t r = 0, cc = 0;
    uint64_t z = 0;
    for (int lvl = 0; lvl < scale; lvl++) {
      if ((lvl & 1) == 0) z = g.next();
      const double u = static_cast<double>((z >> (32 * (lvl & 1))) & 0xffffffffULL) * 0x1.0p-32;
      uint32_t rb = 1, cb = 1;
      if (u < a) { rb = 0; cb = 0; }
      else if (u < ab) { rb = 0; cb = 1; }
      else if (u < abc) { rb = 1; cb = 0; }
      r = (r << 1) | rb;
      cc = (cc << 1) | cb;
    }
    const bool keep = static_cast<int64_t>(r) < m && static_cast<int64_t>(cc) < m;
    er[e] = keep ? static_cast<int>(r) : -1;
    ec[e] = static_cast<int>(cc);
  }
  std::vector<std::atomic<int64_t>> cnt(static_cast<size_t>(m));
  for (auto &x : cnt) x.store(0, std::memory_order_relaxed);
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < edges; e++)
    if (er[e] >= 0) cnt[er[e]].fetch_add(1, std::memory_order_relaxed);
  std::vector<int64_t> start(static_cast<size_t>(m) + 1, 0);
  for (int64_t r = 0; r < m; r++) start[r + 1] = start[r] + cnt[r].load(std::memory_order_relaxed);
  std::vector<int> tmp(static_cast<size_t>(edges));
  for (int64_t r = 0; r < m; r++) cnt[r].store(start[r], std::memory_order_relaxed);
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < edges; e++) {
    if (er[e] < 0) continue;
    const int64_t pos = cnt[er[e]].fetch_add(1, std::memory_order_relaxed);
    tmp[pos] = ec[e];
  }
  std::vector<int>().swap(er);
  std::vector<int>().swap(ec);
  std::vector<int64_t> uniq(static_cast<size_t>(m), 0);
#pragma omp parallel for schedule(dynamic, 1024)
  for (int64_t r = 0; r < m; r++) {
    int *b0 = tmp.data() + start[r], *b1 = tmp.data() + start[r + 1];
    std::sort(b0, b1);
    uniq[r] = std::unique(b0, b1) - b0;
  }
  int64_t nnz = 0;
  for (int64_t r = 0; r < m; r++) nnz += uniq[r];
  if (nnz > INT32_MAX) return -1;
  int *rp = static_cast<int *>(std::malloc(sizeof(int) * (static_cast<size_t>(m) + 1)));
  int *col = static_cast<int *>(std::malloc(sizeof(int) * static_cast<size_t>(nnz > 0 ? nnz : 1)));
  if (!rp || !col) {
    std::free(rp);
    std::free(col);
    return -1;
  }
  rp[0] = 0;
  for (int64_t r = 0; r < m; r++) rp[r + 1] = rp[r] + static_cast<int>(uniq[r]);
#pragma omp parallel for schedule(static)
  for (int64_t r = 0; r < m; r++)
    std::memcpy(col + rp[r], tmp.data() + start[r], sizeof(int) * static_cast<size_t>(uniq[r]));
  *rp_out = rp;
  *col_out = col;
  return nnz;
}

void sblas_gen_free(void *p) { std::free(p); }

// CSR -> CSC (stable: row indices ascending within a column) and CSR -> COO
// row indices; host helpers for building the other formats of one matrix.
int sblas_gen_csr_to_csc(int m, int n, const int *rp, const int *col, const double *val,
                         int *cp, int *ri, double *cval) {
  if (m < 0 || n < 0 || !rp || !cp) return 1;
  const int nnz = rp[m];
  std::fill(cp, cp + n + 1, 0);
  for (int j = 0; j < nnz; j++) cp[col[j] + 1]++;
  for (int c = 0; c < n; c++) cp[c + 1] += cp[c];
  std::vector<int> next(cp, cp + n);
  for (int r = 0; r < m; r++)
    for (int j = rp[r]; j < rp[r + 1]; j++) {
      const int pos = next[col[j]]++;
      ri[pos] = r;
      if (cval) cval[pos] = val[j];
    }
  return 0;
}

int sblas_gen_csr_to_coo_rows(int m, const int *rp, int *row_idx) {
  if (m < 0 || !rp) return 1;
#pragma omp parallel for schedule(static)
  for (int r = 0; r < m; r++)
    for (int j = rp[r]; j < rp[r + 1]; j++) row_idx[j] = r;
  return 0;
}

}  // extern "C"
