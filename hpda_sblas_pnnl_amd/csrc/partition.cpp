// partition.cpp -- chunk descriptors of the reference's partition rules and
// the execution ranges the engine runs them with.
//
// Descriptor rules (reference file:line):
//   rowblock      src/dspmv_mgpu_baseline.cu:64-70 (int (d*m)/C)
//   baseline_coo  src/dspmv_mgpu_baseline_coo.cu:65-73
//   v1            src/dspmv_mgpu_v1.cu:119-157 (int32 products i*nnz)
//   v1_numa       src/dspmv_mgpu_v1_numa.cu:210-255 (level 1, double formula),
//                 :432-497 (level 2 + chunk-0 / last-chunk flag overrides)
//   v1_numa_coo   src/dspmv_mgpu_v1_numa_coo.cu:75-111, :211-250
//   v1_numa_csc   src/dspmv_mgpu_v1_numa_csc.cu:79-108, :220-250
//   v2            src/dspmv_mgpu_v2.cu:114 (nb clamp), :318-366
//   NumaContext   include/spmv_kernel.h:150-202
#include "partition.hpp"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>

namespace sblas {

int search_row_of(int n, const int *a, int64_t alen, int idx) {
  int lo = 0, hi = n;
  while (hi - lo > 1) {
    const int mid = lo + (hi - lo) / 2;
    if (idx < a[mid]) hi = mid; else lo = mid;
  }
  if (a[lo] == idx) return lo;
  if (hi < alen && a[hi] == idx) return hi;
  return lo;
}

int search_first(const int *a, int n, int64_t alen, int key) {
  int lo = 0, hi = n;
  while (hi - lo > 1) {
    const int mid = lo + (hi - lo) / 2;
    if (key <= a[mid]) hi = mid; else lo = mid;
  }
  if (a[lo] == key) return lo;
  if (hi < alen && a[hi] == key) return hi;
  return lo;
}

int search_last(const int *a, int n, int64_t alen, int key) {
  int lo = 0, hi = n;
  while (hi - lo > 1) {
    const int mid = lo + (hi - lo) / 2;
    if (key < a[mid]) hi = mid; else lo = mid;
  }
  if (a[lo] == key) return lo;
  if (hi < alen && a[hi] == key) return hi;
  return lo;
}

namespace {

sblas_chunk_desc make_desc(int64_t si, int64_t ei, int sr, int er, bool sf, bool ef,
                           int mc, int64_t nnzc) {
  sblas_chunk_desc d;
  std::memset(&d, 0, sizeof(d));
  d.start_idx = si;
  d.end_idx = ei;
  d.start_row = sr;
  d.end_row = er;
  d.start_flag = sf ? 1 : 0;
  d.end_flag = ef ? 1 : 0;
  d.m_c = mc;
  d.nnz_c = nnzc;
  return d;
}

// NumaContext: nodes = max(mapping)+1, gpus per node, prefix of gpus before
// each node, and each chunk's index among the chunks of its node.
struct Numa {
  int nodes = 0;
  std::vector<int> node_of, gpus, before, local;
};

bool numa_context(const int *mapping, int C, Numa &ctx) {
  ctx.node_of.assign(C, 0);
  int top = 0;
  for (int i = 0; i < C; i++) {
    const int v = mapping ? mapping[i] : 0;
    if (v < 0 || v > 65535) return false;
    ctx.node_of[i] = v;
    top = std::max(top, v);
  }
  ctx.nodes = top + 1;
  ctx.gpus.assign(ctx.nodes, 0);
  ctx.local.assign(C, 0);
  for (int i = 0; i < C; i++) ctx.local[i] = ctx.gpus[ctx.node_of[i]]++;
  ctx.before.assign(ctx.nodes + 1, 0);
  for (int k = 0; k < ctx.nodes; k++) ctx.before[k + 1] = ctx.before[k] + ctx.gpus[k];
  return true;
}

}  // namespace

int partition_rowblock(int m, const int *rp, int C, sblas_chunk_desc *out) {
  if (m < 0 || C < 1 || !rp || !out) return SBLAS_ERR_ARGS;
  if (static_cast<int64_t>(C) * m > INT_MAX) return SBLAS_ERR_UNDEFINED;
  for (int d = 0; d < C; d++) {
    const int sr = (d * m) / C;
    const int er = ((d + 1) * m) / C - 1;
    const int64_t nnzc = static_cast<int64_t>(rp[er + 1]) - rp[sr];
    out[d] = make_desc(rp[sr], rp[sr] + nnzc - 1, sr, er, false, false, er - sr + 1, nnzc);
  }
  return SBLAS_OK;
}

int partition_baseline_coo(int m, int nnz, const int *ri, int C, sblas_chunk_desc *out) {
  if (m < 0 || nnz < 1 || C < 1 || !ri || !out) return SBLAS_ERR_ARGS;
  if (static_cast<int64_t>(C) * m > INT_MAX) return SBLAS_ERR_UNDEFINED;
  for (int d = 0; d < C; d++) {
    const int sr = (d * m) / C;
    const int er = ((d + 1) * m) / C - 1;
    const int si = search_first(ri, nnz, nnz, sr);
    const int ei = search_last(ri, nnz, nnz, er);
    out[d] = make_desc(si, ei, sr, er, false, false, er - sr + 1,
                       static_cast<int64_t>(ei) - si + 1);
  }
  return SBLAS_OK;
}

int partition_v1(int m, const int *rp, int nnz, int C, sblas_chunk_desc *out) {
  if (m < 1 || nnz < 0 || C < 1 || !rp || !out) return SBLAS_ERR_ARGS;
  if (static_cast<int64_t>(C) * nnz > INT_MAX) return SBLAS_ERR_UNDEFINED;
  for (int i = 0; i < C; i++) {
    const int p0 = i * nnz, p1 = (i + 1) * nnz;
    const int si = static_cast<int>(std::floor(static_cast<double>(p0) / C));
    const int ei = static_cast<int>(std::floor(static_cast<double>(p1) / C)) - 1;
    const int sr = search_row_of(m, rp, static_cast<int64_t>(m) + 1, si);
    const int er = search_row_of(m, rp, static_cast<int64_t>(m) + 1, ei);
    if (er + 1 > m || er - sr + 1 < 0) return SBLAS_ERR_UNDEFINED;
    out[i] = make_desc(si, ei, sr, er, si > rp[sr], ei < rp[er + 1] - 1, er - sr + 1,
                       static_cast<int64_t>(ei) - si + 1);
  }
  return SBLAS_OK;
}

int partition_v1_numa(int m, const int *rp, int nnz, int C, const int *mapping,
                      sblas_chunk_desc *out, sblas_chunk_desc *node_out) {
  if (m < 1 || nnz < 1 || C < 1 || !rp || !out) return SBLAS_ERR_ARGS;
  Numa ctx;
  if (!numa_context(mapping, C, ctx)) return SBLAS_ERR_ARGS;
  const double share = static_cast<double>(nnz) / C;
  for (int k = 0; k < ctx.nodes; k++) {
    if (ctx.gpus[k] == 0) continue;
    const int S = static_cast<int>(std::floor(ctx.before[k] * share));
    const int E = static_cast<int>(std::floor(ctx.before[k + 1] * share)) - 1;
    const int SR = search_row_of(m + 1, rp, static_cast<int64_t>(m) + 1, S);
    const int ER = search_row_of(m + 1, rp, static_cast<int64_t>(m) + 1, E);
    if (ER + 1 > m) return SBLAS_ERR_UNDEFINED;
    const bool SF = S > rp[SR];
    const bool EF = E < rp[ER + 1] - 1;
    const int M = ER - SR + 1, N1 = E - S + 1;
    if (M < 1 || N1 < 0) return SBLAS_ERR_UNDEFINED;
    if (node_out) node_out[k] = make_desc(S, E, SR, ER, SF, EF, M, N1);
    std::vector<int> L(static_cast<size_t>(M) + 1);
    L[0] = 0;
    L[M] = N1;
    for (int j = 1; j < M; j++) L[j] = rp[SR + j] - S;
    const double share2 = static_cast<double>(N1) / ctx.gpus[k];
    for (int d = 0; d < C; d++) {
      if (ctx.node_of[d] != k) continue;
      const int j = ctx.local[d];
      const int si = static_cast<int>(std::floor(j * share2));
      const int ei = static_cast<int>(std::floor((j + 1) * share2)) - 1;
      const int sr = search_row_of(M, L.data(), static_cast<int64_t>(M) + 1, si);
      const int er = search_row_of(M, L.data(), static_cast<int64_t>(M) + 1, ei);
      if (er + 1 > M || er - sr + 1 < 0) return SBLAS_ERR_UNDEFINED;
      const bool sf = (j == 0) ? SF : si > L[sr];
      const bool ef = (j + 1 == ctx.gpus[k]) ? EF : ei < L[er + 1] - 1;
      out[d] = make_desc(static_cast<int64_t>(S) + si, static_cast<int64_t>(S) + ei, SR + sr,
                         SR + er, sf, ef, er - sr + 1, static_cast<int64_t>(ei) - si + 1);
    }
  }
  return SBLAS_OK;
}

int partition_v1_numa_coo(int m, int nnz, const int *ri, int C, const int *mapping,
                          sblas_chunk_desc *out, sblas_chunk_desc *node_out) {
  if (m < 1 || nnz < 1 || C < 1 || !ri || !out) return SBLAS_ERR_ARGS;
  Numa ctx;
  if (!numa_context(mapping, C, ctx)) return SBLAS_ERR_ARGS;
  const double share = static_cast<double>(nnz) / C;
  for (int k = 0; k < ctx.nodes; k++) {
    if (ctx.gpus[k] == 0) continue;
    const int S = static_cast<int>(std::floor(ctx.before[k] * share));
    const int E = static_cast<int>(std::floor(ctx.before[k + 1] * share)) - 1;
    if (S < 0 || S >= nnz || E < 0 || E >= nnz) return SBLAS_ERR_UNDEFINED;
    const int SR = ri[S], ER = ri[E];
    const int N1 = E - S + 1;
    if (N1 < 1) return SBLAS_ERR_UNDEFINED;
    const bool SF = S > search_first(ri, nnz, nnz, SR);
    const bool EF = E < search_last(ri, nnz, nnz, ER);
    if (node_out) node_out[k] = make_desc(S, E, SR, ER, SF, EF, ER - SR + 1, N1);
    // node-local row indices (the reference's rebased staging copy)
    std::vector<int> R(static_cast<size_t>(N1));
    for (int i = 0; i < N1; i++) R[i] = ri[S + i] - SR;
    const double share2 = static_cast<double>(N1) / ctx.gpus[k];
    for (int d = 0; d < C; d++) {
      if (ctx.node_of[d] != k) continue;
      const int j = ctx.local[d];
      const int si = static_cast<int>(std::floor(j * share2));
      const int ei = static_cast<int>(std::floor((j + 1) * share2)) - 1;
      if (si < 0 || si >= N1 || ei < 0 || ei >= N1) return SBLAS_ERR_UNDEFINED;
      const int sr = R[si], er = R[ei];
      const int nnzc = ei - si + 1;
      // level-2 searches cover [0, nnz_c) of the node array, as the reference's do
      const bool sf = (j == 0) ? SF : si > search_first(R.data(), nnzc, N1, sr);
      const bool ef = (j + 1 == ctx.gpus[k]) ? EF : ei < search_last(R.data(), nnzc, N1, er);
      out[d] = make_desc(static_cast<int64_t>(S) + si, static_cast<int64_t>(S) + ei, SR + sr,
                         SR + er, sf, ef, er - sr + 1, nnzc);
    }
  }
  return SBLAS_OK;
}

int partition_v1_numa_csc(int n, const int *cp, long long nnz, int C, const int *mapping,
                          sblas_chunk_desc *out, sblas_chunk_desc *node_out) {
  if (n < 1 || nnz < 1 || nnz > INT_MAX || C < 1 || !cp || !out) return SBLAS_ERR_ARGS;
  Numa ctx;
  if (!numa_context(mapping, C, ctx)) return SBLAS_ERR_ARGS;
  const double share = static_cast<double>(nnz) / C;
  for (int k = 0; k < ctx.nodes; k++) {
    if (ctx.gpus[k] == 0) continue;
    const int S = static_cast<int>(std::floor(ctx.before[k] * share));
    const int E = static_cast<int>(std::floor(ctx.before[k + 1] * share)) - 1;
    const int SC = search_row_of(n + 1, cp, static_cast<int64_t>(n) + 1, S);
    const int EC = search_row_of(n + 1, cp, static_cast<int64_t>(n) + 1, E);
    if (EC + 1 > n) return SBLAS_ERR_UNDEFINED;
    const bool SF = S > cp[SC];
    const bool EF = E < cp[EC + 1] - 1;
    const int N = EC - SC + 1, N1 = E - S + 1;
    if (N < 1 || N1 < 0) return SBLAS_ERR_UNDEFINED;
    if (node_out) node_out[k] = make_desc(S, E, SC, EC, SF, EF, N, N1);
    std::vector<int> L(static_cast<size_t>(N) + 1);
    L[0] = 0;
    L[N] = N1;
    for (int j = 1; j < N; j++) L[j] = cp[SC + j] - S;
    const double share2 = static_cast<double>(N1) / ctx.gpus[k];
    for (int d = 0; d < C; d++) {
      if (ctx.node_of[d] != k) continue;
      const int j = ctx.local[d];
      const int si = static_cast<int>(std::floor(j * share2));
      const int ei = static_cast<int>(std::floor((j + 1) * share2)) - 1;
      const int sc = search_row_of(N, L.data(), static_cast<int64_t>(N) + 1, si);
      const int ec = search_row_of(N, L.data(), static_cast<int64_t>(N) + 1, ei);
      if (ec + 1 > N || ec - sc + 1 < 0) return SBLAS_ERR_UNDEFINED;
      // no chunk-0 / last-chunk overrides on the CSC path
      out[d] = make_desc(static_cast<int64_t>(S) + si, static_cast<int64_t>(S) + ei, SC + sc,
                         SC + ec, si > L[sc], ei < L[ec + 1] - 1, ec - sc + 1,
                         static_cast<int64_t>(ei) - si + 1);
    }
  }
  return SBLAS_OK;
}

static int v2_nb_clamped(int nb, int q) {
  const int cap =
      static_cast<int>(0.8 * 16 * 1e9 / static_cast<double>(sizeof(double) + 2 * sizeof(int))) /
      q;
  return std::min(nb, cap);
}

int v2_num_tasks(int nnz, int nb, int q) {
  if (nnz < 0 || nb < 1 || q < 1) return -SBLAS_ERR_ARGS;
  nb = v2_nb_clamped(nb, q);
  if (nb < 1) return -SBLAS_ERR_ARGS;
  if (static_cast<int64_t>(nnz) + nb - 1 > INT_MAX) return -SBLAS_ERR_UNDEFINED;
  return (nnz + nb - 1) / nb;
}

int partition_v2(int m, const int *rp, int nnz, int nb, int q, sblas_chunk_desc *out,
                 int cap) {
  if (m < 1 || !rp || !out) return -SBLAS_ERR_ARGS;
  const int T = v2_num_tasks(nnz, nb, q);
  if (T < 0) return T;
  if (T > cap) return -SBLAS_ERR_ARGS;
  if (static_cast<int64_t>(T) * nnz > INT_MAX) return -SBLAS_ERR_UNDEFINED;
  for (int t = 0; t < T; t++) {
    const int si = (t * nnz) / T;          // integer division, then floor((double)..)
    const int ei = ((t + 1) * nnz) / T - 1;
    const int sr = search_row_of(m, rp, static_cast<int64_t>(m) + 1, si);
    const int er = search_row_of(m, rp, static_cast<int64_t>(m) + 1, ei);
    if (er + 1 > m || er - sr + 1 < 0) return -SBLAS_ERR_UNDEFINED;
    out[t] = make_desc(si, ei, sr, er, si > rp[sr], ei < rp[er + 1] - 1, er - sr + 1,
                       static_cast<int64_t>(ei) - si + 1);
  }
  return T;
}

// ---------------------------------------------------------------------------
// execution ranges
// ---------------------------------------------------------------------------

// Row containing element e (rows with rp[r] <= e < rp[r+1]); the last such
// row index when empty rows share the offset.
static int row_containing(int m, const int *rp, int64_t e) {
  // upper_bound over rp[0..m] then step back
  const int *it = std::upper_bound(rp, rp + m + 1, static_cast<int>(e));
  int r = static_cast<int>(it - rp) - 1;
  return std::min(std::max(r, 0), m - 1);
}

std::vector<ExecChunk> exec_chunks_from_bounds(int m, const int *rp,
                                               const std::vector<int64_t> &bounds) {
  std::vector<ExecChunk> out;
  std::vector<std::pair<int64_t, int64_t>> ranges;
  for (size_t i = 0; i + 1 < bounds.size(); i++)
    if (bounds[i + 1] > bounds[i]) ranges.push_back({bounds[i], bounds[i + 1]});
  if (ranges.empty()) {
    out.push_back(ExecChunk{0, 0, 0, m});
    return out;
  }
  const size_t C = ranges.size();
  std::vector<int> sr(C), er(C);
  for (size_t c = 0; c < C; c++) {
    sr[c] = row_containing(m, rp, ranges[c].first);
    er[c] = row_containing(m, rp, ranges[c].second - 1);
  }
  for (size_t c = 0; c < C; c++) {
    ExecChunk e;
    e.lo = ranges[c].first;
    e.hi = ranges[c].second;
    e.row_begin = (c == 0) ? 0 : sr[c];
    if (c + 1 == C)
      e.row_end = m;
    else
      e.row_end = (er[c] == sr[c + 1]) ? er[c] + 1 : sr[c + 1];
    out.push_back(e);
  }
  return out;
}

std::vector<int64_t> element_bounds(int part_mode, int m, const int *rp, int64_t nnz, int C,
                                    int nb) {
  std::vector<int64_t> b;
  if (C < 1) C = 1;
  switch (part_mode) {
    case SBLAS_PART_V1_NUMA: {
      // level 1 at one node covers [0, E]; level 2 splits it into C pieces;
      // the execution keeps the elements the double formula drops at the end.
      const double share = static_cast<double>(nnz) / C;
      const int64_t N1 = static_cast<int64_t>(std::floor(C * share));
      const double share2 = static_cast<double>(N1) / C;
      for (int j = 0; j < C; j++) b.push_back(static_cast<int64_t>(std::floor(j * share2)));
      break;
    }
    case SBLAS_PART_V2: {
      if (nb < 1) nb = 1;
      const int64_t T = (nnz + nb - 1) / nb;
      for (int64_t t = 0; t < std::max<int64_t>(T, 1); t++) b.push_back((t * nnz) / std::max<int64_t>(T, 1));
      break;
    }
    case SBLAS_PART_ROWBLOCK: {
      for (int d = 0; d < C; d++)
        b.push_back(rp[static_cast<int64_t>(d) * m / C]);
      break;
    }
    case SBLAS_PART_V1:
    default:
      for (int i = 0; i < C; i++) b.push_back(static_cast<int64_t>(i) * nnz / C);
      break;
  }
  b.push_back(nnz);
  for (size_t i = 1; i < b.size(); i++) b[i] = std::max(b[i], b[i - 1]);
  b[0] = 0;
  return b;
}

}  // namespace sblas
