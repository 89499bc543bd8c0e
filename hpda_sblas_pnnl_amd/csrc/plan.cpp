// plan.cpp -- analysis and execution of SpMV plans on device-resident arrays.
//
// A plan is the engine's counterpart of the reference's per-call work
// (partition src/dspmv_mgpu_v1.cu:119-193, kernel switch :256-278, merge
// :304-319) and of the CSR5 handle (include/anonymouslib_cuda.h:11-53): it
// cuts the matrix into chunks with one of the reference's partition rules,
// picks a kernel and its launch shape from row-length statistics, uploads the
// chunk table and split-row merge lists once, and then runs
// y = alpha*A*x + beta*y on a caller stream as often as asked.
#include "plan.hpp"

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <new>

namespace sblas {

namespace {
thread_local int g_last_error = SBLAS_OK;
thread_local std::string g_last_msg;

int pow2_round(double v, int lo, int hi) {
  if (!(v > 0)) return lo;
  const int p = 1 << static_cast<int>(std::lround(std::log2(v)));
  return std::min(std::max(p, lo), hi);
}

int env_int(const char *name, int fallback) {
  const char *e = std::getenv(name);
  return (e && *e) ? std::atoi(e) : fallback;
}
}  // namespace

void set_error(int status, const std::string &what) {
  g_last_error = status;
  g_last_msg = what;
  static const bool verbose = [] {
    const char *e = std::getenv("SBLAS_LOG");
    return e && std::atoi(e) > 0;
  }();
  if (verbose && status != SBLAS_OK) std::fprintf(stderr, "sblas: %s\n", what.c_str());
}
int last_error() { return g_last_error; }
const char *last_message() { return g_last_msg.c_str(); }
void clear_error() {
  g_last_error = SBLAS_OK;
  g_last_msg.clear();
}

RowStats row_stats(int m, const int *rp) {
  RowStats s;
  if (m <= 0) return s;
  double sum = 0, sum2 = 0;
  int64_t mx = 0, empty = 0;
  for (int r = 0; r < m; r++) {
    const int64_t L = static_cast<int64_t>(rp[r + 1]) - rp[r];
    sum += static_cast<double>(L);
    sum2 += static_cast<double>(L) * static_cast<double>(L);
    mx = std::max(mx, L);
    empty += (L == 0);
  }
  s.mean = sum / m;
  const double var = std::max(0.0, sum2 / m - s.mean * s.mean);
  s.cv = s.mean > 0 ? std::sqrt(var) / s.mean : 0.0;
  s.max_row = mx;
  s.empty_frac = static_cast<double>(empty) / m;
  return s;
}

// Lanes per row: about three quarters of the mean row length, so most rows
// finish in one or two strides of the sub-wave.
int autotune_subwave(const RowStats &s) { return pow2_round(s.mean * 0.75, 2, 64); }

// CSR5-style tile depth: the reference's rule on integer nnz / m
// (include/anonymouslib_cuda.h:294-318: <= 4 -> 4, <= 32 -> nnz/m, <= 256 ->
// 32, else 6 elements per lane), snapped to the depths this kernel is built
// for (8, 16, 32), nearest in log2.
int autotune_sigma(int64_t nnz, int m) {
  const int64_t r = m > 0 ? nnz / m : 0;
  const int ref = r <= 4 ? 4 : (r <= 32 ? static_cast<int>(r) : (r <= 256 ? 32 : 6));
  return ref <= 11 ? 8 : (ref <= 22 ? 16 : 32);
}

namespace {

// Cached decisions of the timed trial, keyed by the row-length profile.
struct TuneKey {
  int val_type, mean_q, cv_q, empty_q, nnz_q;
  bool operator<(const TuneKey &o) const {
    if (val_type != o.val_type) return val_type < o.val_type;
    if (mean_q != o.mean_q) return mean_q < o.mean_q;
    if (cv_q != o.cv_q) return cv_q < o.cv_q;
    if (empty_q != o.empty_q) return empty_q < o.empty_q;
    return nnz_q < o.nnz_q;
  }
};
struct TuneResult {
  int kernel, variant;
  float ms[4];  // trial medians: row blocks, CSR5-style, column-sorted row blocks, segments
};
std::mutex g_tune_mu;
std::map<TuneKey, TuneResult> g_tune_cache;

TuneKey tune_key(int val_type, const RowStats &s, int64_t nnz) {
  auto q = [](double v, double per_octave) {
    return static_cast<int>(std::lround(std::log2(std::max(v, 1e-3)) * per_octave));
  };
  return TuneKey{val_type, q(s.mean, 4), q(1.0 + s.cv, 4),
                 static_cast<int>(std::lround(s.empty_frac * 10)), q(static_cast<double>(nnz), 1)};
}

}  // namespace
// ---------------------------------------------------------------------------

Plan::~Plan() {
  for (void *p : panel.buffers) (void)hipFree(p);
  for (void *p : sorted.buffers) (void)hipFree(p);
  for (void *p : seg.buffers) (void)hipFree(p);
  for (hipEvent_t e : tev) (void)hipEventDestroy(e);
  for (void *p : device_buffers) (void)hipFree(p);
  for (hipStream_t s : streams) (void)hipStreamDestroy(s);
  for (hipEvent_t e : events) (void)hipEventDestroy(e);
}

void *Plan::dev_alloc(size_t bytes) {
  if (bytes == 0) bytes = 8;
  void *p = nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
  device_buffers.push_back(p);
  buffer_bytes.push_back(workspace_bytes);
  workspace_bytes += static_cast<int64_t>(bytes);
  return p;
}

// Vector-kernel launch shape over a set of execution chunks: a block owns
// rows_per_block consecutive rows of one chunk.
VectorLaunch vector_shape(const std::vector<ExecChunk> &ex, double mean, int subwave,
                          int xcd_remap, std::vector<DevChunk> &dev) {
  VectorLaunch L{};
  L.subwave = subwave;
  const int subs = 256 / subwave;
  int64_t rows_total = 0;
  for (const ExecChunk &e : ex) rows_total += e.row_end - e.row_begin;
  // about 4K elements per block, but keep >= ~2048 blocks when rows allow
  int per_sub = static_cast<int>(std::max(1.0, 4096.0 / (subs * std::max(1.0, mean))));
  const int64_t cap = std::max<int64_t>(1, rows_total / (static_cast<int64_t>(subs) * 2048));
  per_sub = static_cast<int>(std::min<int64_t>(per_sub, cap));
  L.rows_per_block = subs * std::max(1, per_sub);
  dev.clear();
  int blk = 0;
  for (const ExecChunk &e : ex) {
    DevChunk d{};
    d.lo = static_cast<int>(e.lo);
    d.hi = static_cast<int>(e.hi);
    d.row_begin = e.row_begin;
    d.row_end = e.row_end;
    d.blk_begin = blk;
    blk += (e.row_end - e.row_begin + L.rows_per_block - 1) / L.rows_per_block;
    dev.push_back(d);
  }
  L.total_blocks = blk;
  L.xcd_remap = xcd_remap;
  L.grid = xcd_remap ? ((blk + 7) / 8) * 8 : blk;
  return L;
}

// Split-row merge lists: every row that more than one chunk touches, with the
// partial slots of those chunks in chunk order (slot 2c = chunk c's first
// row, 2c+1 = its last row).
void merge_lists(const std::vector<ExecChunk> &ex, const int *rp, std::vector<int> &rows,
                 std::vector<int> &ptr, std::vector<int> &slots) {
  rows.clear();
  ptr.assign(1, 0);
  slots.clear();
  for (size_t c = 0; c < ex.size(); c++) {
    const ExecChunk &e = ex[c];
    if (e.row_end <= e.row_begin) continue;
    const int first = e.row_begin, last = e.row_end - 1;
    auto split = [&](int r) { return rp[r] < e.lo || rp[r + 1] > e.hi; };
    auto add = [&](int r, int slot) {
      if (!rows.empty() && rows.back() == r) {
        slots.push_back(slot);
        ptr.back() = static_cast<int>(slots.size());
      } else {
        rows.push_back(r);
        slots.push_back(slot);
        ptr.push_back(static_cast<int>(slots.size()));
      }
    };
    if (split(first)) add(first, static_cast<int>(2 * c));
    if (last != first && split(last)) add(last, static_cast<int>(2 * c + 1));
  }
}

namespace {

template <typename T>
int upload(Plan &p, T **dst, const std::vector<T> &src) {
  *dst = static_cast<T *>(p.dev_alloc(sizeof(T) * std::max<size_t>(src.size(), 1)));
  if (!*dst) {
    set_error(SBLAS_ERR_ALLOC, "plan: device allocation failed");
    return SBLAS_ERR_ALLOC;
  }
  if (!src.empty())
    SBLAS_HIP_TRY(hipMemcpy(*dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice));
  return SBLAS_OK;
}

}  // namespace

int Plan::build_csr_chunks(const int *h_rp) {
  const int C = chunks;
  const int inz = static_cast<int>(nnz);
  const int q = std::max(opts.streams, 1);
  desc_status = SBLAS_OK;
  switch (opts.part_mode) {
    case SBLAS_PART_V1_NUMA:
      descs.assign(C, sblas_chunk_desc{});
      desc_status = nnz > 0 ? partition_v1_numa(m, h_rp, inz, C, nullptr, descs.data(), nullptr)
                            : SBLAS_ERR_UNDEFINED;
      break;
    case SBLAS_PART_V2: {
      const int T = v2_num_tasks(inz, std::max(opts.nb, 1), q);
      if (T > 0) {
        descs.assign(T, sblas_chunk_desc{});
        const int got = partition_v2(m, h_rp, inz, std::max(opts.nb, 1), q, descs.data(), T);
        desc_status = got > 0 ? SBLAS_OK : -got;
      } else {
        descs.clear();
        desc_status = T == 0 ? SBLAS_OK : -T;
      }
      break;
    }
    case SBLAS_PART_ROWBLOCK:
      descs.assign(C, sblas_chunk_desc{});
      desc_status = partition_rowblock(m, h_rp, C, descs.data());
      break;
    case SBLAS_PART_V1:
    default:
      descs.assign(C, sblas_chunk_desc{});
      desc_status = partition_v1(m, h_rp, inz, C, descs.data());
      break;
  }
  int nb_eff = 1;
  if (opts.part_mode == SBLAS_PART_V2) {
    const int cap = static_cast<int>(0.8 * 16 * 1e9 / 16.0) / q;
    nb_eff = std::max(1, std::min(std::max(opts.nb, 1), cap));
  }
  exec = exec_chunks_from_bounds(m, h_rp, element_bounds(opts.part_mode, m, h_rp, nnz, C, nb_eff));
  std::vector<int> mrows, mptr, mslots;
  merge_lists(exec, h_rp, mrows, mptr, mslots);
  ngroups = static_cast<int>(mrows.size());
  split_rows = ngroups;
  int st;
  if ((st = upload(*this, &d_mrows, mrows)) != SBLAS_OK) return st;
  if ((st = upload(*this, &d_mptr, mptr)) != SBLAS_OK) return st;
  if ((st = upload(*this, &d_mslots, mslots)) != SBLAS_OK) return st;
  d_partials = static_cast<double *>(dev_alloc(sizeof(double) * 2 * exec.size() * max_k));
  if (!d_partials) {
    set_error(SBLAS_ERR_ALLOC, "plan: partials allocation failed");
    return SBLAS_ERR_ALLOC;
  }
  return SBLAS_OK;
}

int Plan::setup_vector(const int *h_rp) {
  (void)h_rp;
  // the sub-wave option targets the kernel in use; this form needs 2..64
  int sub = (kernel == SBLAS_KERNEL_VECTOR && variant == 1 && opts.subwave >= 2)
                ? opts.subwave
                : autotune_subwave(stats);
  if (sub < 2 || sub > 64 || (sub & (sub - 1))) {
    set_error(SBLAS_ERR_ARGS, "subwave must be a power of two in [2, 64]");
    return SBLAS_ERR_ARGS;
  }
  int remap = opts.xcd_remap;
  if (remap < 0) remap = env_int("SBLAS_XCD_REMAP", 1);
  vl = vector_shape(exec, stats.mean, sub, remap ? 1 : 0, dev_chunks);
  int st;
  if ((st = upload(*this, &d_chunks, dev_chunks)) != SBLAS_OK) return st;
  if (opts.streams > 1 && exec.size() > 1) {
    // one launch per chunk: per-chunk tables with blk_begin = 0
    std::vector<DevChunk> solo;
    solo_vl.clear();
    for (size_t c = 0; c < exec.size(); c++) {
      std::vector<ExecChunk> one(1, exec[c]);
      std::vector<DevChunk> d1;
      solo_vl.push_back(vector_shape(one, stats.mean, sub, remap ? 1 : 0, d1));
      solo.push_back(d1[0]);
    }
    if ((st = upload(*this, &d_chunks_solo, solo)) != SBLAS_OK) return st;
    for (int i = 0; i < opts.streams; i++) {
      hipStream_t s;
      SBLAS_HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
      streams.push_back(s);
    }
    for (size_t i = 0; i < exec.size() + 1; i++) {
      hipEvent_t e;
      SBLAS_HIP_TRY(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      events.push_back(e);
    }
  }
  return SBLAS_OK;
}

// Row blocks over the execution chunks: consecutive rows are packed while
// their elements fit one block (cap elements, at most kStreamRowsMax rows);
// a row with more elements than that inside its chunk is cut into cap-sized
// single-row pieces.  Every row of [0, m) lands in some block.
std::vector<RowBlock> build_row_blocks(const std::vector<ExecChunk> &ex, const int *rp,
                                       int cap) {
  std::vector<RowBlock> out;
  for (const ExecChunk &ch : ex) {
    const int64_t lo = ch.lo, hi = ch.hi;
    int r = ch.row_begin;
    while (r < ch.row_end) {
      const int64_t s = std::max<int64_t>(rp[r], lo);
      const int64_t e = std::min<int64_t>(rp[r + 1], hi);
      if (e - s > cap) {
        for (int64_t p = s; p < e; p += cap)
          out.push_back(RowBlock{r, r + 1, static_cast<int>(p),
                                 static_cast<int>(std::min<int64_t>(p + cap, e))});
        r++;
        continue;
      }
      const int rb = r;
      r++;
      while (r < ch.row_end && r - rb < kStreamRowsMax &&
             std::min<int64_t>(rp[r + 1], hi) - s <= cap)
        r++;
      const int64_t bhi = std::min<int64_t>(std::max<int64_t>(rp[r], lo), hi);
      out.push_back(RowBlock{rb, r, static_cast<int>(s), static_cast<int>(std::max(bhi, s))});
    }
  }
  return out;
}

int Plan::setup_stream(const int *h_rp) {
  int items = opts.items > 0 ? opts.items : 8;
  if (items != 8 && items != 16) {
    set_error(SBLAS_ERR_ARGS, "items must be 8 or 16");
    return SBLAS_ERR_ARGS;
  }
  const int cap = kStreamBlock * items;
  rblocks = build_row_blocks(exec, h_rp, cap);
  int sub = opts.subwave > 0 ? opts.subwave : pow2_round(stats.mean / 4.0, 1, 64);
  if (sub < 1 || sub > 64 || (sub & (sub - 1))) {
    set_error(SBLAS_ERR_ARGS, "subwave must be a power of two in [1, 64]");
    return SBLAS_ERR_ARGS;
  }
  int remap = opts.xcd_remap;
  if (remap < 0) remap = env_int("SBLAS_XCD_REMAP", 1);
  sl.items = items;
  sl.subwave = sub;
  sl.nblocks = static_cast<int>(rblocks.size());
  sl.xcd_remap = remap ? 1 : 0;
  sl.grid = remap ? ((sl.nblocks + 7) / 8) * 8 : sl.nblocks;
  int st;
  if ((st = upload(*this, &d_rblocks, rblocks)) != SBLAS_OK) return st;
  std::vector<ExecChunk> as_chunks;
  as_chunks.reserve(rblocks.size());
  for (const RowBlock &b : rblocks) as_chunks.push_back(ExecChunk{b.lo, b.hi, b.row_begin, b.row_end});
  std::vector<int> mrows, mptr, mslots;
  merge_lists(as_chunks, h_rp, mrows, mptr, mslots);
  bgroups = static_cast<int>(mrows.size());
  if ((st = upload(*this, &d_bmrows, mrows)) != SBLAS_OK) return st;
  if ((st = upload(*this, &d_bmptr, mptr)) != SBLAS_OK) return st;
  if ((st = upload(*this, &d_bmslots, mslots)) != SBLAS_OK) return st;
  // 2 slots per block, max_k values per slot (multi-RHS blocks use the same list)
  d_bpartials = static_cast<double *>(
      dev_alloc(sizeof(double) * 2 * max_k * std::max<size_t>(rblocks.size(), 1)));
  if (!d_bpartials) {
    set_error(SBLAS_ERR_ALLOC, "plan: partials allocation failed");
    return SBLAS_ERR_ALLOC;
  }
  // Multi-vector products on large matrices with irregular rows take blocks
  // of 16 elements a thread: R-MAT scale 22, k = 2 / 4 / 8 / 16: 0.81 / 1.18 /
  // 1.52 / 2.22 ms against 0.91 / 1.32 / 1.67 / 2.55 with 8.  A banded matrix
  // of 5.6e7 elements (row CV 0.3) ran 1 % slower with them, the power-law
  // one of 3.2e5 elements 1.6x slower (too few blocks).
  if (opts.items == 0 && nnz >= (int64_t{1} << 23) && stats.cv >= 1.0) {
    const std::vector<RowBlock> kb = build_row_blocks(exec, h_rp, kStreamBlock * 16);
    if ((st = upload(*this, &d_kblocks, kb)) != SBLAS_OK) return st;
    as_chunks.clear();
    for (const RowBlock &b : kb) as_chunks.push_back(ExecChunk{b.lo, b.hi, b.row_begin, b.row_end});
    merge_lists(as_chunks, h_rp, mrows, mptr, mslots);
    kgroups = static_cast<int>(mrows.size());
    if ((st = upload(*this, &d_kmrows, mrows)) != SBLAS_OK) return st;
    if ((st = upload(*this, &d_kmptr, mptr)) != SBLAS_OK) return st;
    if ((st = upload(*this, &d_kmslots, mslots)) != SBLAS_OK) return st;
    d_kpartials = static_cast<double *>(
        dev_alloc(sizeof(double) * 2 * max_k * std::max<size_t>(kb.size(), 1)));
    if (!d_kpartials) {
      set_error(SBLAS_ERR_ALLOC, "plan: partials allocation failed");
      return SBLAS_ERR_ALLOC;
    }
    slk = sl;
    slk.items = 16;
    slk.nblocks = static_cast<int>(kb.size());
    slk.grid = slk.xcd_remap ? ((slk.nblocks + 7) / 8) * 8 : slk.nblocks;
  }
  return SBLAS_OK;
}

hipEvent_t *Plan::timing_slot() {
  if (!timing) return nullptr;
  const int cap = 4096;
  if (tcount >= cap) return nullptr;
  while (static_cast<int>(tev.size()) < 2 * (tcount + 1)) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    tev.push_back(e);
  }
  return &tev[2 * tcount++];
}

int Plan::run_spmv(double alpha, const double *x, double beta, double *y, hipStream_t s) {
  return run_spmm(1, alpha, x, beta, y, s);
}

int Plan::run_spmm(int k, double alpha, const double *X, double beta, double *Y,
                   hipStream_t s) {
  launches = 0;
  if (m == 0) return SBLAS_OK;
  if (k < 1 || k > max_k) {
    set_error(SBLAS_ERR_ARGS, "spmm: k out of range");
    return SBLAS_ERR_ARGS;
  }
  int st = SBLAS_OK;
  const int C = static_cast<int>(exec.size());
  if (uses_tiles(k)) {
    // COO segmented reduction or merge-path CSR, then the ordered carry fix-up
    hipEvent_t *te = timing_slot();
    if (te) SBLAS_HIP_TRY(hipEventRecord(te[0], s));
    if (tiles == 0)
      st = launch_scale(m, beta, Y, s);
    else if (kernel == SBLAS_KERNEL_CSR5) {
      st = launch_scale_rows(c5_empty, d_c5_empty, beta, Y, s);
      if (st == SBLAS_OK)
        st = launch_csr5(val_type, sigma, tiles, nnz, d_c5_tile_nz, d_c5_flags, d_c5_rows_nz,
                         d_c5_colT, d_c5_valT, X, Y, alpha, beta, d_carry, s);
    } else if (kernel == SBLAS_KERNEL_COO)
      st = launch_coo_tiles(val_type, tile_sub, m, nnz, tiles, a0, a1, val, X, Y, alpha, beta,
                            d_carry, s);
    else
      st = launch_csr_merge_path(val_type, tile_sub, m, nnz, tiles, a0, a1, val, X, Y, alpha,
                                 beta, d_carry, s);
    if (st != SBLAS_OK) return st;
    if (te) SBLAS_HIP_TRY(hipEventRecord(te[1], s));
    launches++;
    if (tiles > 0 && merge_enabled) {
      st = launch_tile_fixup(carry_count(), d_carry, Y, alpha, beta, s);
      launches++;
    }
    return st;
  }
  if (kernel == SBLAS_KERNEL_COO) {
    set_error(SBLAS_ERR_UNSUPPORTED, "multi-vector products need a CSR, CSC or row-pointer COO plan");
    return SBLAS_ERR_UNSUPPORTED;
  }
  if (variant == 3 && k == 1 && streams.empty()) {
    hipEvent_t *te = timing_slot();
    if (te) SBLAS_HIP_TRY(hipEventRecord(te[0], s));
    st = launch_panel_spmv(val_type, panel, m, n, vl.xcd_remap, X, Y, alpha, beta, s);
    if (st != SBLAS_OK) return st;
    if (te) SBLAS_HIP_TRY(hipEventRecord(te[1], s));
    launches++;
    if (pgroups > 0 && merge_enabled) {
      st = launch_merge(d_pmrows, d_pmptr, d_pmslots, pgroups, panel.partials, 1, Y, alpha, beta,
                        s);
      if (st != SBLAS_OK) return st;
      launches++;
    }
    return SBLAS_OK;
  }
  if (variant == 2 && k == 1 && streams.empty()) {
    hipEvent_t *te = timing_slot();
    if (te) SBLAS_HIP_TRY(hipEventRecord(te[0], s));
    if (x32_rows && (st = launch_x_to_f32(n, X, x32_rows, s)) != SBLAS_OK) return st;
    st = launch_csr_stream(val_type, sl, d_rblocks, a0, a1, val, X, Y, alpha, beta,
                           d_bpartials, s, x32_rows);
    if (st != SBLAS_OK) return st;
    if (te) SBLAS_HIP_TRY(hipEventRecord(te[1], s));
    launches++;
    if (bgroups > 0 && merge_enabled) {
      st = launch_merge(d_bmrows, d_bmptr, d_bmslots, bgroups, d_bpartials, 1, Y, alpha, beta, s);
      if (st != SBLAS_OK) return st;
      launches++;
    }
    return SBLAS_OK;
  }
  if (variant == 6 && k == 1 && streams.empty()) {
    // the virtual rows' sums (their cut rows merged), then y from them; the
    // timing covers all three launches
    hipEvent_t *te = timing_slot();
    if (te) SBLAS_HIP_TRY(hipEventRecord(te[0], s));
    if (x32_rows && (st = launch_x_to_f32(n, X, x32_rows, s)) != SBLAS_OK) return st;
    st = launch_csr_stream(val_type, vsl, d_vblocks, seg.vrp, seg.vcol, seg.vval, X, seg.yv, 1.0,
                           0.0, d_vpartials, s, x32_rows);
    if (st == SBLAS_OK && vgroups > 0)
      st = launch_merge(d_vmrows, d_vmptr, d_vmslots, vgroups, d_vpartials, 1, seg.yv, 1.0, 0.0,
                        s);
    if (st == SBLAS_OK) st = launch_seg_merge(seg, alpha, beta, Y, s);
    if (st != SBLAS_OK) return st;
    if (te) SBLAS_HIP_TRY(hipEventRecord(te[1], s));
    launches += 2 + (vgroups > 0 ? 1 : 0);
    return SBLAS_OK;
  }
  if (variant == 4 && k == 1 && streams.empty()) {
    hipEvent_t *te = timing_slot();
    if (te) SBLAS_HIP_TRY(hipEventRecord(te[0], s));
    st = launch_csr_sorted(val_type, sorted, a0, X, Y, alpha, beta, s);
    if (st != SBLAS_OK) return st;
    if (te) SBLAS_HIP_TRY(hipEventRecord(te[1], s));
    launches++;
    if (sgroups > 0 && merge_enabled) {
      st = launch_merge(d_smrows, d_smptr, d_smslots, sgroups, sorted.partials, 1, Y, alpha, beta,
                        s);
      if (st != SBLAS_OK) return st;
      launches++;
    }
    return SBLAS_OK;
  }
  if (variant >= 2 && k > 1 && streams.empty()) {
    // k lanes per row unless a sub-wave was asked for (banded, k = 4: 0.536
    // against 0.567 ms with the k = 1 sub-wave of 32; R-MAT, k = 2: 0.81
    // against 0.84 with 4; equal elsewhere)
    const bool own = slk.grid > 0;
    StreamLaunch L = own ? slk : sl;
    if (opts.subwave <= 0) L.subwave = k;
    hipEvent_t *te = timing_slot();
    if (te) SBLAS_HIP_TRY(hipEventRecord(te[0], s));
    st = launch_csr_blocks_mrhs(val_type, k, L, own ? d_kblocks : d_rblocks, a0, a1, val, X, Y,
                                alpha, beta, own ? d_kpartials : d_bpartials, s);
    if (st != SBLAS_OK) return st;
    if (te) SBLAS_HIP_TRY(hipEventRecord(te[1], s));
    launches++;
    const int groups = own ? kgroups : bgroups;
    if (groups > 0 && merge_enabled) {
      st = own ? launch_merge(d_kmrows, d_kmptr, d_kmslots, kgroups, d_kpartials, k, Y, alpha,
                              beta, s)
               : launch_merge(d_bmrows, d_bmptr, d_bmslots, bgroups, d_bpartials, k, Y, alpha,
                              beta, s);
      if (st != SBLAS_OK) return st;
      launches++;
    }
    return SBLAS_OK;
  }
  if (!streams.empty()) {
    // chunks over streams; the caller stream waits for all of them
    SBLAS_HIP_TRY(hipEventRecord(events[C], s));
    for (int c = 0; c < C; c++) {
      hipStream_t cs = streams[c % streams.size()];
      SBLAS_HIP_TRY(hipStreamWaitEvent(cs, events[C], 0));
      if (k == 1)
        st = launch_csr_vector(val_type, solo_vl[c], d_chunks_solo + c, 1, a0, a1, val, X, Y,
                               alpha, beta, d_partials + 2 * c, cs);
      else
        st = launch_csr_vector_mrhs(val_type, k, solo_vl[c], d_chunks_solo + c, 1, a0, a1, val,
                                    X, Y, alpha, beta, d_partials + 2 * c * k, cs);
      if (st != SBLAS_OK) return st;
      SBLAS_HIP_TRY(hipEventRecord(events[c], cs));
      launches++;
    }
    for (int c = 0; c < C; c++) SBLAS_HIP_TRY(hipStreamWaitEvent(s, events[c], 0));
  } else {
    hipEvent_t *te = timing_slot();
    if (te) SBLAS_HIP_TRY(hipEventRecord(te[0], s));
    if (k == 1)
      st = launch_csr_vector(val_type, vl, d_chunks, C, a0, a1, val, X, Y, alpha, beta,
                             d_partials, s);
    else
      st = launch_csr_vector_mrhs(val_type, k, vl, d_chunks, C, a0, a1, val, X, Y, alpha,
                                  beta, d_partials, s);
    if (st != SBLAS_OK) return st;
    if (te) SBLAS_HIP_TRY(hipEventRecord(te[1], s));
    launches++;
  }
  if (ngroups > 0 && merge_enabled) {
    st = launch_merge(d_mrows, d_mptr, d_mslots, ngroups, d_partials, k, Y, alpha, beta, s);
    if (st != SBLAS_OK) return st;
    launches++;
  }
  return SBLAS_OK;
}

int Plan::setup_tiles() {
  const int64_t items = (kernel == SBLAS_KERNEL_COO) ? nnz : static_cast<int64_t>(m) + nnz;
  tiles = static_cast<int>((items + kTileSize - 1) / kTileSize);
  tile_sub = opts.subwave > 0 ? opts.subwave : tile_subwave(stats.mean);
  if (tile_sub != 1 && tile_sub != 4 && tile_sub != 16 && tile_sub != 64) tile_sub = 16;
  d_carry = static_cast<TileCarry *>(dev_alloc(sizeof(TileCarry) * std::max(tiles, 1)));
  if (!d_carry) {
    set_error(SBLAS_ERR_ALLOC, "plan: tile carry allocation failed");
    return SBLAS_ERR_ALLOC;
  }
  return SBLAS_OK;
}

// CSR5-style format: tiles of 64 lanes x sigma elements; per tile the index
// of its first row among the non-empty rows (bit 31: the tile begins at a
// row start); per lane a word of row-start bits; the element arrays copied
// in tile-transposed order (padded to whole tiles).
int Plan::setup_csr5(const std::vector<int> &h_rp) {
  const int sg = opts.sigma > 0 ? opts.sigma : autotune_sigma(nnz, m);
  if (sg != 8 && sg != 16 && sg != 32) {
    set_error(SBLAS_ERR_ARGS, "csr5: sigma must be 8, 16 or 32");
    return SBLAS_ERR_ARGS;
  }
  sigma = sg;
  const int64_t T = 64LL * sg;
  tiles = static_cast<int>((nnz + T - 1) / T);
  std::vector<int> rows_nz, empty, nzidx(static_cast<size_t>(std::max(m, 1)), -1);
  for (int r = 0; r < m; r++) {
    if (h_rp[r + 1] > h_rp[r]) {
      nzidx[r] = static_cast<int>(rows_nz.size());
      rows_nz.push_back(r);
    } else {
      empty.push_back(r);
    }
  }
  std::vector<uint32_t> tile_nz(static_cast<size_t>(std::max(tiles, 1)), 0);
  std::vector<uint32_t> flags(static_cast<size_t>(std::max(tiles, 1)) * 64, 0);
  int r = 0;
  for (int p = 0; p < tiles; p++) {
    const int64_t e = static_cast<int64_t>(p) * T;
    while (h_rp[r + 1] <= e) r++;
    tile_nz[p] = static_cast<uint32_t>(nzidx[r]) | (h_rp[r] == e ? 0x80000000u : 0u);
  }
  for (int q : rows_nz) {
    const int64_t e = h_rp[q];
    const int64_t o = e % T;
    if (o == 0) continue;  // tile start: bit 31 of the tile word
    flags[static_cast<size_t>(e / T) * 64 + static_cast<size_t>(o / sg)] |= 1u << (o % sg);
  }
  c5_empty = static_cast<int>(empty.size());
  int st;
  if ((st = upload(*this, &d_c5_tile_nz, tile_nz)) != SBLAS_OK) return st;
  if ((st = upload(*this, &d_c5_flags, flags)) != SBLAS_OK) return st;
  if ((st = upload(*this, &d_c5_rows_nz, rows_nz)) != SBLAS_OK) return st;
  if ((st = upload(*this, &d_c5_empty, empty)) != SBLAS_OK) return st;
  const int64_t padded = static_cast<int64_t>(tiles) * T;
  const size_t vbytes = val_type == SBLAS_VAL_F64 ? 8 : (val_type == SBLAS_VAL_F32 ? 4 : 2);
  d_c5_colT = static_cast<int *>(dev_alloc(sizeof(int) * std::max<int64_t>(padded, 1)));
  d_c5_valT = dev_alloc(vbytes * std::max<int64_t>(padded, 1));
  d_carry = static_cast<TileCarry *>(dev_alloc(sizeof(TileCarry) * std::max(tiles, 1)));
  if (!d_c5_colT || !d_c5_valT || !d_carry) {
    set_error(SBLAS_ERR_ALLOC, "csr5: format allocation failed");
    return SBLAS_ERR_ALLOC;
  }
  if ((st = launch_csr5_transpose(val_type, sg, padded, nnz, a1, val, d_c5_colT, d_c5_valT,
                                  nullptr)) != SBLAS_OK)
    return st;
  SBLAS_HIP_TRY(hipStreamSynchronize(nullptr));
  return SBLAS_OK;
}

void Plan::free_buffers_from(size_t first) {
  for (size_t i = first; i < device_buffers.size(); i++) (void)hipFree(device_buffers[i]);
  device_buffers.resize(std::min(first, device_buffers.size()));
  workspace_bytes = first < buffer_bytes.size() ? buffer_bytes[first] : workspace_bytes;
  buffer_bytes.resize(std::min(first, buffer_bytes.size()));
}

// Column-sorted row blocks over the execution chunks (variant 4), with the
// merge lists of the rows their windows cut.
int Plan::setup_sorted(const int *h_rp) {
  const std::vector<RowBlock> hb = build_row_blocks(exec, h_rp, kSortedCap);
  int sub = opts.subwave > 0 ? opts.subwave : pow2_round(stats.mean / 4.0, 1, 64);
  if (sub < 1 || sub > 64 || (sub & (sub - 1))) {
    set_error(SBLAS_ERR_ARGS, "subwave must be a power of two in [1, 64]");
    return SBLAS_ERR_ARGS;
  }
  int remap = opts.xcd_remap;
  if (remap < 0) remap = env_int("SBLAS_XCD_REMAP", 1);
  int st = sorted_build(val_type, n, a1, val, hb, sub, remap, sorted);
  if (st != SBLAS_OK) return st;
  workspace_bytes += sorted.bytes;
  std::vector<ExecChunk> as_chunks;
  as_chunks.reserve(hb.size());
  for (const RowBlock &r : hb) as_chunks.push_back(ExecChunk{r.lo, r.hi, r.row_begin, r.row_end});
  std::vector<int> mrows, mptr, mslots;
  merge_lists(as_chunks, h_rp, mrows, mptr, mslots);
  sgroups = static_cast<int>(mrows.size());
  if ((st = upload(*this, &d_smrows, mrows)) != SBLAS_OK) return st;
  if ((st = upload(*this, &d_smptr, mptr)) != SBLAS_OK) return st;
  return upload(*this, &d_smslots, mslots);
}

// The column-sorted blocks' device arrays go (their merge lists are small
// and stay).
void Plan::drop_sorted() {
  for (void *p : sorted.buffers) (void)hipFree(p);
  workspace_bytes -= sorted.bytes;
  sorted = SortedFormat{};
  sgroups = 0;
}

// The CSR5-style format's buffers (allocated after `mark`) go.
void Plan::drop_csr5(size_t mark) {
  free_buffers_from(mark);
  d_c5_tile_nz = d_c5_flags = nullptr;
  d_c5_rows_nz = d_c5_empty = d_c5_colT = nullptr;
  d_c5_valT = nullptr;
  d_carry = nullptr;
  tiles = 0;
  sigma = 0;
  c5_empty = 0;
}

// Row blocks over P column segments (variant 6): the segment format, row
// blocks over its virtual rows (the row-block form's elements per thread, a
// sub-wave from their mean length) and the merge lists of the virtual rows
// those blocks cut; every buffer belongs to the format (seg.buffers).
int Plan::setup_segments(int P) {
  std::vector<int> h_vrp;
  int st = seg_build(val_type, m, n, nnz, a0, a1, val, P, seg, h_vrp);
  if (st != SBLAS_OK) return st;
  auto own = [&](const void *src, size_t bytes) -> void * {
    void *p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(bytes, 8)) != hipSuccess) return nullptr;
    seg.buffers.push_back(p);
    seg.bytes += static_cast<int64_t>(std::max<size_t>(bytes, 8));
    if (src && bytes && hipMemcpy(p, src, bytes, hipMemcpyHostToDevice) != hipSuccess)
      return nullptr;
    return p;
  };
  const std::vector<ExecChunk> one{ExecChunk{0, nnz, 0, seg.M}};
  const int items = sl.items > 0 ? sl.items : 8;
  const std::vector<RowBlock> vb = build_row_blocks(one, h_vrp.data(), kStreamBlock * items);
  vsl.items = items;
  vsl.subwave = opts.subwave > 0
                    ? opts.subwave
                    : pow2_round(static_cast<double>(nnz) / std::max(seg.M, 1) / 4.0, 1, 64);
  vsl.nblocks = static_cast<int>(vb.size());
  vsl.xcd_remap = 1;
  vsl.grid = ((vsl.nblocks + 7) / 8) * 8;
  std::vector<ExecChunk> as_chunks;
  as_chunks.reserve(vb.size());
  for (const RowBlock &r : vb) as_chunks.push_back(ExecChunk{r.lo, r.hi, r.row_begin, r.row_end});
  std::vector<int> mrows, mptr, mslots;
  merge_lists(as_chunks, h_vrp.data(), mrows, mptr, mslots);
  vgroups = static_cast<int>(mrows.size());
  d_vblocks = static_cast<RowBlock *>(own(vb.data(), sizeof(RowBlock) * vb.size()));
  d_vmrows = static_cast<int *>(own(mrows.data(), sizeof(int) * mrows.size()));
  d_vmptr = static_cast<int *>(own(mptr.data(), sizeof(int) * mptr.size()));
  d_vmslots = static_cast<int *>(own(mslots.data(), sizeof(int) * mslots.size()));
  d_vpartials = static_cast<double *>(own(nullptr, sizeof(double) * 2 * std::max<size_t>(vb.size(), 1)));
  workspace_bytes += seg.bytes;
  if (!d_vblocks || !d_vmrows || !d_vmptr || !d_vmslots || !d_vpartials) {
    set_error(SBLAS_ERR_ALLOC, "plan: column segment buffers");
    return SBLAS_ERR_ALLOC;
  }
  return SBLAS_OK;
}

void Plan::drop_segments() {
  for (void *p : seg.buffers) (void)hipFree(p);
  workspace_bytes -= seg.bytes;
  seg = SegmentFormat{};
  d_vblocks = nullptr;
  d_vmrows = d_vmptr = d_vmslots = nullptr;
  d_vpartials = nullptr;
  vgroups = 0;
  vsl = StreamLaunch{};
}

// Column segments for row blocks over x: about 2 MB of FP64 x each, half an
// XCD's L2, so the segment stays there beside the streamed elements
// (SBLAS_SEGMENTS overrides), at least 2.  R-MAT scale 22 (32 MB of x), all
// three launches, against 0.491 ms unsegmented (FP32 values 0.419): 8 segments
// 0.416 ms (0.364), 12 0.411, 16 0.401 (0.355), 24 0.404, 32 0.411.
int segments_for(int n) {
  const int env = env_int("SBLAS_SEGMENTS", 0);
  if (env > 0) return std::min(env, kSegmentsMax);
  const int64_t bytes = static_cast<int64_t>(n) * 8;
  return static_cast<int>(std::min<int64_t>(kSegmentsMax,
                                            std::max<int64_t>(2, (bytes + (2 << 20) - 1) >> 21)));
}

// Times forms (kernel, variant) of the plan's own matrix with x = 0 and a
// scratch y: launches cycle through the forms so a clock ramp affects all
// alike, untimed for at least 20 ms, then five timed rounds; ms[i] is form
// i's median.  The plan's kernel and variant are left at the last form.
int Plan::trial(const std::vector<std::pair<int, int>> &forms, std::vector<float> &ms) {
  const size_t F = forms.size();
  ms.assign(F, 0.f);
  double *tx = nullptr, *ty = nullptr;
  hipStream_t ts = nullptr;
  std::vector<hipEvent_t> ev(F + 1, nullptr);
  int st = SBLAS_OK;
  auto cleanup = [&]() {
    if (ts) (void)hipStreamSynchronize(ts);
    for (hipEvent_t e : ev)
      if (e) (void)hipEventDestroy(e);
    if (ts) (void)hipStreamDestroy(ts);
    if (tx) (void)hipFree(tx);
    if (ty) (void)hipFree(ty);
  };
  if (hipMalloc(&tx, sizeof(double) * std::max(n, 1)) != hipSuccess ||
      hipMalloc(&ty, sizeof(double) * std::max(m, 1)) != hipSuccess ||
      hipStreamCreateWithFlags(&ts, hipStreamNonBlocking) != hipSuccess) {
    cleanup();
    set_error(SBLAS_ERR_ALLOC, "autotune: trial buffers");
    return SBLAS_ERR_ALLOC;
  }
  for (hipEvent_t &e : ev)
    if (hipEventCreate(&e) != hipSuccess) {
      cleanup();
      set_error(SBLAS_ERR_HIP, "autotune: event creation failed");
      return SBLAS_ERR_HIP;
    }
  if (hipMemsetAsync(tx, 0, sizeof(double) * std::max(n, 1), ts) != hipSuccess ||
      hipMemsetAsync(ty, 0, sizeof(double) * std::max(m, 1), ts) != hipSuccess) {
    cleanup();
    set_error(SBLAS_ERR_HIP, "autotune: trial buffer set-up failed");
    return SBLAS_ERR_HIP;
  }
  const bool was_timing = timing;
  timing = false;
  auto run_as = [&](const std::pair<int, int> &f) {
    kernel = f.first;
    variant = f.second;
    return run_spmm(1, 1.0, tx, 0.0, ty, ts);
  };
  const auto t0 = std::chrono::steady_clock::now();
  do {
    for (int i = 0; i < 4 && st == SBLAS_OK; i++)
      for (size_t f = 0; f < F && st == SBLAS_OK; f++) st = run_as(forms[f]);
    if (st == SBLAS_OK && hipStreamSynchronize(ts) != hipSuccess) st = SBLAS_ERR_HIP;
  } while (st == SBLAS_OK &&
           std::chrono::steady_clock::now() - t0 < std::chrono::milliseconds(20));
  std::vector<std::vector<float>> t(F);
  for (int rep = 0; rep < 5 && st == SBLAS_OK; rep++) {
    if (hipEventRecord(ev[0], ts) != hipSuccess) st = SBLAS_ERR_HIP;
    for (size_t f = 0; f < F && st == SBLAS_OK; f++) {
      st = run_as(forms[f]);
      if (st == SBLAS_OK && hipEventRecord(ev[f + 1], ts) != hipSuccess) st = SBLAS_ERR_HIP;
    }
    if (st == SBLAS_OK && hipEventSynchronize(ev[F]) != hipSuccess) st = SBLAS_ERR_HIP;
    for (size_t f = 0; f < F && st == SBLAS_OK; f++) {
      float v = 0.f;
      if (hipEventElapsedTime(&v, ev[f], ev[f + 1]) != hipSuccess) st = SBLAS_ERR_HIP;
      t[f].push_back(v);
    }
  }
  timing = was_timing;
  cleanup();
  if (st != SBLAS_OK) {
    if (st == SBLAS_ERR_HIP) set_error(SBLAS_ERR_HIP, "autotune: trial launch failed");
    return st;
  }
  for (size_t f = 0; f < F; f++) {
    std::sort(t[f].begin(), t[f].end());
    ms[f] = t[f][t[f].size() / 2];
  }
  return SBLAS_OK;
}

// Kernel choice and launch set-up for a CSR matrix whose row pointer is h_rp
// (the caller's CSR, or the row-major form built from a CSC matrix).
//
// With kernel AUTO the autotuner decides from the matrix's structure and row
// statistics: the column-panel form of kernel 1 where the column structure
// pays for it (a structural test over the per-panel column ranges, measured
// crossover; slices in row order for FP64 values, in sorted groups where x
// is staged in FP32); otherwise the row-block form of kernel 1, its
// column-sorted variant or the CSR5-style kernel 3 (tile depth sigma from the
// reference's rule), whichever a timed trial on this matrix finds fastest.
// Trials run for matrices of at least
// 2^20 elements (smaller ones take row blocks) and are cached by row profile
// (value type, mean, coefficient of variation, empty-row fraction, size),
// so later plans of the same profile reuse the decision without a trial.
// The merge path (kernel 2) is never faster than both on the measured
// matrices and runs only when asked for.  SBLAS_AUTOTUNE=0 turns trials off.
int Plan::finish_csr(const std::vector<int> &h_rp) {
  int st = choose_csr(h_rp);
  if (st != SBLAS_OK) return st;
  // narrow values in row blocks: x gathered from an FP32 copy written by each
  // product unless asked not to (narrow_x 1; SBLAS_NARROW_X overrides the
  // auto setting).  R-MAT scale 22: FP32 values 1.10x -> 1.18x the FP64 plan,
  // BF16 values 1.14x -> 1.22x, normwise error 2.2e-8.  Auto leaves matrices
  // under 2^20 elements in FP64 x: their product is mostly launch latency,
  // and the copy's own launch made the power-law n = 2e4 matrix 0.70x.
  int nxo = opts.narrow_x;
  if (nxo == 0) {
    nxo = env_int("SBLAS_NARROW_X", 0);
    if (nxo <= 0) nxo = nnz >= (int64_t{1} << 20) ? 2 : 1;
  }
  if ((variant == 2 || variant == 6) && kernel == SBLAS_KERNEL_VECTOR &&
      val_type != SBLAS_VAL_F64 && nxo == 2 && n > 0) {
    x32_rows = static_cast<float *>(dev_alloc(sizeof(float) * static_cast<size_t>(n)));
    if (!x32_rows) {
      set_error(SBLAS_ERR_ALLOC, "plan: FP32 x copy allocation failed");
      return SBLAS_ERR_ALLOC;
    }
  }
  return SBLAS_OK;
}

int Plan::choose_csr(const std::vector<int> &h_rp) {
  stats = row_stats(m, h_rp.data());
  const bool auto_kernel = opts.kernel == SBLAS_KERNEL_AUTO;
  kernel = auto_kernel ? SBLAS_KERNEL_VECTOR : opts.kernel;
  tuned = auto_kernel ? 1 : 0;
  if (kernel < SBLAS_KERNEL_VECTOR || kernel > SBLAS_KERNEL_CSR5) {
    set_error(SBLAS_ERR_KERNEL, "CSR plans take kernel 1, 2 or 3");
    return SBLAS_ERR_KERNEL;
  }
  // kernel 1 form: row blocks staged in LDS unless the sub-wave form is asked
  // for; the sub-wave form also serves multi-RHS blocks and per-chunk streams
  variant = (opts.variant == 1 || opts.streams > 1) ? 1
            : (opts.variant == 4 || opts.variant == 6) ? opts.variant
                                                        : 2;
  if (opts.variant == 0) variant = env_int("SBLAS_VARIANT", variant);
  int st;
  if ((st = build_csr_chunks(h_rp.data())) != SBLAS_OK) return st;
  if ((st = setup_vector(h_rp.data())) != SBLAS_OK) return st;
  if (kernel == SBLAS_KERNEL_MERGE) return setup_tiles();
  if (kernel == SBLAS_KERNEL_CSR5) return setup_csr5(h_rp);
  // the panel form sums each row in one thread: not for matrices with very long rows
  const bool panel_ok = opts.variant == 3 || stats.max_row <= 8192;
  if (variant == 2 && (opts.variant == 0 || opts.variant == 3) && opts.streams <= 1 && nnz > 0 &&
      panel_ok) {
    // column-panel form: forced by variant 3, else taken when it pays
    int cus = 256;
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device);
    const int min_panels = 2 * std::max(cus, 1);
    // narrow values stage x in FP32 unless asked not to (SBLAS_NARROW_X
    // overrides the auto setting)
    int nxo = opts.narrow_x;
    if (nxo == 0) nxo = env_int("SBLAS_NARROW_X", 0) > 0 ? env_int("SBLAS_NARROW_X", 0) : 2;
    const bool narrow_x = val_type != SBLAS_VAL_F64 && nxo == 2;
    // slice layout: sorted groups where x is staged in FP32, rows in order
    // otherwise, unless asked (SBLAS_PANEL_LAYOUT overrides the auto setting)
    int lo_opt = opts.panel_layout;
    if (lo_opt == 0) lo_opt = env_int("SBLAS_PANEL_LAYOUT", 0);
    const int layout = lo_opt == 2   ? kPanelSorted
                       : lo_opt == 1 ? kPanelRowOrder
                       : narrow_x    ? kPanelSorted
                                     : kPanelRowOrder;
    const int rows =
        opts.panel_rows > 0
            ? opts.panel_rows
            : panel_rows_for(val_type, m, device, narrow_x, layout, static_cast<int>(exec.size()));
    bool chosen = false;
    st = panel_build(val_type, rows, m, n, nnz, a0, a1, val, exec, h_rp.data(),
                     opts.variant == 3, min_panels, panel, &chosen, nullptr, layout);
    // sorted-layout panels sized for one round may hold slices too sparse
    // for the form: taller ones (while enough panels remain) before giving up
    for (int r2 = 2 * rows; st == SBLAS_OK && !chosen && layout == kPanelSorted &&
                            opts.panel_rows <= 0 && r2 <= 1024 && m / r2 >= min_panels;
         r2 *= 2) {
      panel = PanelFormat{};
      st = panel_build(val_type, r2, m, n, nnz, a0, a1, val, exec, h_rp.data(), false, min_panels,
                       panel, &chosen, nullptr, layout);
    }
    if (st == SBLAS_ERR_UNSUPPORTED && opts.variant == 0) {
      // a slice too large for 16-bit offsets: keep the row-block form
      for (void *p : panel.buffers) (void)hipFree(p);
      panel.buffers.clear();
      panel.npanels = 0;
      clear_error();
      chosen = false;
    } else if (st != SBLAS_OK) {
      return st;
    }
    if (chosen && narrow_x) {
      // FP32 copy of x, zero-padded to whole column panels (the pad is read
      // by the x window loads, never used)
      const int64_t padded = (static_cast<int64_t>(n) + kPanelCols - 1) / kPanelCols * kPanelCols;
      void *p = nullptr;
      if (hipMalloc(&p, sizeof(float) * std::max<int64_t>(padded, 4)) != hipSuccess) {
        set_error(SBLAS_ERR_ALLOC, "plan: FP32 x copy allocation failed");
        return SBLAS_ERR_ALLOC;
      }
      panel.buffers.push_back(p);
      workspace_bytes += static_cast<int64_t>(sizeof(float)) * std::max<int64_t>(padded, 4);
      SBLAS_HIP_TRY(hipMemset(p, 0, sizeof(float) * std::max<int64_t>(padded, 4)));
      panel.x32 = static_cast<float *>(p);
    }
    if (chosen) {
      variant = 3;
      int remap = opts.xcd_remap;
      if (remap < 0) remap = env_int("SBLAS_XCD_REMAP", 1);
      vl.xcd_remap = remap ? 1 : 0;
      // rows cut by chunk edges: partials of their panels, merged in panel order
      std::vector<int> mrows, mptr, mslots;
      merge_lists(panel.hpanels, h_rp.data(), mrows, mptr, mslots);
      pgroups = static_cast<int>(mrows.size());
      if ((st = upload(*this, &d_pmrows, mrows)) != SBLAS_OK) return st;
      if ((st = upload(*this, &d_pmptr, mptr)) != SBLAS_OK) return st;
      if ((st = upload(*this, &d_pmslots, mslots)) != SBLAS_OK) return st;
      return setup_stream(h_rp.data());  // row blocks serve multi-RHS products
    }
  }
  if (variant != 2 && variant != 4 && variant != 6) return SBLAS_OK;
  const int asked = variant;
  variant = 2;
  if ((st = setup_stream(h_rp.data())) != SBLAS_OK) return st;  // also serves multi-RHS
  if (asked == 4) {
    variant = 4;
    return setup_sorted(h_rp.data());
  }
  if (asked == 6) {
    if (nnz == 0) return SBLAS_OK;  // nothing to segment: row blocks
    if ((st = setup_segments(segments_for(n))) != SBLAS_OK) return st;
    variant = 6;
    return SBLAS_OK;
  }
  // kernel AUTO, no column panels: row blocks, column-sorted row blocks or
  // the CSR5-style form
  if (!auto_kernel || opts.variant != 0 || opts.streams > 1 || nnz < (int64_t{1} << 20))
    return SBLAS_OK;
  const TuneKey key = tune_key(val_type, stats, nnz);
  TuneResult found{};
  bool cached = false;
  {
    std::lock_guard<std::mutex> lk(g_tune_mu);
    auto it = g_tune_cache.find(key);
    if (it != g_tune_cache.end()) {
      found = it->second;
      cached = true;
    }
  }
  if (cached) {
    tuned = 3;
    for (int i = 0; i < 4; i++) trial_ms[i] = found.ms[i];
    if (found.kernel == SBLAS_KERNEL_CSR5) {
      kernel = SBLAS_KERNEL_CSR5;
      return setup_csr5(h_rp);
    }
    if (found.variant == 4) {
      variant = 4;
      return setup_sorted(h_rp.data());
    }
    if (found.variant == 6) {
      st = setup_segments(segments_for(n));
      if (st == SBLAS_ERR_ALLOC || st == SBLAS_ERR_UNSUPPORTED) {  // row blocks instead
        clear_error();
        drop_segments();
        return SBLAS_OK;
      }
      if (st != SBLAS_OK) return st;
      variant = 6;
      return SBLAS_OK;
    }
    return SBLAS_OK;
  }
  if (env_int("SBLAS_AUTOTUNE", 1) == 0) return SBLAS_OK;
  // the sorted blocks' merge lists go into device_buffers before the mark,
  // so dropping the CSR5 format (everything after the mark) leaves them
  if ((st = setup_sorted(h_rp.data())) != SBLAS_OK) return st;
  const size_t mark = device_buffers.size();
  if ((st = setup_csr5(h_rp)) != SBLAS_OK) return st;
  std::vector<std::pair<int, int>> forms{
      {SBLAS_KERNEL_VECTOR, 2}, {SBLAS_KERNEL_CSR5, 2}, {SBLAS_KERNEL_VECTOR, 4}};
  // column segments where x is several times an XCD's L2 (4 MB), unless the
  // format build's count tables (16 bytes per segment and row) would pass
  // 4 GB
  if (static_cast<int64_t>(n) * 8 >= (int64_t{16} << 20) &&
      static_cast<int64_t>(segments_for(n)) * m <= (int64_t{1} << 28)) {
    st = setup_segments(segments_for(n));
    if (st == SBLAS_ERR_ALLOC || st == SBLAS_ERR_UNSUPPORTED) {
      clear_error();
      drop_segments();
    } else if (st != SBLAS_OK) {
      return st;
    } else {
      forms.push_back({SBLAS_KERNEL_VECTOR, 6});
    }
  }
  std::vector<float> ms;
  if ((st = trial(forms, ms)) != SBLAS_OK) return st;
  tuned = 2;
  size_t best = 0;
  for (size_t f = 0; f < forms.size(); f++) {
    trial_ms[f] = ms[f];
    if (ms[f] < ms[best]) best = f;
  }
  kernel = forms[best].first;
  variant = forms[best].second;
  if (kernel != SBLAS_KERNEL_CSR5) drop_csr5(mark);
  if (variant != 4) drop_sorted();
  if (variant != 6 && !seg.buffers.empty()) drop_segments();
  std::lock_guard<std::mutex> lk(g_tune_mu);
  g_tune_cache[key] = TuneResult{kernel, variant, {trial_ms[0], trial_ms[1], trial_ms[2], trial_ms[3]}};
  return SBLAS_OK;
}

bool Plan::uses_tiles(int k) const {
  return kernel == SBLAS_KERNEL_COO ||
         ((kernel == SBLAS_KERNEL_MERGE || kernel == SBLAS_KERNEL_CSR5) && k == 1);
}

bool Plan::has_merge(int k) const {
  if (uses_tiles(k)) return tiles > 0;
  if (variant == 6 && k == 1 && streams.empty()) return false;  // merges run inside the product
  if (variant == 3 && k == 1 && streams.empty()) return pgroups > 0;
  if (variant == 4 && k == 1 && streams.empty()) return sgroups > 0;
  if (variant >= 2 && streams.empty())  // row blocks (any k)
    return (k > 1 && slk.grid > 0) ? kgroups > 0 : bgroups > 0;
  return ngroups > 0;
}

int Plan::run_merge(int k, double alpha, double beta, double *Y, hipStream_t s) {
  if (variant == 6 && k == 1 && streams.empty()) return SBLAS_OK;
  if (uses_tiles(k))
    return tiles > 0 ? launch_tile_fixup(carry_count(), d_carry, Y, alpha, beta, s) : SBLAS_OK;
  if (variant == 3 && k == 1 && streams.empty())
    return pgroups > 0 ? launch_merge(d_pmrows, d_pmptr, d_pmslots, pgroups, panel.partials, 1, Y,
                                      alpha, beta, s)
                       : SBLAS_OK;
  if (variant == 4 && k == 1 && streams.empty())
    return sgroups > 0 ? launch_merge(d_smrows, d_smptr, d_smslots, sgroups, sorted.partials, 1, Y,
                                      alpha, beta, s)
                       : SBLAS_OK;
  if (variant >= 2 && streams.empty() && k > 1 && slk.grid > 0)
    return kgroups > 0 ? launch_merge(d_kmrows, d_kmptr, d_kmslots, kgroups, d_kpartials, k, Y,
                                      alpha, beta, s)
                       : SBLAS_OK;
  if (variant >= 2 && streams.empty())
    return bgroups > 0 ? launch_merge(d_bmrows, d_bmptr, d_bmslots, bgroups, d_bpartials, k, Y,
                                      alpha, beta, s)
                       : SBLAS_OK;
  return ngroups > 0
             ? launch_merge(d_mrows, d_mptr, d_mslots, ngroups, d_partials, k, Y, alpha, beta, s)
             : SBLAS_OK;
}

}  // namespace sblas

// ===========================================================================
// C API
// ===========================================================================
using sblas::Plan;

extern "C" {

int sblas_last_error(void) { return sblas::last_error(); }

const char *sblas_error_string(int status) {
  switch (status) {
    case SBLAS_OK: return "ok";
    case SBLAS_ERR_ARGS: return "invalid arguments";
    case SBLAS_ERR_NGPU: return "ngpu must be 1";
    case SBLAS_ERR_KERNEL: return "kernel must be 1, 2 or 3";
    case SBLAS_ERR_UNDEFINED: return "reference arithmetic undefined for this input";
    case SBLAS_ERR_HIP: return "HIP runtime error";
    case SBLAS_ERR_ALLOC: return "allocation failed";
    case SBLAS_ERR_UNSUPPORTED: return "unsupported entry point";
    case SBLAS_ERR_NO_DEVICE: return "no HIP device";
    default: return "unknown status";
  }
}

const char *sblas_version(void) { return "sblas-mi355x 0.1 (gfx950)"; }

void sblas_opts_default(sblas_opts *o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->kernel = SBLAS_KERNEL_AUTO;
  o->chunks = 0;
  o->part_mode = SBLAS_PART_V1;
  o->nb = 0;
  o->streams = 1;
  o->subwave = 0;
  o->sigma = 0;
  o->xcd_remap = -1;
}

static int plan_common(int format, int m, int n, int64_t nnz, const int *a0, const int *a1,
                       const void *val, int val_type, const sblas_opts *opts,
                       sblas_plan *out) {
  sblas::clear_error();
  if (!out) {
    sblas::set_error(SBLAS_ERR_ARGS, "plan: null output handle");
    return SBLAS_ERR_ARGS;
  }
  *out = nullptr;
  if (m < 0 || n < 0 || nnz < 0 || nnz > INT_MAX - 4096 || !a0 || (nnz > 0 && (!a1 || !val)) ||
      val_type < SBLAS_VAL_F64 || val_type > SBLAS_VAL_BF16) {
    sblas::set_error(SBLAS_ERR_ARGS, "plan: invalid sizes, pointers or value type");
    return SBLAS_ERR_ARGS;
  }
  sblas_plan h = new (std::nothrow) sblas_plan_s;
  if (!h) {
    sblas::set_error(SBLAS_ERR_ALLOC, "plan: host allocation failed");
    return SBLAS_ERR_ALLOC;
  }
  Plan &p = h->impl;
  p.format = format;
  p.val_type = val_type;
  p.m = m;
  p.n = n;
  p.nnz = nnz;
  p.a0 = a0;
  p.a1 = a1;
  p.val = val;
  if (opts) p.opts = *opts; else sblas_opts_default(&p.opts);
  if (hipGetDevice(&p.device) != hipSuccess) {
    delete h;
    sblas::set_error(SBLAS_ERR_NO_DEVICE, "plan: no HIP device");
    return SBLAS_ERR_NO_DEVICE;
  }
  // chunks = 0: SBLAS_CHUNKS when set, else the autotuner's rule -- one chunk
  // for a single launch (its XCD-contiguous block order already gives each
  // XCD a contiguous eighth of the rows, the role the reference's per-GPU
  // chunks play), one chunk per stream when chunks run over several streams
  int C = p.opts.chunks > 0 ? p.opts.chunks : sblas::env_int("SBLAS_CHUNKS", 0);
  if (C <= 0) C = std::max(p.opts.streams, 1);
  p.chunks = std::max(1, C);
  *out = h;
  return SBLAS_OK;
}

int sblas_plan_create_csr(int m, int n, int64_t nnz, const int *d_row_ptr, const int *d_col,
                          const void *d_val, int val_type, const sblas_opts *opts,
                          sblas_plan *plan) {
  int st = plan_common(SBLAS_FMT_CSR, m, n, nnz, d_row_ptr, d_col, d_val, val_type, opts, plan);
  if (st != SBLAS_OK) return st;
  Plan &p = (*plan)->impl;
  std::vector<int> h_rp(static_cast<size_t>(m) + 1);
  if (hipMemcpy(h_rp.data(), d_row_ptr, sizeof(int) * h_rp.size(), hipMemcpyDeviceToHost) !=
      hipSuccess) {
    sblas::set_error(SBLAS_ERR_HIP, "plan: row pointer download failed");
    delete *plan;
    *plan = nullptr;
    return SBLAS_ERR_HIP;
  }
  auto fail = [&](int code) {
    delete *plan;
    *plan = nullptr;
    return code;
  };
  for (int r = 0; r < m; r++)
    if (h_rp[r + 1] < h_rp[r]) {
      sblas::set_error(SBLAS_ERR_ARGS, "plan: row_ptr must be non-decreasing");
      return fail(SBLAS_ERR_ARGS);
    }
  if (h_rp[0] != 0 || h_rp[m] != nnz) {
    sblas::set_error(SBLAS_ERR_ARGS, "plan: row_ptr[0] must be 0 and row_ptr[m] == nnz");
    return fail(SBLAS_ERR_ARGS);
  }
  bool ok = true;
  if ((st = sblas::check_indices(d_col, nnz, 0, n, false, nullptr, &ok)) != SBLAS_OK)
    return fail(st);
  if (!ok) {
    sblas::set_error(SBLAS_ERR_ARGS, "plan: a column index is outside [0, n)");
    return fail(SBLAS_ERR_ARGS);
  }
  if ((st = p.finish_csr(h_rp)) != SBLAS_OK) return fail(st);
  return SBLAS_OK;
}

int sblas_plan_create_coo(int m, int n, int64_t nnz, const int *d_row_idx, const int *d_col_idx,
                          const void *d_val, int val_type, const sblas_opts *opts,
                          sblas_plan *plan) {
  int st = plan_common(SBLAS_FMT_COO, m, n, nnz, d_row_idx, d_col_idx, d_val, val_type, opts,
                       plan);
  if (st != SBLAS_OK) return st;
  Plan &p = (*plan)->impl;
  auto fail = [&](int code) {
    delete *plan;
    *plan = nullptr;
    return code;
  };
  bool ok_r = true, ok_c = true;
  if ((st = sblas::check_indices(d_row_idx, nnz, 0, m, true, nullptr, &ok_r)) != SBLAS_OK ||
      (st = sblas::check_indices(d_col_idx, nnz, 0, n, false, nullptr, &ok_c)) != SBLAS_OK)
    return fail(st);
  if (!ok_r || !ok_c) {
    sblas::set_error(SBLAS_ERR_ARGS,
                     "plan: COO row indices must be sorted in [0, m) and columns in [0, n)");
    return fail(SBLAS_ERR_ARGS);
  }
  if (p.opts.kernel == SBLAS_KERNEL_COO) {
    // the segmented reduction over the row-sorted COO arrays as given
    p.kernel = SBLAS_KERNEL_COO;
    // descriptors of the COO rules need the row indices on the host:
    // sblas_partition_v1_numa_coo / sblas_partition_baseline_coo give them
    p.desc_status = SBLAS_ERR_UNSUPPORTED;
    p.stats.mean = m > 0 ? static_cast<double>(nnz) / m : 0.0;
    if ((st = p.setup_tiles()) != SBLAS_OK) return fail(st);
    return SBLAS_OK;
  }
  // the reference's COO path (coo2csr, src/spmv_helper.cu:478-491, then the
  // CSR kernel, src/dspmv_mgpu_v1_numa_coo.cu:460-476): row pointers built on
  // the device from the sorted row indices, then the CSR forms over the
  // caller's column and value arrays (the row indices are not read again)
  p.d_csr_rp = static_cast<int *>(p.dev_alloc(sizeof(int) * (static_cast<size_t>(m) + 1)));
  if (!p.d_csr_rp) {
    sblas::set_error(SBLAS_ERR_ALLOC, "plan: COO row pointer allocation failed");
    return fail(SBLAS_ERR_ALLOC);
  }
  if ((st = sblas::ptr_from_sorted_device(m, nnz, d_row_idx, p.d_csr_rp, nullptr)) != SBLAS_OK)
    return fail(st);
  std::vector<int> h_rp(static_cast<size_t>(m) + 1);
  if (hipMemcpy(h_rp.data(), p.d_csr_rp, sizeof(int) * h_rp.size(), hipMemcpyDeviceToHost) !=
      hipSuccess) {
    sblas::set_error(SBLAS_ERR_HIP, "plan: row pointer download failed");
    return fail(SBLAS_ERR_HIP);
  }
  p.a0 = p.d_csr_rp;
  if ((st = p.finish_csr(h_rp)) != SBLAS_OK) return fail(st);
  p.desc_status = SBLAS_ERR_UNSUPPORTED;  // the COO rules: see above
  p.descs.clear();
  return SBLAS_OK;
}

int sblas_plan_create_csc(int m, int n, int64_t nnz, const int *d_col_ptr, const int *d_row_idx,
                          const void *d_val, int val_type, const sblas_opts *opts,
                          sblas_plan *plan) {
  int st = plan_common(SBLAS_FMT_CSC, m, n, nnz, d_col_ptr, d_row_idx, d_val, val_type, opts,
                       plan);
  if (st != SBLAS_OK) return st;
  Plan &p = (*plan)->impl;
  auto fail = [&](int code) {
    delete *plan;
    *plan = nullptr;
    return code;
  };
  std::vector<int> h_cp(static_cast<size_t>(n) + 1);
  if (hipMemcpy(h_cp.data(), d_col_ptr, sizeof(int) * h_cp.size(), hipMemcpyDeviceToHost) !=
      hipSuccess) {
    sblas::set_error(SBLAS_ERR_HIP, "plan: column pointer download failed");
    return fail(SBLAS_ERR_HIP);
  }
  for (int c = 0; c < n; c++)
    if (h_cp[c + 1] < h_cp[c]) {
      sblas::set_error(SBLAS_ERR_ARGS, "plan: col_ptr must be non-decreasing");
      return fail(SBLAS_ERR_ARGS);
    }
  if (h_cp[0] != 0 || h_cp[n] != nnz) {
    sblas::set_error(SBLAS_ERR_ARGS, "plan: col_ptr[0] must be 0 and col_ptr[n] == nnz");
    return fail(SBLAS_ERR_ARGS);
  }
  bool ok = true;
  if ((st = sblas::check_indices(d_row_idx, nnz, 0, m, false, nullptr, &ok)) != SBLAS_OK)
    return fail(st);
  if (!ok) {
    sblas::set_error(SBLAS_ERR_ARGS, "plan: a row index is outside [0, m)");
    return fail(SBLAS_ERR_ARGS);
  }
  // column-partition descriptors of the reference's CSC rules
  {
    const int C = p.chunks;
    p.descs.assign(C, sblas_chunk_desc{});
    p.desc_status = (p.opts.part_mode == SBLAS_PART_ROWBLOCK)
                        ? sblas::partition_rowblock(n, h_cp.data(), C, p.descs.data())
                        : (nnz > 0 ? sblas::partition_v1_numa_csc(n, h_cp.data(), nnz, C, nullptr,
                                                                   p.descs.data(), nullptr)
                                   : SBLAS_ERR_UNDEFINED);
  }
  // row-major form, built once on the device
  const size_t vbytes = val_type == SBLAS_VAL_F64 ? 8 : (val_type == SBLAS_VAL_F32 ? 4 : 2);
  p.d_csr_rp = static_cast<int *>(p.dev_alloc(sizeof(int) * (static_cast<size_t>(m) + 1)));
  p.d_csr_col = static_cast<int *>(p.dev_alloc(sizeof(int) * std::max<int64_t>(nnz, 1)));
  p.d_csr_val = p.dev_alloc(vbytes * std::max<int64_t>(nnz, 1));
  if (!p.d_csr_rp || !p.d_csr_col || !p.d_csr_val) {
    sblas::set_error(SBLAS_ERR_ALLOC, "plan: CSC transpose allocation failed");
    return fail(SBLAS_ERR_ALLOC);
  }
  if ((st = sblas::csc_to_csr_device(m, n, static_cast<int>(nnz), d_col_ptr, d_row_idx, d_val,
                                     val_type, p.d_csr_rp, p.d_csr_col, p.d_csr_val, nullptr)) !=
      SBLAS_OK)
    return fail(st);
  std::vector<int> h_rp(static_cast<size_t>(m) + 1);
  if (hipMemcpy(h_rp.data(), p.d_csr_rp, sizeof(int) * h_rp.size(), hipMemcpyDeviceToHost) !=
      hipSuccess) {
    sblas::set_error(SBLAS_ERR_HIP, "plan: row pointer download failed");
    return fail(SBLAS_ERR_HIP);
  }
  p.a0 = p.d_csr_rp;
  p.a1 = p.d_csr_col;
  p.val = p.d_csr_val;
  const std::vector<sblas_chunk_desc> col_descs = p.descs;
  const int col_status = p.desc_status;
  sblas_opts o = p.opts;
  p.opts.part_mode = SBLAS_PART_V1;  // the row-major form runs row chunks
  if ((st = p.finish_csr(h_rp)) != SBLAS_OK) return fail(st);
  p.opts = o;
  p.descs = col_descs;
  p.desc_status = col_status;
  return SBLAS_OK;
}

int sblas_plan_spmv(sblas_plan plan, double alpha, const double *d_x, double beta, double *d_y,
                    void *stream) {
  sblas::clear_error();
  if (!plan || (!d_y && plan->impl.m > 0) || (!d_x && plan->impl.nnz > 0)) {
    sblas::set_error(SBLAS_ERR_ARGS, "spmv: null plan or vector");
    return SBLAS_ERR_ARGS;
  }
  return plan->impl.run_spmv(alpha, d_x, beta, d_y, static_cast<hipStream_t>(stream));
}

int sblas_plan_spmm(sblas_plan plan, int k, double alpha, const double *d_X, double beta,
                    double *d_Y, void *stream) {
  sblas::clear_error();
  if (!plan || (!d_Y && plan->impl.m > 0) || (!d_X && plan->impl.nnz > 0)) {
    sblas::set_error(SBLAS_ERR_ARGS, "spmm: null plan or block");
    return SBLAS_ERR_ARGS;
  }
  return plan->impl.run_spmm(k, alpha, d_X, beta, d_Y, static_cast<hipStream_t>(stream));
}

int sblas_plan_get_info(sblas_plan plan, sblas_plan_info *info) {
  if (!plan || !info) return SBLAS_ERR_ARGS;
  const Plan &p = plan->impl;
  std::memset(info, 0, sizeof(*info));
  info->format = p.format;
  info->val_type = p.val_type;
  info->kernel = p.kernel;
  info->chunks = static_cast<int>(p.exec.size());
  info->subwave = p.vl.subwave;
  info->sigma = p.sigma;
  info->xcd_remap = p.vl.xcd_remap;
  info->split_rows = p.split_rows;
  info->m = p.m;
  info->n = p.n;
  info->nnz = p.nnz;
  info->workspace_bytes = p.workspace_bytes;
  info->mean_row = p.stats.mean;
  info->cv_row = p.stats.cv;
  info->max_row = p.stats.max_row;
  info->launches = p.launches;
  info->variant = p.variant;
  info->row_blocks = static_cast<int>(p.rblocks.size());
  info->items = p.sl.items;
  info->panel_rows = p.variant == 3 ? p.panel.rows : 0;
  info->panel_slices = p.variant == 3 ? p.panel.nslices : 0;
  info->tuned = p.tuned;
  for (int i = 0; i < 3; i++) info->trial_ms[i] = p.trial_ms[i];
  info->panel_layout = p.variant == 3 ? (p.panel.layout == sblas::kPanelSorted ? 2 : 1) : 0;
  info->narrow_x = (p.variant == 3 && p.panel.x32) || p.x32_rows ? 1 : 0;
  info->panel_elements = p.variant == 3 ? p.panel.padded : 0;
  if (p.variant == 2) info->subwave = p.sl.subwave;
  if (p.variant == 6) {
    info->subwave = p.vsl.subwave;
    info->row_blocks = p.vsl.nblocks;
  }
  info->segments = p.variant == 6 ? p.seg.P : 0;
  info->virtual_rows = p.variant == 6 ? p.seg.M : 0;
  info->trial_seg_ms = p.trial_ms[3];
  return SBLAS_OK;
}

int sblas_plan_get_chunks(sblas_plan plan, sblas_chunk_desc *out, int cap) {
  if (!plan) return -SBLAS_ERR_ARGS;
  const Plan &p = plan->impl;
  if (p.desc_status != SBLAS_OK) return -p.desc_status;
  const int n = static_cast<int>(p.descs.size());
  if (out) std::memcpy(out, p.descs.data(), sizeof(sblas_chunk_desc) * std::min(n, cap));
  return n;
}

// Host-only view of the row blocks a plan would build (not in sblas.h; used
// by the CPU tests to check the block invariants).  out holds 4 ints per
// block: row_begin, row_end, lo, hi.  Returns the block count.
int sblas_internal_row_blocks(int m, const int *rp, int64_t nnz, int C, int part_mode,
                              int items, int *out, int cap) {
  const std::vector<int64_t> bounds = sblas::element_bounds(part_mode, m, rp, nnz, C, 1);
  const std::vector<sblas::ExecChunk> ex = sblas::exec_chunks_from_bounds(m, rp, bounds);
  const std::vector<sblas::RowBlock> rb =
      sblas::build_row_blocks(ex, rp, sblas::kStreamBlock * items);
  const int n = static_cast<int>(rb.size());
  for (int i = 0; i < std::min(n, cap); i++) {
    out[4 * i] = rb[i].row_begin;
    out[4 * i + 1] = rb[i].row_end;
    out[4 * i + 2] = rb[i].lo;
    out[4 * i + 3] = rb[i].hi;
  }
  return n;
}

// Host-only view of the CSR5-style tile depth rule (not in sblas.h; used by
// the CPU tests).
int sblas_internal_autotune_sigma(int64_t nnz, int m) { return sblas::autotune_sigma(nnz, m); }

int sblas_plan_set_timing(sblas_plan plan, int on) {
  if (!plan) return SBLAS_ERR_ARGS;
  plan->impl.timing = on != 0;
  plan->impl.tcount = 0;
  return SBLAS_OK;
}

int sblas_plan_kernel_times(sblas_plan plan, float *ms, int cap) {
  if (!plan) return -SBLAS_ERR_ARGS;
  Plan &p = plan->impl;
  const int n = std::min(p.tcount, std::max(cap, 0));
  for (int i = 0; i < n; i++) {
    if (hipEventSynchronize(p.tev[2 * i + 1]) != hipSuccess) return -SBLAS_ERR_HIP;
    float t = 0.f;
    if (hipEventElapsedTime(&t, p.tev[2 * i], p.tev[2 * i + 1]) != hipSuccess)
      return -SBLAS_ERR_HIP;
    if (ms) ms[i] = t;
  }
  p.tcount = 0;
  return n;
}

int sblas_plan_destroy(sblas_plan plan) {
  delete plan;
  return SBLAS_OK;
}

}  // extern "C"
