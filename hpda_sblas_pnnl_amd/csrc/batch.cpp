// batch.cpp -- batched SpMV over many CSR matrices in one persistent launch,
// with a per-matrix autotuner and hipGraph replay.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <new>

#include "plan.hpp"

namespace sblas {
namespace {

// Autotuner decision for one row-length profile (the row form).  Lanes per
// row follow the mean row length (a quarter of it, so a sub-wave strides each
// row about four times through LDS); matrices above 2^24 elements are cut
// into 8 nnz chunks (v1 rule) before packing row blocks; smaller ones keep
// one chunk.  Decisions are cached by the rounded profile so equal-shaped
// matrices agree.  Whether a matrix takes the column-panel form instead is
// decided from its column structure when the batch is created.
struct Choice {
  int subwave = 16, items = 8, chunks = 1;
};

int env_int_b(const char *name, int fallback) {
  const char *e = std::getenv(name);
  return (e && *e) ? std::atoi(e) : fallback;
}

int pow2_round_b(double v, int lo, int hi) {
  if (!(v > 0)) return lo;
  const int p = 1 << static_cast<int>(std::lround(std::log2(v)));
  return std::min(std::max(p, lo), hi);
}

Choice decide(const RowStats &s, int64_t nnz) {
  static std::mutex mu;
  static std::map<std::pair<int, int>, Choice> cache;
  const std::pair<int, int> key{static_cast<int>(std::lround(std::log2(std::max(1.0, s.mean)) * 4)),
                                nnz > (int64_t{1} << 24) ? 1 : 0};
  std::lock_guard<std::mutex> lock(mu);
  auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  Choice c;
  c.subwave = pow2_round_b(s.mean / 4.0, 1, 64);
  c.items = 8;
  c.chunks = key.second ? 8 : 1;
  cache[key] = c;
  return c;
}

// Appends a vector of n entries to a list for the FP32 conversion: its chunks
// of kX32Chunk entries follow the list's chunks so far.
void add_x32(std::vector<sblas::BatchX32> &segs, std::vector<int> &chunks, const double *x,
             float *x32, int64_t n) {
  const int seg = static_cast<int>(segs.size());
  segs.push_back(sblas::BatchX32{x, x32, n, static_cast<int>(chunks.size())});
  const int64_t count = (n + sblas::kX32Chunk - 1) / sblas::kX32Chunk;
  chunks.insert(chunks.end(), static_cast<size_t>(count), seg);
}

}  // namespace
}  // namespace sblas

struct sblas_batch_s {
  std::vector<sblas::BatchMatrix> mats;
  std::vector<sblas::BatchItem> items;
  std::vector<sblas_batch_choice> choices;
  std::vector<sblas::PanelFormat> panels;  // column-panel forms, one per matrix (may be empty)
  sblas::BatchMatrix *d_mats = nullptr;
  sblas::BatchItem *d_items = nullptr;
  double *d_partials = nullptr;
  int *d_ctr = nullptr;
  int *d_qbound = nullptr;
  int *d_gmat = nullptr, *d_grow = nullptr, *d_gptr = nullptr, *d_gslot = nullptr;
  int groups = 0;
  // narrow-x matrices in column-panel form (sorted-group layout, FP32 copy
  // of x): each runs as its own launch of the plan's panel kernel, its row
  // panels sized so that they are all resident at once
  std::vector<int> solo;
  // FP32 copies of x for row blocks over narrow values, written by one
  // launch ahead of the row blocks
  std::vector<sblas::BatchX32> xsegs;
  sblas::BatchX32 *d_xsegs = nullptr;
  std::vector<int> xchunks;  // segment of each chunk of the conversion
  int *d_xchunks = nullptr;
  // and those of the matrices with launches of their own (one launch ahead
  // of theirs)
  std::vector<sblas::BatchX32> solo_xsegs;
  sblas::BatchX32 *d_solo_xsegs = nullptr;
  std::vector<int> solo_xchunks;
  int *d_solo_xchunks = nullptr;
  // item list: [0, n_panel) row panels in row order, then row blocks
  int n_panel = 0;
  int grid_panel = 0, grid_rows = 0;
  int device = 0;
  int items_per_thread = 8;
  std::vector<void *> bufs;
  hipStream_t cap_stream = nullptr;
  // the row-order panels and the row blocks run beside the sorted panels
  hipStream_t side = nullptr, side2 = nullptr;
  hipEvent_t fork = nullptr, join = nullptr, join2 = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  ~sblas_batch_s() {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (cap_stream) (void)hipStreamDestroy(cap_stream);
    if (side) (void)hipStreamDestroy(side);
    if (side2) (void)hipStreamDestroy(side2);
    if (fork) (void)hipEventDestroy(fork);
    if (join) (void)hipEventDestroy(join);
    if (join2) (void)hipEventDestroy(join2);
    for (void *p : bufs) (void)hipFree(p);
    for (sblas::PanelFormat &f : panels)
      for (void *p : f.buffers) (void)hipFree(p);
  }
  template <typename T>
  int upload(T **dst, const std::vector<T> &src) {
    void *p = nullptr;
    const size_t bytes = sizeof(T) * std::max<size_t>(src.size(), 1);
    if (hipMalloc(&p, bytes) != hipSuccess) {
      sblas::set_error(SBLAS_ERR_ALLOC, "batch: device allocation failed");
      return SBLAS_ERR_ALLOC;
    }
    bufs.push_back(p);
    *dst = static_cast<T *>(p);
    if (!src.empty() &&
        hipMemcpy(p, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice) != hipSuccess) {
      sblas::set_error(SBLAS_ERR_HIP, "batch: upload failed");
      return SBLAS_ERR_HIP;
    }
    return SBLAS_OK;
  }
  int enqueue(hipStream_t s) {
    // the matrices with their own launches on s, one after another; the
    // row-order panels and the row blocks each in one persistent launch with
    // its own queues (qbound) and counters (ctr), on side streams forked
    // from s and joined back before the merge, so the launches share the GPU
    // instead of queueing
    constexpr int Q = sblas::kBatchQueues + 1;
    const int n_rows = static_cast<int>(items.size()) - n_panel;
    const int used = (solo.empty() ? 0 : 1) + (grid_panel > 0) + (grid_rows > 0);
    // with one kind only, it runs on s itself
    hipStream_t ps = (used > 1 && !solo.empty()) ? side2 : s;
    hipStream_t rs = (used > 1) ? side : s;
    // the FP32 copies of x of the matrices with launches of their own come
    // first, alone on the GPU (0.04 ms for the 64-matrix FP32 batch): behind
    // the fork they shared the GPU with the persistent panel launch and the
    // first of those launches began only after 0.23 ms
    int st = sblas::launch_batch_x32(d_solo_xsegs, d_solo_xchunks,
                                     static_cast<int>(solo_xchunks.size()), s);
    if (st != SBLAS_OK) return st;
    if (used > 1) {
      SBLAS_HIP_TRY(hipEventRecord(fork, s));
      if (grid_panel > 0 && ps != s) SBLAS_HIP_TRY(hipStreamWaitEvent(ps, fork, 0));
      if (grid_rows > 0 && rs != s) SBLAS_HIP_TRY(hipStreamWaitEvent(rs, fork, 0));
    }
    // beside the matrices with launches of their own, the row blocks run one
    // item per workgroup, so CUs pass to those launches as items finish (a
    // persistent launch held every CU for 1.9 ms while they waited: 64-matrix
    // FP32 batch 4.66 ms, 4.49 ms this way); SBLAS_BATCH_FLAT overrides (0
    // persistent, 1 flat row blocks, 2 flat row panels too)
    int flat = sblas::env_int_b("SBLAS_BATCH_FLAT", -1);
    if (flat < 0) flat = solo.empty() ? 0 : 1;
    st = sblas::launch_batch(sblas::kBatchPanel, items_per_thread, grid_panel, n_panel,
                                 d_items, d_mats, d_partials, d_qbound, d_ctr, ps, flat >= 2);
    if (st != SBLAS_OK) return st;
    st = sblas::launch_batch_x32(d_xsegs, d_xchunks, static_cast<int>(xchunks.size()), rs);
    if (st != SBLAS_OK) return st;
    st = sblas::launch_batch(sblas::kBatchRowBlock, items_per_thread, grid_rows, n_rows,
                             d_items + n_panel, d_mats, d_partials, d_qbound + Q, d_ctr + Q, rs,
                             flat >= 1);
    if (st != SBLAS_OK) return st;
    for (int i : solo) {  // one launch each
      const sblas::BatchMatrix &M = mats[i];
      st = sblas::launch_panel_spmv(M.val_type, panels[i], M.m, M.n, 1, M.x, M.y, M.alpha, M.beta,
                                    s, true);
      if (st != SBLAS_OK) return st;
    }
    if (used > 1) {
      if (grid_panel > 0 && ps != s) {
        SBLAS_HIP_TRY(hipEventRecord(join2, ps));
        SBLAS_HIP_TRY(hipStreamWaitEvent(s, join2, 0));
      }
      if (grid_rows > 0 && rs != s) {
        SBLAS_HIP_TRY(hipEventRecord(join, rs));
        SBLAS_HIP_TRY(hipStreamWaitEvent(s, join, 0));
      }
    }
    return sblas::launch_batch_merge(groups, d_gmat, d_grow, d_gptr, d_gslot, d_mats, d_partials,
                                     s);
  }
};

extern "C" {

int sblas_batch_create(int count, const sblas_batch_item *items, const sblas_opts *opts,
                       sblas_batch *out) {
  using namespace sblas;
  clear_error();
  if (!out || count < 0 || (count > 0 && !items)) {
    set_error(SBLAS_ERR_ARGS, "batch: bad arguments");
    return SBLAS_ERR_ARGS;
  }
  *out = nullptr;
  sblas_batch b = new (std::nothrow) sblas_batch_s;
  if (!b) {
    set_error(SBLAS_ERR_ALLOC, "batch: host allocation failed");
    return SBLAS_ERR_ALLOC;
  }
  if (hipGetDevice(&b->device) != hipSuccess) b->device = 0;
  int cus = 256;
  (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device);
  cus = std::max(cus, 1);
  auto fail = [&](int code) {
    delete b;
    return code;
  };
  std::vector<int> gmat, grow, gptr(1, 0), gslot;
  // work items of each matrix in its own order
  struct Pending {
    BatchItem item;
    int64_t work;        // elements
    int lo, hi;          // element window
    int row_begin, row_end;
  };
  std::vector<std::vector<Pending>> pending(count);
  std::vector<std::vector<int>> h_rps(count);
  for (int i = 0; i < count; i++) {
    const sblas_batch_item &it = items[i];
    if (it.m < 0 || it.n < 0 || it.nnz < 0 || it.nnz > INT32_MAX - 4096 || !it.d_row_ptr ||
        (it.nnz > 0 && (!it.d_col || !it.d_val || !it.d_x)) || (it.m > 0 && !it.d_y) ||
        it.val_type < SBLAS_VAL_F64 || it.val_type > SBLAS_VAL_BF16 ||
        !(it.k == 0 || it.k == 1 || it.k == 2 || it.k == 4 || it.k == 8 || it.k == 16)) {
      set_error(SBLAS_ERR_ARGS, "batch: bad item " + std::to_string(i));
      return fail(SBLAS_ERR_ARGS);
    }
    const int kv = it.k <= 1 ? 1 : it.k;
    std::vector<int> h_rp(static_cast<size_t>(it.m) + 1);
    if (hipMemcpy(h_rp.data(), it.d_row_ptr, sizeof(int) * h_rp.size(), hipMemcpyDeviceToHost) !=
        hipSuccess) {
      set_error(SBLAS_ERR_HIP, "batch: row pointer download failed");
      return fail(SBLAS_ERR_HIP);
    }
    if (h_rp[0] != 0 || h_rp[it.m] != it.nnz) {
      set_error(SBLAS_ERR_ARGS, "batch: row_ptr[0] must be 0 and row_ptr[m] == nnz");
      return fail(SBLAS_ERR_ARGS);
    }
    for (int r = 0; r < it.m; r++)
      if (h_rp[r + 1] < h_rp[r]) {
        set_error(SBLAS_ERR_ARGS, "batch: row_ptr must be non-decreasing");
        return fail(SBLAS_ERR_ARGS);
      }
    bool ok = true;
    int st = check_indices(it.d_col, it.nnz, 0, it.n, false, nullptr, &ok);
    if (st != SBLAS_OK) return fail(st);
    if (!ok) {
      set_error(SBLAS_ERR_ARGS, "batch: column index outside [0, n) in item " + std::to_string(i));
      return fail(SBLAS_ERR_ARGS);
    }
    const RowStats rs = row_stats(it.m, h_rp.data());
    Choice c = decide(rs, it.nnz);
    if (opts && opts->subwave > 0) c.subwave = opts->subwave;
    if (opts && opts->chunks > 0) c.chunks = opts->chunks;
    b->items_per_thread = (opts && opts->items == 16) ? 16 : 8;
    c.items = b->items_per_thread;

    BatchMatrix M{};
    M.rp = it.d_row_ptr;
    M.col = it.d_col;
    M.val = it.d_val;
    M.x = it.d_x;
    M.y = it.d_y;
    M.alpha = it.alpha;
    M.beta = it.beta;
    M.val_type = it.val_type;
    M.subwave = c.subwave;
    M.m = it.m;
    M.n = it.n;
    M.k = kv;

    // column-panel form when the matrix's column structure pays for it (one
    // chunk: rows are never cut; 16-bit slice offsets only)
    const int want = opts ? opts->variant : 0;
    b->panels.emplace_back();
    PanelFormat &f = b->panels.back();
    bool panel = false;
    // narrow values: x staged from an FP32 copy, the sorted-group layout and
    // a launch of its own (the plan's kernel, row panels sized so that all
    // are resident at once), unless asked not to.  In the persistent launch
    // a workgroup takes many panels one after another, and the sorted body,
    // which waits for its loads slice by slice, was slower there (64-matrix
    // FP32 batch: 5.84 ms against 4.67 ms staged in FP64).
    int nxo = opts ? opts->narrow_x : 0;
    if (nxo == 0) nxo = env_int_b("SBLAS_NARROW_X", 0) > 0 ? env_int_b("SBLAS_NARROW_X", 0) : 2;
    bool narrow = it.val_type != SBLAS_VAL_F64 && nxo == 2;
    // a launch of its own only for a matrix whose row panels give every CU
    // a workgroup (smaller ones left a few workgroups running for
    // milliseconds); the others stage x in FP64 in the persistent launch
    const int solo_rows =
        narrow ? panel_rows_for(it.val_type, std::max(it.m, 1), b->device, true, kPanelSorted, 1)
               : 0;
    if (narrow && !(opts && opts->panel_rows > 0) &&
        (static_cast<int64_t>(it.m) + solo_rows - 1) / solo_rows < cus)
      narrow = false;
    if (kv == 1 && it.m > 0 && it.nnz > 0 && want != 1 && want != 2 &&
        (want == 3 || rs.max_row <= 8192)) {
      const std::vector<ExecChunk> one{ExecChunk{0, it.nnz, 0, it.m}};
      int rows = (opts && opts->panel_rows > 0) ? std::min(opts->panel_rows, kBatchPanelRowsMax)
                 : narrow ? solo_rows
                          : kBatchPanelRowsMax;
      int layout = narrow ? kPanelSorted : kPanelRowOrder;
      int st = panel_build(it.val_type, rows, it.m, it.n, it.nnz, it.d_row_ptr, it.d_col,
                           it.d_val, one, h_rp.data(), want == 3, 1, f, &panel, nullptr, layout);
      // panels sized for one round may hold slices too sparse for the form:
      // taller ones while every CU still gets one
      while (st == SBLAS_OK && !panel && narrow && rows < 1024 && !(opts && opts->panel_rows > 0) &&
             (static_cast<int64_t>(it.m) + 2 * rows - 1) / (2 * rows) >= cus) {
        f = PanelFormat{};
        rows = std::min(2 * rows, 1024);
        st = panel_build(it.val_type, rows, it.m, it.n, it.nnz, it.d_row_ptr, it.d_col, it.d_val,
                         one, h_rp.data(), want == 3, 1, f, &panel, nullptr, layout);
      }
      if (st == SBLAS_OK && !panel && narrow && !(opts && opts->panel_rows > 0)) {
        // no launch of its own: the persistent launch's row panels, x in FP64
        narrow = false;
        f = PanelFormat{};
        rows = kBatchPanelRowsMax;
        layout = kPanelRowOrder;
        st = panel_build(it.val_type, rows, it.m, it.n, it.nnz, it.d_row_ptr, it.d_col, it.d_val,
                         one, h_rp.data(), want == 3, 1, f, &panel, nullptr, layout);
      }
      // row order with slices beyond 16-bit offsets (the batch kernel reads
      // 16-bit ones only): shorter row panels, down to 192 rows, before
      // giving up (the sorted layout has no such limit)
      while (st == SBLAS_OK && panel && layout == kPanelRowOrder && !f.srp && rows > 192 &&
             !(opts && opts->panel_rows > 0)) {
        for (void *q : f.buffers) (void)hipFree(q);
        f = PanelFormat{};
        rows /= 2;
        st = panel_build(it.val_type, rows, it.m, it.n, it.nnz, it.d_row_ptr, it.d_col, it.d_val,
                         one, h_rp.data(), want == 3, 1, f, &panel, nullptr, layout);
      }
      if (st == SBLAS_OK && panel && layout == kPanelRowOrder && !f.srp)
        panel = false;  // slices beyond 16-bit offsets
      if (st == SBLAS_ERR_UNSUPPORTED && want != 3) {
        clear_error();
        st = SBLAS_OK;
        panel = false;
      }
      if (st != SBLAS_OK) return fail(st);
      if (!panel) {
        for (void *q : f.buffers) (void)hipFree(q);
        f = PanelFormat{};
      }
    }
    if (panel) {
      M.pinfo = f.info;
      M.slice_base = f.slice_base;
      M.slice_off = f.slice_off;
      M.srp = f.srp;
      M.pval = f.pval;
      M.pcol = f.pcol;
      M.prows = f.rows;
      c.chunks = 1;
      if (narrow) {
        // FP32 copy of x, zero-padded to whole column panels, written by
        // the launch itself
        b->solo.push_back(i);
        const int64_t padded = (static_cast<int64_t>(it.n) + kPanelCols - 1) / kPanelCols *
                               kPanelCols;
        void *p = nullptr;
        if (hipMalloc(&p, sizeof(float) * std::max<int64_t>(padded, 4)) != hipSuccess) {
          set_error(SBLAS_ERR_ALLOC, "batch: FP32 x copy allocation failed");
          return fail(SBLAS_ERR_ALLOC);
        }
        f.buffers.push_back(p);
        if (hipMemset(p, 0, sizeof(float) * std::max<int64_t>(padded, 4)) != hipSuccess) {
          set_error(SBLAS_ERR_HIP, "batch: FP32 x copy set-up failed");
          return fail(SBLAS_ERR_HIP);
        }
        f.x32 = static_cast<float *>(p);
        add_x32(b->solo_xsegs, b->solo_xchunks, it.d_x, f.x32, it.n);
      }
    }
    b->mats.push_back(M);
    if (it.m == 0) {
      b->choices.push_back(sblas_batch_choice{SBLAS_KERNEL_VECTOR, c.subwave, c.items, c.chunks, 0,
                                              2, rs.mean, rs.cv, rs.max_row, 0, 0});
      continue;
    }
    std::vector<Pending> &mine = pending[i];
    if (panel && narrow) {
      // its own launch: no items
    } else if (panel) {
      for (size_t q = 0; q < f.hpanels.size(); q++) {
        const ExecChunk &e = f.hpanels[q];
        mine.push_back(Pending{BatchItem{i, kBatchPanel, static_cast<int>(q), 0, 0, 0, 0, 0},
                               e.hi - e.lo, static_cast<int>(e.lo), static_cast<int>(e.hi),
                               e.row_begin, e.row_end});
      }
    } else {
      // narrow values in row blocks gather x from an FP32 copy (as a plan
      // does for matrices of 2^20 elements or more)
      if (it.val_type != SBLAS_VAL_F64 && nxo == 2 && kv == 1 && it.nnz >= (int64_t{1} << 20) &&
          it.n > 0) {
        void *p = nullptr;
        if (hipMalloc(&p, sizeof(float) * static_cast<size_t>(it.n)) != hipSuccess) {
          set_error(SBLAS_ERR_ALLOC, "batch: FP32 x copy allocation failed");
          return fail(SBLAS_ERR_ALLOC);
        }
        b->bufs.push_back(p);
        b->mats.back().x32 = static_cast<const float *>(p);
        add_x32(b->xsegs, b->xchunks, it.d_x, static_cast<float *>(p), it.n);
      }
      const std::vector<int64_t> bounds =
          element_bounds(SBLAS_PART_V1, it.m, h_rp.data(), it.nnz, c.chunks, 1);
      const std::vector<ExecChunk> ex = exec_chunks_from_bounds(it.m, h_rp.data(), bounds);
      for (const RowBlock &r : build_row_blocks(ex, h_rp.data(), kStreamBlock * c.items))
        mine.push_back(Pending{BatchItem{i, kBatchRowBlock, r.row_begin, r.row_end, r.lo, r.hi,
                                         0, 0},
                               r.hi - r.lo, r.lo, r.hi, r.row_begin, r.row_end});
    }
    b->choices.push_back(sblas_batch_choice{SBLAS_KERNEL_VECTOR, c.subwave, c.items, c.chunks,
                                            static_cast<int>(panel && narrow ? f.npanels
                                                                             : mine.size()),
                                            panel ? 3 : 2, rs.mean,
                                            rs.cv, rs.max_row, panel ? f.rows : 0, 0});
    h_rps[i] = std::move(h_rp);
  }
  // list order: the row-panel items, then the row-block items (one
  // persistent launch each); within a kind matrix by matrix, each matrix's
  // items in its own order (row panels of one matrix stay neighbours, so
  // they share x in one XCD's L2)
  std::vector<std::pair<int, int>> order;  // (matrix, index in its list)
  for (int kind : {kBatchPanel, kBatchRowBlock})
    for (int i = 0; i < count; i++)
      for (int q = 0; q < static_cast<int>(pending[i].size()); q++)
        if (pending[i][q].item.kind == kind) order.emplace_back(i, q);
  std::vector<std::vector<int>> where(count);
  for (int i = 0; i < count; i++) where[i].assign(pending[i].size(), -1);
  int64_t nslots = 0;
  for (const auto &o : order) {
    where[o.first][o.second] = static_cast<int>(b->items.size());
    BatchItem item = pending[o.first][o.second].item;
    const int kv = b->mats[o.first].k;
    item.pslot = static_cast<int>(nslots);  // 2 slots of kv partials each
    nslots += 2 * kv;
    b->items.push_back(item);
  }
  if (nslots > INT32_MAX) {
    set_error(SBLAS_ERR_ARGS, "batch: too many work items");
    return fail(SBLAS_ERR_ARGS);
  }
  // per kind, eight queues of about equal work, contiguous in that kind's
  // part of the list (bounds relative to its first item)
  size_t n_panel = 0;
  while (n_panel < b->items.size() && b->items[n_panel].kind == kBatchPanel) n_panel++;
  b->n_panel = static_cast<int>(n_panel);
  std::vector<int> qbound;
  const size_t cut[3] = {0, n_panel, order.size()};
  for (int part = 0; part < 2; part++) {
    const size_t w0 = cut[part], w1 = cut[part + 1];
    auto work = [&](size_t w) {  // elements + per-item cost
      return pending[order[w].first][order[w].second].work * b->mats[order[w].first].k + 256;
    };
    int64_t total_work = 0;
    for (size_t w = w0; w < w1; w++) total_work += work(w);
    std::vector<int> qb(kBatchQueues + 1, static_cast<int>(w1 - w0));
    qb[0] = 0;
    int64_t run = 0;
    int q = 1;
    for (size_t w = w0; w < w1 && q < kBatchQueues; w++) {
      while (q < kBatchQueues && run >= total_work * q / kBatchQueues)
        qb[q++] = static_cast<int>(w - w0);
      run += work(w);
    }
    qbound.insert(qbound.end(), qb.begin(), qb.end());
  }
  // rows cut by an item's element window: partial slots pslot (first row) and
  // pslot + k (last row), listed per row in the matrix's own item order
  for (int i = 0; i < count; i++) {
    const std::vector<int> &h_rp = h_rps[i];
    for (size_t q = 0; q < pending[i].size(); q++) {
      const Pending &r = pending[i][q];
      const int w = where[i][q];
      auto split = [&](int row) { return h_rp[row] < r.lo || h_rp[row + 1] > r.hi; };
      auto add = [&](int row, int slot) {
        if (!grow.empty() && gmat.back() == i && grow.back() == row) {
          gslot.push_back(slot);
          gptr.back() = static_cast<int>(gslot.size());
        } else {
          gmat.push_back(i);
          grow.push_back(row);
          gslot.push_back(slot);
          gptr.push_back(static_cast<int>(gslot.size()));
        }
      };
      const int first = r.row_begin, last = r.row_end - 1;
      if (first > last) continue;
      const int base = b->items[w].pslot, kv = b->mats[i].k;
      if (split(first)) add(first, base);
      if (last != first && split(last)) add(last, base + kv);
    }
  }
  b->groups = static_cast<int>(grow.size());
  int st;
  if ((st = b->upload(&b->d_mats, b->mats)) != SBLAS_OK ||
      (st = b->upload(&b->d_items, b->items)) != SBLAS_OK ||
      (st = b->upload(&b->d_gmat, gmat)) != SBLAS_OK ||
      (st = b->upload(&b->d_grow, grow)) != SBLAS_OK ||
      (st = b->upload(&b->d_gptr, gptr)) != SBLAS_OK ||
      (st = b->upload(&b->d_gslot, gslot)) != SBLAS_OK ||
      (st = b->upload(&b->d_xsegs, b->xsegs)) != SBLAS_OK ||
      (st = b->upload(&b->d_xchunks, b->xchunks)) != SBLAS_OK ||
      (st = b->upload(&b->d_solo_xsegs, b->solo_xsegs)) != SBLAS_OK ||
      (st = b->upload(&b->d_solo_xchunks, b->solo_xchunks)) != SBLAS_OK)
    return fail(st);
  std::vector<double> zeros(static_cast<size_t>(std::max<int64_t>(nslots, 1)), 0.0);
  if ((st = b->upload(&b->d_partials, zeros)) != SBLAS_OK) return fail(st);
  if ((st = b->upload(&b->d_ctr, std::vector<int>(2 * (kBatchQueues + 1), 0))) != SBLAS_OK ||
      (st = b->upload(&b->d_qbound, qbound)) != SBLAS_OK)
    return fail(st);
  // per kind the resident workgroups, never more than there are items.  With
  // another kind beside them, panels take one workgroup per CU less than fit:
  // the panel workgroups fill a CU's LDS, and the other launch would wait for
  // the panel launch to finish (64-matrix batch: 4.87 -> 4.69 ms)
  const int n_rows = static_cast<int>(b->items.size()) - b->n_panel;
  const int kinds = (b->solo.empty() ? 0 : 1) + (b->n_panel > 0) + (n_rows > 0);
  const int occ_panel = batch_occupancy(kBatchPanel, b->items_per_thread);
  const int panel_per_cu = (kinds > 1 && occ_panel > 1) ? occ_panel - 1 : occ_panel;
  b->grid_panel = std::min<int>(b->n_panel, std::max(1, cus) * panel_per_cu);
  b->grid_rows = std::min<int>(n_rows, std::max(1, cus) *
                                           batch_occupancy(kBatchRowBlock, b->items_per_thread));
  if (hipStreamCreateWithFlags(&b->side, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&b->side2, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&b->join, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&b->join2, hipEventDisableTiming) != hipSuccess) {
    set_error(SBLAS_ERR_HIP, "batch: stream or event creation failed");
    return fail(SBLAS_ERR_HIP);
  }
  *out = b;
  return SBLAS_OK;
}

int sblas_batch_run(sblas_batch b, void *stream) {
  sblas::clear_error();
  if (!b) return SBLAS_ERR_ARGS;
  return b->enqueue(static_cast<hipStream_t>(stream));
}

int sblas_batch_graph_launch(sblas_batch b, void *stream) {
  sblas::clear_error();
  if (!b) return SBLAS_ERR_ARGS;
  if (!b->graph_exec) {
    // capture on a private stream (kernel launches only; nothing to clear)
    SBLAS_HIP_TRY(hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking));
    SBLAS_HIP_TRY(hipStreamBeginCapture(b->cap_stream, hipStreamCaptureModeThreadLocal));
    const int st = b->enqueue(b->cap_stream);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(b->cap_stream, &g);
    if (st != SBLAS_OK) {
      if (g) (void)hipGraphDestroy(g);
      return st;
    }
    if (e != hipSuccess) {
      sblas::set_error(SBLAS_ERR_HIP, std::string("batch capture: ") + hipGetErrorString(e));
      return SBLAS_ERR_HIP;
    }
    b->graph = g;
    SBLAS_HIP_TRY(hipGraphInstantiate(&b->graph_exec, b->graph, nullptr, nullptr, 0));
  }
  SBLAS_HIP_TRY(hipGraphLaunch(b->graph_exec, static_cast<hipStream_t>(stream)));
  return SBLAS_OK;
}

int sblas_batch_get_choice(sblas_batch b, int index, sblas_batch_choice *out) {
  if (!b || !out || index < 0 || index >= static_cast<int>(b->choices.size()))
    return SBLAS_ERR_ARGS;
  *out = b->choices[index];
  return SBLAS_OK;
}

int sblas_batch_get_shape(sblas_batch b, int *work_items, int *grid, int *launches) {
  if (!b) return SBLAS_ERR_ARGS;
  if (work_items) *work_items = static_cast<int>(b->items.size());
  if (grid) *grid = std::max(b->grid_panel, b->grid_rows);
  if (launches)
    *launches = (b->grid_panel > 0 ? 1 : 0) + (b->grid_rows > 0 ? 1 : 0) +
                (b->groups > 0 ? 1 : 0) + static_cast<int>(b->solo.size()) +
                (b->solo.empty() ? 0 : 1) + (b->xsegs.empty() ? 0 : 1);
  return SBLAS_OK;
}

int sblas_batch_destroy(sblas_batch b) {
  delete b;
  return SBLAS_OK;
}

}  // extern "C"
