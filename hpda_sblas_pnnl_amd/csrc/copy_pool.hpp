// copy_pool.hpp -- a fixed set of host threads that split one memcpy between
// them.  Started once and reused for every copy, so a copy pays no thread
// start-up: the pinned staging of host arrays copies each 64 MB piece of a
// caller's array into a pinned slot with it.  Plain C++ (no HIP).
#pragma once

#include <algorithm>
#include <condition_variable>
#include <cstddef>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

namespace sblas {

class CopyPool {
 public:
  // `threads` copies in parallel: the caller and threads - 1 workers.  If
  // starting a worker fails, the ones started are stopped and joined before
  // the error propagates.
  explicit CopyPool(unsigned threads) {
    try {
      for (unsigned i = 1; i < std::max(threads, 1u); i++) workers_.emplace_back([this, i] { work(i); });
    } catch (...) {
      stop_all();
      throw;
    }
  }
  CopyPool(const CopyPool &) = delete;
  CopyPool &operator=(const CopyPool &) = delete;
  ~CopyPool() { stop_all(); }

  // dst[0, bytes) = src[0, bytes); returns when every part is copied.  Copies
  // below `serial_below` bytes are done by the caller alone.
  void copy(void *dst, const void *src, size_t bytes, size_t serial_below = size_t{4} << 20) {
    if (workers_.empty() || bytes < serial_below) {
      if (bytes) std::memcpy(dst, src, bytes);
      return;
    }
    {
      std::lock_guard<std::mutex> lk(m_);
      dst_ = static_cast<char *>(dst);
      src_ = static_cast<const char *>(src);
      bytes_ = bytes;
      pending_ = static_cast<unsigned>(workers_.size());
      gen_++;
    }
    go_.notify_all();
    part(0, dst_, src_, bytes);
    std::unique_lock<std::mutex> lk(m_);
    done_.wait(lk, [this] { return pending_ == 0; });
  }

  unsigned threads() const { return static_cast<unsigned>(workers_.size()) + 1; }

 private:
  void stop_all() {
    {
      std::lock_guard<std::mutex> lk(m_);
      stop_ = true;
    }
    go_.notify_all();
    for (std::thread &t : workers_) t.join();
    workers_.clear();
  }
  // part i of n = threads(): a contiguous range, the last one shorter
  void part(unsigned i, char *dst, const char *src, size_t bytes) const {
    const size_t n = threads(), per = (bytes + n - 1) / n;
    const size_t lo = std::min(bytes, i * per), hi = std::min(bytes, lo + per);
    if (hi > lo) std::memcpy(dst + lo, src + lo, hi - lo);
  }
  void work(unsigned id) {
    size_t seen = 0;
    for (;;) {
      std::unique_lock<std::mutex> lk(m_);
      go_.wait(lk, [&] { return stop_ || gen_ != seen; });
      if (stop_) return;
      seen = gen_;
      char *d = dst_;
      const char *s = src_;
      const size_t b = bytes_;
      lk.unlock();
      part(id, d, s, b);
      lk.lock();
      // the caller waits for every worker before it publishes the next copy,
      // so each worker sees each copy exactly once
      if (--pending_ == 0) done_.notify_one();
    }
  }

  std::vector<std::thread> workers_;
  std::mutex m_;
  std::condition_variable go_, done_;
  size_t gen_ = 0;
  unsigned pending_ = 0;
  bool stop_ = false;
  char *dst_ = nullptr;
  const char *src_ = nullptr;
  size_t bytes_ = 0;
};

}  // namespace sblas
