// host_util.hpp -- host-side helpers of the synchronous entry points that
// take host arrays (the reference's drop-in calls and ingestion helpers).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <memory>
#include <vector>

#include "copy_pool.hpp"

namespace sblas {

// Makes `dev` current for the scope and restores the caller's device after.
struct DeviceGuard {
  int prev = -1;
  bool ok = false;
  explicit DeviceGuard(int dev) {
    if (hipGetDevice(&prev) != hipSuccess) return;
    ok = hipSetDevice(dev) == hipSuccess;
  }
  ~DeviceGuard() {
    if (prev >= 0) (void)hipSetDevice(prev);
  }
  DeviceGuard(const DeviceGuard &) = delete;
  DeviceGuard &operator=(const DeviceGuard &) = delete;
};

// Device allocations freed when the scope ends.
struct DeviceArrays {
  std::vector<void *> ptrs;
  DeviceArrays() = default;
  DeviceArrays(const DeviceArrays &) = delete;
  DeviceArrays &operator=(const DeviceArrays &) = delete;
  ~DeviceArrays() {
    for (void *p : ptrs) (void)hipFree(p);
  }
  template <typename T>
  T *alloc(size_t count) {
    void *p = nullptr;
    if (hipMalloc(&p, sizeof(T) * (count ? count : 1)) != hipSuccess) return nullptr;
    ptrs.push_back(p);
    return static_cast<T *>(p);
  }
};

// A non-blocking stream destroyed when the scope ends (nullptr if creation failed).
struct OwnedStream {
  hipStream_t s = nullptr;
  OwnedStream() { (void)hipStreamCreateWithFlags(&s, hipStreamNonBlocking); }
  ~OwnedStream() {
    if (s) (void)hipStreamDestroy(s);
  }
  OwnedStream(const OwnedStream &) = delete;
  OwnedStream &operator=(const OwnedStream &) = delete;
};

// Host-to-device copies of pageable host arrays through a ring of pinned
// buffers (staging.cpp): a pool of host threads, started once per ring,
// fills one slot while earlier slots' transfers run.  to_device issues the
// transfers on s and returns once the last piece is queued; a slot is
// refilled only after its previous transfer has completed.
struct PinnedRing {
  std::vector<void *> slots;
  std::vector<hipEvent_t> events;
  std::vector<bool> used;
  std::unique_ptr<CopyPool> pool;
  size_t next = 0;
  PinnedRing() = default;
  PinnedRing(const PinnedRing &) = delete;
  PinnedRing &operator=(const PinnedRing &) = delete;
  ~PinnedRing();
  int init();
  int to_device(void *dst, const void *src, size_t bytes, hipStream_t s);
  int wait_all();  // every transfer issued so far has completed
};

// One call's use of the process-wide staging ring of the current device.
// The ring (3 x 64 MB pinned, a pool of copy threads) is created by the first
// call that stages and kept until the process ends.  A call stages only when
// its uploads total at least kStageMinBytes, the ring is free (another
// thread's call holding it makes this one copy directly) and the pinned
// allocation succeeds; otherwise to_device copies straight from the pageable
// array.  The destructor waits for this call's staged transfers before
// handing the ring on.
class StagingLease {
 public:
  static constexpr size_t kStageMinBytes = size_t{8} << 20;
  explicit StagingLease(size_t total_bytes);
  ~StagingLease();
  StagingLease(const StagingLease &) = delete;
  StagingLease &operator=(const StagingLease &) = delete;
  int to_device(void *dst, const void *src, size_t bytes, hipStream_t s);
  bool staged() const { return ring_ != nullptr; }

 private:
  PinnedRing *ring_ = nullptr;
  int device_ = -1;
};

}  // namespace sblas
