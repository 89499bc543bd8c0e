// staging.cpp -- host-to-device copies of the caller's (pageable) arrays
// through a ring of pinned buffers.
//
// A copy from pageable memory is staged by the runtime at about half the PCIe
// rate.  Here each copy is cut into pieces; a pool of host threads (started
// with the ring, so a piece pays no thread start-up) copies a piece into a
// free pinned slot while earlier slots' DMA transfers run, and a slot is
// refilled only after the event recorded behind its last transfer has
// completed.  Transfers go to the stream of the copy, so the kernels that
// consume them can start while later pieces are still on their way.
#include <algorithm>
#include <cstring>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "host_util.hpp"
#include "internal.hpp"

namespace sblas {
namespace {

constexpr size_t kSlotBytes = size_t{64} << 20;  // one piece
constexpr int kSlots = 3;
constexpr unsigned kThreadsMax = 16;               // host copy threads (the caller among them)
constexpr int kMaxDevices = 64;

// The process-wide rings, one per device, each behind its own mutex.  Rings
// are never destroyed: their pinned buffers live until the process ends, and
// no HIP call runs from a static destructor.
struct RingHolder {
  std::mutex mu;
  PinnedRing *ring = nullptr;
  bool alloc_failed = false;  // pinned allocation failed once: copy directly from then on
};
RingHolder g_rings[kMaxDevices];

}  // namespace

PinnedRing::~PinnedRing() {
  for (size_t k = 0; k < slots.size(); k++) {
    if (used[k]) (void)hipEventSynchronize(events[k]);
    (void)hipHostFree(slots[k]);
    (void)hipEventDestroy(events[k]);
  }
}

int PinnedRing::init() {
  if (!slots.empty()) return SBLAS_OK;
  if (!pool) {  // no exception leaves the C entry points: without threads, copies run serially
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    try {
      pool.reset(new CopyPool(std::min(kThreadsMax, hw)));
    } catch (...) {
      pool.reset();
    }
  }
  for (int k = 0; k < kSlots; k++) {
    void *p = nullptr;
    hipEvent_t e = nullptr;
    if (hipHostMalloc(&p, kSlotBytes, hipHostMallocDefault) != hipSuccess) {
      set_error(SBLAS_ERR_ALLOC, "pinned staging: allocation failed");
      return SBLAS_ERR_ALLOC;
    }
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
      (void)hipHostFree(p);
      set_error(SBLAS_ERR_HIP, "pinned staging: event creation failed");
      return SBLAS_ERR_HIP;
    }
    slots.push_back(p);
    events.push_back(e);
    used.push_back(false);
  }
  return SBLAS_OK;
}

int PinnedRing::wait_all() {
  for (size_t k = 0; k < slots.size(); k++)
    if (used[k]) SBLAS_HIP_TRY(hipEventSynchronize(events[k]));
  return SBLAS_OK;
}

int PinnedRing::to_device(void *dst, const void *src, size_t bytes, hipStream_t s) {
  if (bytes == 0) return SBLAS_OK;
  int st = init();
  if (st != SBLAS_OK) return st;
  for (size_t off = 0; off < bytes; off += kSlotBytes) {
    const size_t len = std::min(kSlotBytes, bytes - off);
    const size_t k = next++ % slots.size();
    if (used[k]) SBLAS_HIP_TRY(hipEventSynchronize(events[k]));  // its last transfer is done
    if (pool)
      pool->copy(slots[k], static_cast<const char *>(src) + off, len);
    else
      std::memcpy(slots[k], static_cast<const char *>(src) + off, len);
    SBLAS_HIP_TRY(hipMemcpyAsync(static_cast<char *>(dst) + off, slots[k], len,
                                 hipMemcpyHostToDevice, s));
    SBLAS_HIP_TRY(hipEventRecord(events[k], s));
    used[k] = true;
  }
  return SBLAS_OK;
}

StagingLease::StagingLease(size_t total_bytes) {
  if (total_bytes < kStageMinBytes) return;
  if (hipGetDevice(&device_) != hipSuccess || device_ < 0 || device_ >= kMaxDevices) return;
  RingHolder &h = g_rings[device_];
  if (!h.mu.try_lock()) return;  // another call is staging on this device
  if (h.alloc_failed) {
    h.mu.unlock();
    return;
  }
  if (!h.ring) {
    PinnedRing *r = new (std::nothrow) PinnedRing;
    if (!r || r->init() != SBLAS_OK) {
      // pinned memory unavailable: the calls copy from pageable memory
      delete r;
      h.alloc_failed = true;
      h.mu.unlock();
      clear_error();
      return;
    }
    h.ring = r;
  }
  ring_ = h.ring;
}

StagingLease::~StagingLease() {
  if (!ring_) return;
  (void)ring_->wait_all();
  g_rings[device_].mu.unlock();
}

int StagingLease::to_device(void *dst, const void *src, size_t bytes, hipStream_t s) {
  if (bytes == 0) return SBLAS_OK;
  if (ring_) return ring_->to_device(dst, src, bytes, s);
  SBLAS_HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, s));
  return SBLAS_OK;
}

}  // namespace sblas
