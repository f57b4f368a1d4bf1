// Device memory for the library's own buffers (operators, solver scratch,
// batches): hipMalloc / hipFree behind an exact-size cache of freed blocks.
//
// A serving loop releases and re-creates operators and batches of the same
// lattice many times.  Plain hipMalloc of gigabyte buffers can stall for
// seconds right after large frees, and hipFree synchronises the device, which
// a graph capture running on another thread does not survive.  Freed blocks
// are therefore kept per (device, size) and handed out again to a request of
// the same size; at most an eighth of the device's memory (and 32 GB) is
// kept that way, beyond which blocks are freed.  An allocation that fails returns
// every cached block of the device to the runtime and tries once more, and
// wc_release_cached_memory does the same on request.
//
// hipFree waits for all work on the device before the block can be used
// again.  A cached block is handed out again at once, possibly to work on
// another stream, so a free waits for the block's device just the same
// before it caches the block: kernels still queued on the old owner's
// streams cannot run into the new owner's data.  Frees happen when
// operators and batches are destroyed or their scratch grows.
//
// A device-wide wait (or hipFree, which waits too) breaks a graph capture
// in progress on any thread.  Captures of CG runs therefore hold g_capmu
// shared (wc_capture_begin / end), and every device-wide wait and hipFree
// here holds it exclusively: it waits for running captures to end and keeps
// new ones from starting meanwhile.  A thread's own frees inside a run of
// iterations (scratch that grows on the run's first iteration, possibly
// inside its capture) are held until the run has ended on its stream
// (wc_alloc_defer_*).  Scratch that grows outside a run was last used on
// the stream of the call that grows it: wc_dev_free_on waits for that stream
// alone.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <shared_mutex>
#include <unordered_map>
#include <utility>
#include <vector>

#include "wc_internal.h"

namespace {

constexpr size_t kCacheMax = (size_t)32 << 30;  // bytes kept per device at most

struct Live {
  int dev;
  size_t bytes;
};

// lock order: g_capmu before g_mu
std::shared_mutex g_capmu;  // shared: a capture in progress; exclusive: a device-wide wait or hipFree
std::mutex g_mu;
std::unordered_map<void *, Live> g_live;                  // blocks handed out
std::map<std::pair<int, size_t>, std::vector<void *>> g_free;  // cached, by (device, size)
std::map<int, size_t> g_cached;                           // cached bytes per device
std::map<int, size_t> g_cap;                              // cache limit per device

thread_local int t_capturing = 0;  // this thread holds g_capmu shared
// frees held by this thread until its run of iterations has ended
thread_local int t_defer = 0;
thread_local std::vector<void *> t_pending;

// the cache limit of dev, computed on its first allocation (dev current then)
size_t cap_of(int dev) {
  auto it = g_cap.find(dev);
  if (it != g_cap.end()) return it->second;
  size_t free_b = 0, total = 0;
  size_t cap = (hipMemGetInfo(&free_b, &total) == hipSuccess) ? total / 8 : 0;
  cap = std::min(cap, kCacheMax);
  g_cap[dev] = cap;
  return cap;
}

// hipFree of the cached blocks of dev (g_capmu exclusive and g_mu held)
size_t flush_locked(int dev) {
  size_t n = 0;
  for (auto it = g_free.begin(); it != g_free.end();) {
    if (it->first.first != dev) {
      ++it;
      continue;
    }
    for (void *p : it->second) {
      (void)hipFree(p);
      n += it->first.second;
    }
    it = g_free.erase(it);
  }
  g_cached[dev] = 0;
  return n;
}

// caches a block no work uses any more, or frees it past the cache limit
hipError_t release_block(void *p) {
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_live.find(p);
    if (it == g_live.end()) return hipErrorInvalidValue;  // freed twice
    const Live l = it->second;
    g_live.erase(it);
    if (g_cached[l.dev] + l.bytes <= cap_of(l.dev)) {
      g_free[{l.dev, l.bytes}].push_back(p);
      g_cached[l.dev] += l.bytes;
      return hipSuccess;
    }
  }
  // no longer known to anyone: free it when no capture runs
  std::unique_lock<std::shared_mutex> q(g_capmu);
  return hipFree(p);
}

}  // namespace

void wc_capture_begin() {
  g_capmu.lock_shared();
  ++t_capturing;
}

void wc_capture_end() {
  --t_capturing;
  g_capmu.unlock_shared();
}

void wc_alloc_defer_begin() { ++t_defer; }

void wc_alloc_defer_end(bool synced) {
  if (--t_defer > 0) return;
  std::vector<void *> held;
  held.swap(t_pending);
  for (void *p : held) (void)(synced ? release_block(p) : wc_dev_free(p));
}

hipError_t wc_dev_malloc_raw(void **p, size_t bytes) {
  *p = nullptr;
  if (bytes == 0) return hipSuccess;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  bool flush = false;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    cap_of(dev);  // the device's cache limit, taken while it is the current one
    auto it = g_free.find({dev, bytes});
    if (it != g_free.end() && !it->second.empty()) {
      *p = it->second.back();
      it->second.pop_back();
      g_cached[dev] -= bytes;
      g_live[*p] = {dev, bytes};
      return hipSuccess;
    }
    e = hipMalloc(p, bytes);
    if (e == hipSuccess) {
      g_live[*p] = {dev, bytes};
      return hipSuccess;
    }
    (void)hipGetLastError();  // the failed attempt's error is not the caller's
    flush = g_cached[dev] > 0 && t_capturing == 0;
  }
  *p = nullptr;
  if (!flush) return e;
  // hand the cached blocks back to the runtime and try once more
  std::unique_lock<std::shared_mutex> q(g_capmu);
  std::lock_guard<std::mutex> lk(g_mu);
  flush_locked(dev);
  e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    return e;
  }
  g_live[*p] = {dev, bytes};
  return hipSuccess;
}

hipError_t wc_dev_free(const void *cp) {
  void *p = const_cast<void *>(cp);
  if (!p) return hipSuccess;
  int dev = -1;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_live.find(p);
    if (it != g_live.end()) dev = it->second.dev;
  }
  if (t_defer > 0 && dev >= 0) {
    t_pending.push_back(p);
    return hipSuccess;
  }
  if (t_capturing > 0) return hipErrorInvalidValue;  // a wait here would break this thread's capture
  std::unique_lock<std::shared_mutex> q(g_capmu);
  if (dev < 0) return hipFree(p);  // not ours: hand it straight back
  // the work still using the block ends first, as hipFree would wait for it
  int cur = dev;
  (void)hipGetDevice(&cur);
  if (cur != dev) (void)hipSetDevice(dev);
  (void)hipDeviceSynchronize();
  if (cur != dev) (void)hipSetDevice(cur);
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_live.find(p);
  if (it == g_live.end()) return hipErrorInvalidValue;  // freed twice meanwhile
  const Live l = it->second;
  g_live.erase(it);
  if (g_cached[l.dev] + l.bytes <= cap_of(l.dev)) {
    g_free[{l.dev, l.bytes}].push_back(p);
    g_cached[l.dev] += l.bytes;
    return hipSuccess;
  }
  return hipFree(p);
}

void wc_device_wait() {
  std::unique_lock<std::shared_mutex> q(g_capmu);
  (void)hipDeviceSynchronize();
}

hipError_t wc_host_free(void *p) {
  if (!p) return hipSuccess;
  if (t_capturing > 0) return hipErrorInvalidValue;  // would cut this thread's capture
  std::unique_lock<std::shared_mutex> q(g_capmu);
  return hipHostFree(p);
}

hipError_t wc_dev_free_idle(const void *cp) {
  void *p = const_cast<void *>(cp);
  if (!p) return hipSuccess;
  bool ours;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    ours = g_live.count(p) > 0;
  }
  if (!ours || t_defer > 0 || t_capturing > 0) return wc_dev_free(p);
  return release_block(p);
}

hipError_t wc_dev_free_on(const void *cp, hipStream_t s) {
  void *p = const_cast<void *>(cp);
  if (!p) return hipSuccess;
  bool ours;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    ours = g_live.count(p) > 0;
  }
  if (!ours || t_defer > 0) return wc_dev_free(p);
  if (t_capturing > 0) return hipErrorInvalidValue;
  // s itself being captured by the caller: a wait would cut that capture, so
  // the block is cached at once (callers that capture use the allocation-free
  // entry point, which never gets here)
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone) return release_block(p);
  // its last users ran on s: a stream-local wait, which no capture on
  // another stream minds
  const hipError_t e = hipStreamSynchronize(s);
  if (e != hipSuccess) return e;
  return release_block(p);
}

extern "C" int wc_release_cached_memory(long *released_bytes) {
  int dev = 0;
  WC_HIP(hipGetDevice(&dev));
  size_t n = 0;
  {
    std::unique_lock<std::shared_mutex> q(g_capmu);
    std::lock_guard<std::mutex> lk(g_mu);
    n = flush_locked(dev);
  }
  if (released_bytes) *released_bytes = (long)n;
  return WC_OK;
}
