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
// operators and batches are destroyed or their scratch grows, not inside
// solves.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <mutex>
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

std::mutex g_mu;
std::unordered_map<void *, Live> g_live;                  // blocks handed out
std::map<std::pair<int, size_t>, std::vector<void *>> g_free;  // cached, by (device, size)
std::map<int, size_t> g_cached;                           // cached bytes per device
std::map<int, size_t> g_cap;                              // cache limit per device

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

// returns the cached blocks of dev to the runtime (g_mu held)
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

}  // namespace

hipError_t wc_dev_malloc_raw(void **p, size_t bytes) {
  *p = nullptr;
  if (bytes == 0) return hipSuccess;
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
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
  if (e != hipSuccess && g_cached[dev] > 0) {
    (void)hipGetLastError();  // the failed attempt's error is not the caller's
    flush_locked(dev);
    e = hipMalloc(p, bytes);
  }
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
    if (it == g_live.end()) return hipFree(p);  // not ours: hand it straight back
    dev = it->second.dev;
  }
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

extern "C" int wc_release_cached_memory(long *released_bytes) {
  int dev = 0;
  WC_HIP(hipGetDevice(&dev));
  size_t n = 0;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    n = flush_locked(dev);
  }
  if (released_bytes) *released_bytes = (long)n;
  return WC_OK;
}
