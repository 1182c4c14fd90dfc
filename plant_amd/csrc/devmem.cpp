// Caching device allocator behind DBuf (see dev.hpp).
#include <map>
#include <mutex>
#include <utility>

#include "dev.hpp"

namespace chl {
namespace {

std::mutex g_mu;
std::multimap<size_t, std::pair<void*, int>> g_cache;  // bytes -> (pointer, device)
size_t g_cached = 0;

// Blocks kept at most: half of the device memory.
size_t cache_cap() {
  static size_t cap = [] {
    size_t f = 0, t = 0;
    if (hipMemGetInfo(&f, &t) != hipSuccess) return (size_t)0;
    return t / 2;
  }();
  return cap;
}

}  // namespace

hipError_t dev_alloc(void** p, size_t bytes, size_t* got) {
  int dev = 0;
  (void)hipGetDevice(&dev);
  {
    std::lock_guard<std::mutex> lk(g_mu);
    for (auto it = g_cache.lower_bound(bytes); it != g_cache.end() && it->first <= 2 * bytes; ++it) {
      if (it->second.second != dev) continue;
      *p = it->second.first;
      *got = it->first;
      g_cached -= it->first;
      g_cache.erase(it);
      return hipSuccess;
    }
  }
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    dev_cache_flush();
    e = hipMalloc(p, bytes);
  }
  if (e == hipSuccess) *got = bytes;
  return e;
}

void dev_free(void* p, size_t bytes) {
  if (!p) return;
  int dev = 0;
  (void)hipGetDevice(&dev);
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_cached + bytes <= cache_cap()) {
      g_cache.emplace(bytes, std::make_pair(p, dev));
      g_cached += bytes;
      return;
    }
  }
  (void)hipFree(p);
}

void dev_cache_flush() {
  std::lock_guard<std::mutex> lk(g_mu);
  int cur = 0;
  (void)hipGetDevice(&cur);
  for (auto& kv : g_cache) {
    (void)hipSetDevice(kv.second.second);
    (void)hipFree(kv.second.first);
  }
  (void)hipSetDevice(cur);
  g_cache.clear();
  g_cached = 0;
}

}  // namespace chl
