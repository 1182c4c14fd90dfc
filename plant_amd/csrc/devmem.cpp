// Caching device allocator behind DBuf (see dev.hpp).
#include <algorithm>
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

size_t dev_cached_bytes() {
  int dev = 0;
  (void)hipGetDevice(&dev);
  std::lock_guard<std::mutex> lk(g_mu);
  size_t b = 0;
  for (auto& kv : g_cache)
    if (kv.second.second == dev) b += kv.first;
  return b;
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

// ---- ChunkArray ---------------------------------------------------------------

namespace {
hipMemAllocationProp device_prop(int dev) {
  hipMemAllocationProp prop = {};
  prop.type = hipMemAllocationTypePinned;
  prop.location.type = hipMemLocationTypeDevice;
  prop.location.id = dev;
  return prop;
}
size_t round_up(size_t x, size_t g) { return (x + g - 1) / g * g; }
}  // namespace

hipError_t ChunkArray::alloc(size_t bytes, size_t chunk_bytes) {
  release();
  (void)hipGetDevice(&device);
  hipMemAllocationProp prop = device_prop(device);
  size_t gran = 0;
  hipError_t e = hipMemGetAllocationGranularity(&gran, &prop, hipMemAllocationGranularityRecommended);
  if (e == hipSuccess && gran) {
    chunk = round_up(std::max<size_t>(chunk_bytes, 1), std::max<size_t>(gran, (size_t)2 << 20));
    const size_t total = round_up(std::max<size_t>(bytes, 1), chunk);
    void* ptr = nullptr;
    e = hipMemAddressReserve(&ptr, total, 0, nullptr, 0);
    if (e == hipSuccess && ptr) {
      base = (char*)ptr;
      reserved = total;
      vmm = true;
      for (size_t off = 0; off < total && e == hipSuccess; off += chunk) {
        hipMemGenericAllocationHandle_t h{};
        e = hipMemCreate(&h, chunk, &prop, 0);
        if (e != hipSuccess) {
          (void)hipGetLastError();
          dev_cache_flush();
          e = hipMemCreate(&h, chunk, &prop, 0);
          if (e != hipSuccess) break;
        }
        e = hipMemMap(base + off, chunk, 0, h, 0);
        if (e != hipSuccess) {
          (void)hipMemRelease(h);
          break;
        }
        maps.push_back(h);
        mapped = off + chunk;
      }
      if (e == hipSuccess) {
        hipMemAccessDesc desc = {};
        desc.location = prop.location;
        desc.flags = hipMemAccessFlagsProtReadWrite;
        e = hipMemSetAccess(base, mapped, &desc, 1);
      }
      if (e == hipSuccess) return hipSuccess;
      (void)hipGetLastError();
      release();
    }
    (void)hipGetLastError();
  }
  // plain allocation (shrink keeps it whole)
  void* p = nullptr;
  e = hipMalloc(&p, bytes ? bytes : 1);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    dev_cache_flush();
    e = hipMalloc(&p, bytes ? bytes : 1);
    if (e != hipSuccess) return e;
  }
  base = (char*)p;
  mapped = bytes ? bytes : 1;
  vmm = false;
  return hipSuccess;
}

hipError_t ChunkArray::shrink(size_t bytes) {
  if (!vmm || !base) return hipSuccess;
  const size_t keep = std::max<size_t>(1, (bytes + chunk - 1) / chunk);
  while (maps.size() > keep) {
    const size_t off = (maps.size() - 1) * chunk;
    hipError_t e = hipMemUnmap(base + off, chunk);
    if (e != hipSuccess) return e;
    e = hipMemRelease(maps.back());
    if (e != hipSuccess) return e;
    maps.pop_back();
    mapped = off;
  }
  return hipSuccess;
}

void ChunkArray::release() {
  if (!base) return;
  if (vmm) {
    (void)hipDeviceSynchronize();
    for (size_t k = 0; k < maps.size(); ++k) {
      (void)hipMemUnmap(base + k * chunk, chunk);
      (void)hipMemRelease(maps[k]);
    }
    (void)hipMemAddressFree(base, reserved);
  } else {
    (void)hipFree(base);
  }
  maps.clear();
  base = nullptr;
  reserved = mapped = chunk = 0;
  vmm = false;
}

}  // namespace chl
