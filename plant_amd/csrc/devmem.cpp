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

// ---- GrowArray ----------------------------------------------------------------

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

hipError_t GrowArray::grow(size_t bytes, size_t keep_bytes, hipStream_t s) {
  if (base && bytes <= mapped) return hipSuccess;
  if (!base) {
    (void)hipGetDevice(&device);
    // a virtual range as large as the device, mapped as needed
    hipMemAllocationProp prop = device_prop(device);
    size_t gran = 0, f = 0, t = 0;
    if (hipMemGetInfo(&f, &t) == hipSuccess &&
        hipMemGetAllocationGranularity(&gran, &prop, hipMemAllocationGranularityRecommended) == hipSuccess &&
        gran) {
      const size_t res = round_up(t, std::max<size_t>(gran, (size_t)2 << 20));
      void* ptr = nullptr;
      if (hipMemAddressReserve(&ptr, res, 0, nullptr, 0) == hipSuccess && ptr) {
        base = (char*)ptr;
        reserved = res;
        vmm = true;
      }
    }
    (void)hipGetLastError();
  }
  if (vmm) {
    hipError_t e = vmm_map(bytes);
    if (e == hipSuccess) return hipSuccess;
    (void)hipGetLastError();
    // the runtime refused: continue as a plain allocation (the kept prefix is
    // copied out of the mapped range below, then the range is released)
  }
  // fallback: a larger allocation, the kept prefix copied over
  size_t want = std::max(bytes, mapped + mapped / 4);
  void* nb = nullptr;
  hipError_t e = hipMalloc(&nb, want);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    dev_cache_flush();
    want = bytes;
    e = hipMalloc(&nb, want);
    if (e != hipSuccess) return e;
  }
  if (base && keep_bytes) {
    e = hipMemcpyAsync(nb, base, std::min(keep_bytes, mapped), hipMemcpyDeviceToDevice, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) {
      (void)hipFree(nb);
      return e;
    }
  }
  release();
  base = (char*)nb;
  mapped = want;
  return hipSuccess;
}

// Maps physical memory behind the range up to at least `bytes`; access is set
// on the whole mapped range after each new mapping.
hipError_t GrowArray::vmm_map(size_t bytes) {
  hipMemAllocationProp prop = device_prop(device);
  size_t gran = 0;
  hipError_t e = hipMemGetAllocationGranularity(&gran, &prop, hipMemAllocationGranularityRecommended);
  if (e != hipSuccess || !gran) return e != hipSuccess ? e : hipErrorInvalidValue;
  gran = std::max<size_t>(gran, (size_t)2 << 20);
  // at least a quarter of what is mapped (few mappings) and 32 MB
  const size_t need = bytes - mapped;
  size_t step = round_up(std::max(need, std::max(mapped / 4, (size_t)32 << 20)), gran);
  if (mapped + step > reserved) step = round_up(need, gran);
  if (mapped + step > reserved) return hipErrorOutOfMemory;
  hipMemGenericAllocationHandle_t h{};
  e = hipMemCreate(&h, step, &prop, 0);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    dev_cache_flush();
    step = round_up(need, gran);
    e = hipMemCreate(&h, step, &prop, 0);
    if (e != hipSuccess) return e;
  }
  e = hipMemMap(base + mapped, step, 0, h, 0);
  if (e != hipSuccess) {
    (void)hipMemRelease(h);
    return e;
  }
  hipMemAccessDesc desc = {};
  desc.location = prop.location;
  desc.flags = hipMemAccessFlagsProtReadWrite;
  e = hipMemSetAccess(base, mapped + step, &desc, 1);
  if (e != hipSuccess) {
    (void)hipMemUnmap(base + mapped, step);
    (void)hipMemRelease(h);
    return e;
  }
  maps.emplace_back(step, h);
  mapped += step;
  return hipSuccess;
}

void GrowArray::release() {
  if (!base) return;
  if (vmm) {
    (void)hipDeviceSynchronize();
    size_t off = 0;
    for (auto& m : maps) {
      (void)hipMemUnmap(base + off, m.first);
      (void)hipMemRelease(m.second);
      off += m.first;
    }
    (void)hipMemAddressFree(base, reserved);
  } else {
    (void)hipFree(base);
  }
  maps.clear();
  base = nullptr;
  reserved = mapped = 0;
  vmm = false;
}

}  // namespace chl
