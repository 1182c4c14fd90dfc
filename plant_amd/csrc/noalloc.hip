// The host-pointer distance query without any device allocation, for callers
// that cannot allocate (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc): chl_query stages the pairs and distances in
// device buffers the label handle allocates and grows; this variant stages them
// in `scratch`, a file-scope device array of fixed size, so a call allocates
// nothing. The device address of the array is taken once per device with
// hipGetSymbolAddress (the array's name is not a device address in host code)
// and kept; every call on a device shares that device's array, so calls that
// overlap on one device race (the caller serialises them).
#include <mutex>

#include "../../include/chl.h"
#include "dev.hpp"
#include "internal.hpp"

namespace chl {
void labels_views(const chl_labels* l, QueryView* qv, uint32_t* n, int* device);
hipStream_t labels_stream(const chl_labels* l);

namespace {

// kNoAllocMaxPairs pairs (8 bytes each) then as many distances (4 bytes each)
constexpr int64_t kNoAllocMaxPairs = (int64_t)1 << 20;
static __device__ __align__(256) uint8_t scratch[kNoAllocMaxPairs * 12];

constexpr int kMaxDevices = 64;
std::mutex g_mu;
uint8_t* g_ptr[kMaxDevices];  // per device: the address of `scratch` there (null until taken)

int scratch_of(int device, uint8_t** out) {
  if (device < 0 || device >= kMaxDevices) return set_error(CHL_E_DEVICE, "device %d out of range", device);
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[device]) {
    void* p = nullptr;
    CHL_HIP(hipSetDevice(device));
    CHL_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return set_error(CHL_E_DEVICE, "no device address for the query scratch");
    g_ptr[device] = static_cast<uint8_t*>(p);
  }
  *out = g_ptr[device];
  return CHL_OK;
}

}  // namespace
}  // namespace chl

using namespace chl;

extern "C" {

int64_t chl_query_noalloc_max(void) { return kNoAllocMaxPairs; }

int chl_query_noalloc(const chl_labels* l, const uint32_t* pairs, int64_t q, int32_t* out) {
  if (!l || q < 0 || (q > 0 && (!pairs || !out))) return set_error(CHL_E_INVAL, "bad arguments");
  if (q > kNoAllocMaxPairs)
    return set_error(CHL_E_RANGE, "%lld pairs: chl_query_noalloc takes at most %lld per call", (long long)q,
                     (long long)kNoAllocMaxPairs);
  if (q == 0) return CHL_OK;
  QueryView qv;
  uint32_t n = 0;
  int device = 0;
  labels_views(l, &qv, &n, &device);
  for (int64_t k = 0; k < 2 * q; ++k)
    if (pairs[k] >= n) return set_error(CHL_E_RANGE, "vertex %u out of range", pairs[k]);
  uint8_t* base = nullptr;
  int rc = scratch_of(device, &base);
  if (rc) return rc;
  CHL_HIP(hipSetDevice(device));
  uint32_t* d_pairs = reinterpret_cast<uint32_t*>(base);
  int32_t* d_out = reinterpret_cast<int32_t*>(base + kNoAllocMaxPairs * 8);
  hipStream_t s = labels_stream(l);
  CHL_HIP(hipMemcpyAsync(d_pairs, pairs, sizeof(uint32_t) * 2 * q, hipMemcpyHostToDevice, s));
  CHL_HIP(launch_query(qv, d_pairs, n, (uint64_t)q, d_out, s));
  CHL_HIP(hipMemcpyAsync(out, d_out, sizeof(int32_t) * q, hipMemcpyDeviceToHost, s));
  CHL_HIP(hipStreamSynchronize(s));
  return CHL_OK;
}

}  // extern "C"
