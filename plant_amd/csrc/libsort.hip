// The one library primitive of the build: rocPRIM's radix sort through hipCUB,
// for the device orders' final sort of (score or degree key, id) pairs
// (order_device.cpp). Every label-path kernel is hand-written.
#include <hipcub/hipcub.hpp>

#include "dev.hpp"

namespace chl {

hipError_t sort_pairs_u64(void* tmp, size_t& tmp_bytes, const uint64_t* kin, uint64_t* kout, const uint64_t* vin,
                          uint64_t* vout, uint64_t m, int end_bit, hipStream_t s) {
  return hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, kin, kout, vin, vout, (int)m, 0, end_bit, s);
}

}  // namespace chl
