// Label-store kernels: appending a batch of labels to the per-vertex rows,
// in-batch cleaning, and the merge-join distance query.
//
// The store keeps, per side, one row per vertex (rank ids) in CSR form:
// off[n+1] and ent[T] with ent = (hub << 32) | dist, rows hub-ascending. A batch
// covers a contiguous rank range above every hub already stored, so appending
// the batch's sorted segment after each row keeps rows sorted, as the
// superstep append of shared_memory/akiba.hpp:197-208 does.
#include <hipcub/hipcub.hpp>

#include "dev.hpp"

namespace chl {
namespace {

constexpr uint32_t kInfDist = 2147483647u;

// Lower bound of (v << 32) in the sorted batch keys, one thread per vertex.
__global__ void seg_offsets_kernel(const uint64_t* keys, uint64_t m, uint32_t n, uint64_t* seg) {
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v <= n; v += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t target = v << 32;
    uint64_t lo = 0, hi = m;
    while (lo < hi) {
      const uint64_t mid = (lo + hi) >> 1;
      if (keys[mid] < target) lo = mid + 1;
      else hi = mid;
    }
    seg[v] = lo;
  }
}

__global__ void max_row_kernel(const uint64_t* off, uint32_t n, unsigned long long* out) {
  unsigned long long mx = 0;
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x) {
    const unsigned long long l = off[v + 1] - off[v];
    mx = l > mx ? l : mx;
  }
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned long long other = __shfl_xor(mx, o);
    mx = other > mx ? other : mx;
  }
  if ((threadIdx.x & 63) == 0 && mx) atomicMax(out, mx);
}

__global__ void switch_test_kernel(const uint32_t* tree_size, const uint32_t* label_count, uint32_t R, double ratio,
                                   unsigned long long* flag) {
  for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < R; r += gridDim.x * blockDim.x) {
    // akiba.hpp:297 divides tree size by label count as floats
    const double t = (double)tree_size[r], l = (double)label_count[r];
    if (t / l > ratio) atomicOr(flag, 1ull);
  }
}

__global__ void merge_offsets_kernel(const uint64_t* old_off, const uint64_t* excl, uint32_t n, uint64_t* new_off) {
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v <= n; v += (uint64_t)gridDim.x * blockDim.x)
    new_off[v] = old_off[v] + excl[v];
}

// In-place merge of a batch into the store, one block of vertices [a, b) at a
// time from the last block down: the block's old rows were first copied to tmp
// (they start at old_off[a]); row v goes to new_off[v] >= old_off[v] followed by
// the batch's entries of v. Later blocks are already in place and earlier rows
// lie below new_off[a], so no unread entry is overwritten.
__global__ __launch_bounds__(256) void place_rows_kernel(const uint64_t* old_off, const uint64_t* tmp, uint64_t tmp_base,
                                                         const uint64_t* new_off, const uint64_t* excl,
                                                         const uint64_t* bat, uint32_t a, uint32_t b, uint64_t* ent) {
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t w0 = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) >> 6;
  const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t v = a + w0; v < b; v += nw) {
    const uint64_t ob = old_off[v] - tmp_base, len = old_off[v + 1] - old_off[v], nb = new_off[v];
    for (uint64_t j = lane; j < len; j += 64) ent[nb + j] = tmp[ob + j];
    const uint64_t bb = excl[v], be = excl[v + 1];
    for (uint64_t j = lane; j < be - bb; j += 64) ent[nb + len + j] = bat[bb + j];
  }
}

// Order-independent digest of a label side: the sum over all entries (hub, d) of
// vertex v's row of a 64-bit mix of (v, hub, d) (the splitmix64 finalizer).
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__global__ __launch_bounds__(256) void checksum_kernel(const uint64_t* off, const uint64_t* ent, const uint32_t* row_of,
                                                       uint32_t n, unsigned long long* out) {
  unsigned long long acc = 0;
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t w0 = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) >> 6;
  const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t v = w0; v < n; v += nw) {
    const uint32_t r = row_of[v];
    const uint64_t key = (uint64_t)v * 0x9E3779B97F4A7C15ull;
    for (uint64_t k = off[r] + lane; k < off[r + 1]; k += 64) acc += mix64(ent[k] ^ key);
  }
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
  if (lane == 0 && acc) atomicAdd(out, acc);
}

__global__ void copy_u64_kernel(const uint64_t* src, uint64_t count, uint64_t* dst) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < count; i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// clean_cover (shared_memory/labeling.hpp:78-88) of candidate k = (v, h, d):
// merge v's batch labels below h with the opposite-side batch labels of h below
// h; covered if a common hub x gives L(v)[x] + L(h)[x] <= d.
__global__ void clean_kernel(const uint64_t* key, const uint64_t* ent, uint64_t m, const uint64_t* seg_off,
                             const uint64_t* oent, const uint64_t* oseg_off, uint8_t* keep, unsigned long long* ctr) {
  unsigned long long reads = 0;
  for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < m; k += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t v = (uint32_t)(key[k] >> 32);
    const uint64_t e = ent[k];
    const uint32_t h = (uint32_t)(e >> 32);
    const uint32_t d = (uint32_t)e;
    uint64_t i = seg_off[v], ie = k;
    uint64_t j = oseg_off[h], je = oseg_off[h + 1];
    bool cov = false;
    while (i < ie && j < je) {
      const uint64_t a = ent[i], b = oent[j];
      const uint32_t ha = (uint32_t)(a >> 32), hb = (uint32_t)(b >> 32);
      ++reads;
      if (ha >= h || hb >= h) break;
      if (ha == hb) {
        if ((uint32_t)a + (uint32_t)b <= d) {
          cov = true;
          break;
        }
        ++i;
        ++j;
      } else if (ha < hb) {
        ++i;
      } else {
        ++j;
      }
    }
    keep[k] = cov ? 0 : 1;
  }
  if (reads) atomicAdd(&ctr[CTR_PRUNE_ENT], reads);
}

// Query: one wavefront per pair. The side-1 row of u is staged into LDS (in
// tiles of kQT entries), each lane takes entries of the side-0 row of v and
// binary-searches their hub among the staged hubs; the answer is the minimum
// sum over common hubs, INT_MAX when there is none (labeling.hpp:56-65).
// Every lane issues all its loads of a tile (kQPer entries) and of v's chunk
// before it uses any of them, so a pair costs a few memory latencies, not one
// per 64 entries, and its kQPer searches advance in lock step (their LDS reads
// overlap).
constexpr int kQWaves = 4;
constexpr int kQT = 1024;
constexpr int kQPer = kQT / 64;

__device__ __forceinline__ void load_chunk(const uint64_t* ent, uint64_t base, uint64_t end, uint32_t lane,
                                           uint64_t (&e)[kQPer]) {
#pragma unroll
  for (int i = 0; i < kQPer; ++i) {
    const uint64_t j = base + lane + 64u * i;
    e[i] = j < end ? ent[j] : ~0ull;  // hub all ones: past every staged hub
  }
}

// rng[2 v], rng[2 v + 1] = the bounds of vertex v's row (row row_of[v]).
__global__ void row_ranges_kernel(const uint64_t* off, const uint32_t* row_of, uint32_t n, uint64_t* rng) {
  for (size_t v = blockIdx.x * (size_t)blockDim.x + threadIdx.x; v < n; v += (size_t)gridDim.x * blockDim.x) {
    const uint32_t r = row_of[v];
    reinterpret_cast<ulonglong2*>(rng)[v] = make_ulonglong2(off[r], off[r + 1]);
  }
}

// rng1 / rng0: per vertex id the (begin, end) of its side-1 / side-0 row, so a
// pair reaches its rows with one 16-byte load per endpoint.
__global__ __launch_bounds__(256) void query_kernel(const uint64_t* rng1, const uint64_t* ent1, const uint64_t* rng0,
                                                    const uint64_t* ent0, const uint32_t* pairs, uint32_t n,
                                                    uint64_t q, int32_t* out) {
  __shared__ uint32_t s_h[kQWaves][kQT];
  __shared__ uint32_t s_d[kQWaves][kQT];
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t k = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) >> 6; k < q; k += nw) {
    const uint32_t pu = pairs[2 * k], pv = pairs[2 * k + 1];
    if (pu >= n || pv >= n) {
      if (lane == 0) out[k] = -1;
      continue;
    }
    const ulonglong2 ru = reinterpret_cast<const ulonglong2*>(rng1)[pu];
    const ulonglong2 rv = reinterpret_cast<const ulonglong2*>(rng0)[pv];
    const uint64_t ab = ru.x, ae = ru.y, bb = rv.x, be = rv.y;
    uint32_t best = kInfDist;
    for (uint64_t t0 = ab; t0 < ae; t0 += kQT) {
      const uint32_t tl = (uint32_t)((ae - t0) < kQT ? (ae - t0) : kQT);
      uint64_t eu[kQPer], ev[kQPer];
      load_chunk(ent1, t0, ae, lane, eu);
      load_chunk(ent0, bb, be, lane, ev);
#pragma unroll
      for (int i = 0; i < kQPer; ++i) {
        const uint32_t j = lane + 64u * i;
        if (j < tl) {
          s_h[wave][j] = (uint32_t)(eu[i] >> 32);
          s_d[wave][j] = (uint32_t)eu[i];
        }
      }
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      const uint32_t lo_h = s_h[wave][0], hi_h = s_h[wave][tl - 1];
      for (uint64_t c0 = bb;;) {
        // kQPer branch-free lower-bound searches in lock step: every step is one
        // LDS read per search, all independent, so their latencies overlap
        uint32_t base[kQPer];
        bool past = true;
#pragma unroll
        for (int i = 0; i < kQPer; ++i) {
          base[i] = 0;
          past = past && (uint32_t)(ev[i] >> 32) > hi_h;
        }
        for (uint32_t len = tl; len > 1;) {
          const uint32_t half = len >> 1;
#pragma unroll
          for (int i = 0; i < kQPer; ++i) {
            const uint32_t h = (uint32_t)(ev[i] >> 32);
            base[i] = s_h[wave][base[i] + half] < h ? base[i] + half : base[i];
          }
          len -= half;
        }
#pragma unroll
        for (int i = 0; i < kQPer; ++i) {
          const uint32_t h = (uint32_t)(ev[i] >> 32);
          // base = last index with hub < h (or 0); the match, if any, is there or next
          const uint32_t at = s_h[wave][base[i]] < h ? base[i] + 1 : base[i];
          if (h >= lo_h && h <= hi_h && at < tl && s_h[wave][at] == h) {
            const uint32_t sum = s_d[wave][at] + (uint32_t)ev[i];
            best = sum < best ? sum : best;
          }
        }
        // the side-0 row is hub-ascending: stop once every lane is past this tile
        c0 += kQT;
        if (__all(past) || c0 >= be) break;
        load_chunk(ent0, c0, be, lane, ev);
      }
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
    // wave minimum
    for (int o = 32; o > 0; o >>= 1) {
      const uint32_t other = __shfl_xor(best, o);
      best = other < best ? other : best;
    }
    if (lane == 0) out[k] = (int32_t)best;
  }
}

static unsigned grid_for(uint64_t items, unsigned per_block, unsigned cap) {
  uint64_t b = (items + per_block - 1) / per_block;
  if (b == 0) b = 1;
  return (unsigned)(b < cap ? b : cap);
}

}  // namespace

hipError_t launch_seg_offsets(const uint64_t* keys, uint64_t m, uint32_t n, uint64_t* seg, hipStream_t s) {
  hipLaunchKernelGGL(seg_offsets_kernel, dim3(grid_for((uint64_t)n + 1, 256, 16384)), dim3(256), 0, s, keys, m, n,
                     seg);
  return hipGetLastError();
}

hipError_t launch_max_row(const uint64_t* off, uint32_t n, unsigned long long* out, hipStream_t s) {
  hipLaunchKernelGGL(max_row_kernel, dim3(grid_for(n, 256, 4096)), dim3(256), 0, s, off, n, out);
  return hipGetLastError();
}

hipError_t launch_switch_test(const uint32_t* tree_size, const uint32_t* label_count, uint32_t R, double ratio,
                              unsigned long long* flag, hipStream_t s) {
  if (!R) return hipSuccess;
  hipLaunchKernelGGL(switch_test_kernel, dim3(grid_for(R, 256, 4096)), dim3(256), 0, s, tree_size, label_count, R,
                     ratio, flag);
  return hipGetLastError();
}

hipError_t launch_merge_offsets(const uint64_t* old_off, const uint64_t* excl, uint32_t n, uint64_t* new_off,
                                hipStream_t s) {
  hipLaunchKernelGGL(merge_offsets_kernel, dim3(grid_for((uint64_t)n + 1, 256, 16384)), dim3(256), 0, s, old_off,
                     excl, n, new_off);
  return hipGetLastError();
}


hipError_t launch_place_rows(const uint64_t* old_off, const uint64_t* tmp, uint64_t tmp_base, const uint64_t* new_off,
                             const uint64_t* excl, const uint64_t* bat, uint32_t a, uint32_t b, uint64_t* ent,
                             hipStream_t s) {
  if (b <= a) return hipSuccess;
  hipLaunchKernelGGL(place_rows_kernel, dim3(grid_for((uint64_t)(b - a) * 64, 256, 65536)), dim3(256), 0, s, old_off,
                     tmp, tmp_base, new_off, excl, bat, a, b, ent);
  return hipGetLastError();
}

hipError_t launch_checksum(const uint64_t* off, const uint64_t* ent, const uint32_t* row_of, uint32_t n,
                           unsigned long long* out, hipStream_t s) {
  hipLaunchKernelGGL(checksum_kernel, dim3(grid_for((uint64_t)n * 64, 256, 16384)), dim3(256), 0, s, off, ent, row_of,
                     n, out);
  return hipGetLastError();
}

hipError_t launch_copy_u64(const uint64_t* src, uint64_t count, uint64_t* dst, hipStream_t s) {
  if (!count) return hipSuccess;
  hipLaunchKernelGGL(copy_u64_kernel, dim3(grid_for(count, 256, 16384)), dim3(256), 0, s, src, count, dst);
  return hipGetLastError();
}

hipError_t launch_clean(const uint64_t* key, const uint64_t* ent, uint64_t m, const uint64_t* seg_off,
                        const uint64_t* oent, const uint64_t* oseg_off, uint8_t* keep, unsigned long long* ctr,
                        hipStream_t s) {
  if (!m) return hipSuccess;
  hipLaunchKernelGGL(clean_kernel, dim3(grid_for(m, 256, 16384)), dim3(256), 0, s, key, ent, m, seg_off, oent,
                     oseg_off, keep, ctr);
  return hipGetLastError();
}

hipError_t launch_query(const uint64_t* rng1, const uint64_t* ent1, const uint64_t* rng0, const uint64_t* ent0,
                        const uint32_t* pairs, uint32_t n, uint64_t q, int32_t* out, hipStream_t s) {
  if (!q) return hipSuccess;
  hipLaunchKernelGGL(query_kernel, dim3(grid_for(q * 64, 256, 8192)), dim3(256), 0, s, rng1, ent1, rng0, ent0, pairs, n,
                     q, out);
  return hipGetLastError();
}

hipError_t launch_row_ranges(const uint64_t* off, const uint32_t* row_of, uint32_t n, uint64_t* rng, hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(row_ranges_kernel, dim3(grid_for(n, 256, 8192)), dim3(256), 0, s, off, row_of, n, rng);
  return hipGetLastError();
}

// ---- library primitives (rocPRIM through hipCUB) ------------------------------

hipError_t sort_pairs_u64(void* tmp, size_t& tmp_bytes, const uint64_t* kin, uint64_t* kout, const uint64_t* vin,
                          uint64_t* vout, uint64_t m, int end_bit, hipStream_t s) {
  return hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, kin, kout, vin, vout, (int)m, 0, end_bit, s);
}

hipError_t exclusive_scan_u32_u64(void* tmp, size_t& tmp_bytes, const uint32_t* in, uint64_t* out, uint64_t m,
                                  hipStream_t s) {
  return hipcub::DeviceScan::ExclusiveScan(tmp, tmp_bytes, in, out, hipcub::Sum(), (uint64_t)0, (int)m, s);
}

hipError_t select_flagged_u64(void* tmp, size_t& tmp_bytes, const uint64_t* in, const uint8_t* flags, uint64_t* out,
                              uint64_t* nsel, uint64_t m, hipStream_t s) {
  return hipcub::DeviceSelect::Flagged(tmp, tmp_bytes, in, flags, out, nsel, (int)m, s);
}

}  // namespace chl
