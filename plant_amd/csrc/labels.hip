// Label-store kernels: building a batch's segment (count, scan, scatter, per-row
// sort), in-batch cleaning and compaction, merging segments, digests, and the
// merge-join distance query over segmented sides.
//
// A batch covers a contiguous hub-rank range above every hub already stored, as
// a superstep of shared_memory/akiba.hpp:152-208 does, so its labels form one
// new segment: each row is the batch's part of a vertex's label, hub-ascending.
// The segment is built in O(batch): the tree kernel writes each label's row
// and (hub, dist); cnt[row] counts them, a scan gives the segment's row
// offsets, a scatter places every label in its row, and a per-row sort puts
// the row in hub order (labels of one row arrive from different trees in any
// order). No row of the existing store moves.
#include "dev.hpp"

namespace chl {
namespace {

constexpr uint32_t kInfDist = 2147483647u;

unsigned grid_for(uint64_t items, unsigned per_block, unsigned cap) {
  uint64_t b = (items + per_block - 1) / per_block;
  if (b == 0) b = 1;
  return (unsigned)(b < cap ? b : cap);
}

// ---- count, scan, scatter ------------------------------------------------------

__global__ void count_rows_kernel(const uint32_t* row, uint64_t m, uint32_t* cnt) {
  for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < m; k += (uint64_t)gridDim.x * blockDim.x)
    atomicAdd(&cnt[row[k]], 1u);
}

// Exclusive scan in tiles of kScanTile items: a block per tile, each thread
// owning kScanPer consecutive items.
constexpr int kScanThreads = 256;
constexpr int kScanPer = 8;
constexpr uint64_t kScanTile = (uint64_t)kScanThreads * kScanPer;

__device__ __forceinline__ uint64_t wave_incl_scan(uint64_t x, uint32_t lane) {
  for (uint32_t o = 1; o < 64; o <<= 1) {
    const uint64_t y = __shfl_up(x, o);
    if (lane >= o) x += y;
  }
  return x;
}

// Block-wide exclusive scan of one value per thread; *total = the block sum.
__device__ __forceinline__ uint64_t block_excl_scan(uint64_t x, uint64_t* s_w, uint64_t* total) {
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t incl = wave_incl_scan(x, lane);
  if (lane == 63) s_w[wave] = incl;
  __syncthreads();
  uint64_t before = 0, sum = 0;
  for (uint32_t w = 0; w < kScanThreads / 64; ++w) {
    if (w < wave) before += s_w[w];
    sum += s_w[w];
  }
  __syncthreads();
  *total = sum;
  return before + incl - x;
}

template <class T>
__global__ __launch_bounds__(kScanThreads) void tile_sums_kernel(const T* in, uint64_t count, uint64_t* sums) {
  __shared__ uint64_t s_w[kScanThreads / 64];
  const uint64_t base = blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanPer;
  uint64_t x = 0;
#pragma unroll
  for (int i = 0; i < kScanPer; ++i)
    if (base + i < count) x += (uint64_t)in[base + i];
  uint64_t total = 0;
  (void)block_excl_scan(x, s_w, &total);
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// out[i] = prefix[tile] + the exclusive scan inside the tile; the block of the
// last tile also writes out[count], the total.
template <class T>
__global__ __launch_bounds__(kScanThreads) void tile_scan_kernel(const T* in, uint64_t count, const uint64_t* prefix,
                                                                 uint64_t* out) {
  __shared__ uint64_t s_w[kScanThreads / 64];
  const uint64_t base = blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanPer;
  uint64_t v[kScanPer];
  uint64_t x = 0;
#pragma unroll
  for (int i = 0; i < kScanPer; ++i) {
    v[i] = base + i < count ? (uint64_t)in[base + i] : 0;
    x += v[i];
  }
  uint64_t total = 0;
  uint64_t run = block_excl_scan(x, s_w, &total) + (prefix ? prefix[blockIdx.x] : 0);
#pragma unroll
  for (int i = 0; i < kScanPer; ++i) {
    if (base + i < count) out[base + i] = run;
    run += v[i];
  }
  if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) out[count] = (prefix ? prefix[blockIdx.x] : 0) + total;
}

template <class T>
hipError_t scan_any(const T* in, uint64_t count, uint64_t* out, uint64_t* tmp, hipStream_t s) {
  if (count == 0) return launch_fill_u64(out, 0, 1, s);
  const uint64_t tiles = (count + kScanTile - 1) / kScanTile;
  if (tiles == 1) {
    hipLaunchKernelGGL((tile_scan_kernel<T>), dim3(1), dim3(kScanThreads), 0, s, in, count, nullptr, out);
    return hipGetLastError();
  }
  uint64_t* sums = tmp;
  uint64_t* prefix = tmp + tiles;
  hipLaunchKernelGGL((tile_sums_kernel<T>), dim3((unsigned)tiles), dim3(kScanThreads), 0, s, in, count, sums);
  hipError_t e = scan_any<uint64_t>(sums, tiles, prefix, tmp + 2 * tiles + 1, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((tile_scan_kernel<T>), dim3((unsigned)tiles), dim3(kScanThreads), 0, s, in, count, prefix, out);
  return hipGetLastError();
}

__global__ void scatter_labels_kernel(const uint32_t* row, const uint64_t* hd, uint64_t m, unsigned long long* cur,
                                      uint32_t fmt, uint32_t h0, uint32_t* key, uint8_t* lo, uint64_t* ent) {
  for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < m; k += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t e = hd[k];
    const uint64_t pos = atomicAdd(&cur[row[k]], 1ull);
    if (fmt == SEG_F8) {
      ent[pos] = e;
    } else {
      const uint32_t h = (uint32_t)(e >> 32), d = (uint32_t)e;
      key[pos] = ((h - h0) << 8) | (d >> 8);
      lo[pos] = (uint8_t)(d & 0xFFu);
    }
  }
}

// ---- per-row sort ----------------------------------------------------------------
// The sort key of an entry is a 64-bit composite with the hub in its high bits
// (hubs are distinct within a row): the entry itself in SEG_F8, key << 8 | lo
// in SEG_F5.

__device__ __forceinline__ uint64_t comp_get(uint32_t fmt, const uint32_t* key, const uint8_t* lo, const uint64_t* ent,
                                             uint64_t k) {
  return fmt == SEG_F8 ? ent[k] : (((uint64_t)key[k] << 8) | lo[k]);
}
__device__ __forceinline__ void comp_put(uint32_t fmt, uint32_t* key, uint8_t* lo, uint64_t* ent, uint64_t k,
                                         uint64_t x) {
  if (fmt == SEG_F8) {
    ent[k] = x;
  } else {
    key[k] = (uint32_t)(x >> 8);
    lo[k] = (uint8_t)(x & 0xFFu);
  }
}

constexpr uint64_t kSmallRow = 16;  // rows up to this length: insertion sort by one thread
constexpr uint32_t kLdsSort = 4096;  // rows up to this length: bitonic sort in LDS

// One thread per row: short rows sorted in place, longer rows listed in
// longq[1..] (longq[0] counts them; zeroed first).
__global__ void sort_small_rows_kernel(const uint64_t* off, uint32_t rows, uint32_t fmt, uint32_t* key, uint8_t* lo,
                                       uint64_t* ent, uint32_t* longq) {
  for (uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; r < rows; r += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t b = off[r], len = off[r + 1] - b;
    if (len <= 1) continue;
    if (len > kSmallRow) {
      longq[1 + atomicAdd(&longq[0], 1u)] = (uint32_t)r;
      continue;
    }
    for (uint64_t i = 1; i < len; ++i) {
      const uint64_t x = comp_get(fmt, key, lo, ent, b + i);
      uint64_t j = i;
      while (j > 0) {
        const uint64_t y = comp_get(fmt, key, lo, ent, b + j - 1);
        if (y <= x) break;
        comp_put(fmt, key, lo, ent, b + j, y);
        --j;
      }
      if (j != i) comp_put(fmt, key, lo, ent, b + j, x);
    }
  }
}

// Bitonic network in the form where every comparison puts the smaller value at
// the lower index: for each block size k the first step compares i with its
// mirror i ^ (k - 1) inside the block, the later steps i with i ^ j. Positions
// at or past len hold +infinity, which never moves down, so a row of any length
// sorts in place within a power-of-two span P.
// Pair t of a step: its lower index i and partner.
__device__ __forceinline__ void bitonic_pair(uint32_t t, uint32_t k, uint32_t j, uint32_t& i, uint32_t& part) {
  i = 2 * j * (t / j) + (t % j);
  part = (j == k / 2) ? (i ^ (k - 1)) : (i + j);
}

__global__ __launch_bounds__(256) void sort_long_rows_kernel(const uint64_t* off, uint32_t fmt, uint32_t* key,
                                                             uint8_t* lo, uint64_t* ent, const uint32_t* longq) {
  __shared__ uint64_t s[kLdsSort];
  const uint32_t tid = threadIdx.x;
  const uint32_t count = longq[0];
  for (uint32_t q = blockIdx.x; q < count; q += gridDim.x) {
    const uint32_t r = longq[1 + q];
    const uint64_t b = off[r];
    const uint32_t len = (uint32_t)(off[r + 1] - b);
    uint32_t P = 1;
    while (P < len) P <<= 1;
    if (len <= kLdsSort) {
      for (uint32_t t = tid; t < P; t += 256) s[t] = t < len ? comp_get(fmt, key, lo, ent, b + t) : ~0ull;
      __syncthreads();
      for (uint32_t k = 2; k <= P; k <<= 1)
        for (uint32_t j = k / 2; j > 0; j >>= 1) {
          for (uint32_t t = tid; t < P / 2; t += 256) {
            uint32_t i, part;
            bitonic_pair(t, k, j, i, part);
            const uint64_t a = s[i], c = s[part];
            if (a > c) {
              s[i] = c;
              s[part] = a;
            }
          }
          __syncthreads();
        }
      for (uint32_t t = tid; t < len; t += 256) comp_put(fmt, key, lo, ent, b + t, s[t]);
    } else {
      // in global memory (rare: a batch gives one row more than kLdsSort labels)
      for (uint32_t k = 2; k <= P; k <<= 1)
        for (uint32_t j = k / 2; j > 0; j >>= 1) {
          for (uint32_t t = tid; t < P / 2; t += 256) {
            uint32_t i, part;
            bitonic_pair(t, k, j, i, part);
            if (part >= len) continue;  // +infinity there: nothing moves
            const uint64_t a = comp_get(fmt, key, lo, ent, b + i), c = comp_get(fmt, key, lo, ent, b + part);
            if (a > c) {
              comp_put(fmt, key, lo, ent, b + i, c);
              comp_put(fmt, key, lo, ent, b + part, a);
            }
          }
          __syncthreads();
        }
    }
    __syncthreads();
  }
}

// ---- in-batch cleaning -----------------------------------------------------------
// clean_cover (shared_memory/labeling.hpp:78-88) over the batch labels. One
// wavefront per row v of this side: v's row (hub-ascending) is staged in LDS.
// For a candidate (h, d) of it the lanes take 64 entries at a time of row h of
// the other side's batch labels and look each hub x < h up among v's entries
// before the candidate (exactly v's batch hubs below h); the candidate is
// covered when some x gives L(v)[x] + L(h)[x] <= d. The first chunks of
// kCleanAhead candidates are loaded together, so their memory latencies
// overlap; a candidate still open after its first chunk walks on alone. A
// candidate that is kept has read its whole other row below h, as the
// reference's merge does: those reads dominate the kernel.
constexpr int kCleanWaves = 4;
constexpr uint32_t kCleanTile = 1024;
constexpr uint32_t kCleanAhead = 4;

__device__ __forceinline__ uint32_t lds_lower_bound(const uint32_t* a, uint32_t len, uint32_t x) {
  uint32_t lo = 0, hi = len;
  while (lo < hi) {
    const uint32_t mid = (lo + hi) >> 1;
    if (a[mid] < x) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Whether entry (x, dx) of the other row covers candidate c = (h, d): x is among
// v's entries before c with L(v)[x] + dx <= d.
__device__ __forceinline__ bool covers(const SegView& seg, uint64_t rb, bool staged, const uint32_t* sh,
                                       const uint32_t* sd, uint32_t c, uint32_t d, uint32_t x, uint32_t dx) {
  uint32_t hv = 0xFFFFFFFFu, dv = 0;
  if (staged) {
    const uint32_t at = lds_lower_bound(sh, c, x);
    if (at < c) {
      hv = sh[at];
      dv = sd[at];
    }
  } else {
    uint32_t lo = 0, hi = c;
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      uint32_t mh, md;
      seg_entry(seg, rb + mid, mh, md);
      if (mh < x) lo = mid + 1;
      else hi = mid;
    }
    if (lo < c) seg_entry(seg, rb + lo, hv, dv);
  }
  return hv == x && dv + dx <= d;
}

__global__ __launch_bounds__(256) void clean_kernel(SegView seg, SegView oseg, uint32_t rows, uint8_t* keep,
                                                    unsigned long long* ctr) {
  __shared__ uint32_t s_h[kCleanWaves][kCleanTile];
  __shared__ uint32_t s_d[kCleanWaves][kCleanTile];
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  unsigned long long reads = 0;
  for (uint64_t v = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) >> 6; v < rows; v += nw) {
    const uint64_t rb = seg.off[v], re = seg.off[v + 1];
    const uint32_t len = (uint32_t)(re - rb);
    if (len == 0) continue;
    const bool staged = len <= kCleanTile;
    const uint32_t* sh = s_h[wave];
    const uint32_t* sd = s_d[wave];
    if (staged) {
      for (uint32_t j = lane; j < len; j += 64) {
        uint32_t h, d;
        seg_entry(seg, rb + j, h, d);
        s_h[wave][j] = h;
        s_d[wave][j] = d;
      }
    }
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    for (uint32_t c0 = 0; c0 < len; c0 += 64) {
      // each lane fetches one candidate and the bounds of its other row
      uint32_t ch = 0, cd = 0;
      uint64_t ob = 0, oe = 0;
      if (c0 + lane < len) {
        if (staged) {
          ch = sh[c0 + lane];
          cd = sd[c0 + lane];
        } else {
          seg_entry(seg, rb + c0 + lane, ch, cd);
        }
        ob = oseg.off[ch];
        oe = oseg.off[ch + 1];
      }
      uint64_t kept = 0;  // keep bits of these candidates
      const uint32_t cn = min(64u, len - c0);
      for (uint32_t i0 = 0; i0 < cn; i0 += kCleanAhead) {
        // the first chunks of kCleanAhead candidates, loaded together
        uint32_t x[kCleanAhead], dx[kCleanAhead], h[kCleanAhead], d[kCleanAhead];
        uint64_t b[kCleanAhead], e[kCleanAhead];
        bool in[kCleanAhead];
#pragma unroll
        for (uint32_t k = 0; k < kCleanAhead; ++k) {
          const uint32_t i = min(i0 + k, cn - 1);
          h[k] = __shfl(ch, i);
          d[k] = __shfl(cd, i);
          b[k] = __shfl(ob, i);
          e[k] = __shfl(oe, i);
          in[k] = i0 + k < cn && b[k] + lane < e[k];
          x[k] = 0xFFFFFFFFu;
          dx[k] = 0;
          if (in[k]) seg_entry(oseg, b[k] + lane, x[k], dx[k]);
        }
#pragma unroll
        for (uint32_t k = 0; k < kCleanAhead; ++k) {
          if (i0 + k >= cn) break;
          const uint32_t c = c0 + i0 + k;
          reads += in[k] ? 1 : 0;
          bool past = x[k] >= h[k];
          bool covered = __any(!past && covers(seg, rb, staged, sh, sd, c, d[k], x[k], dx[k]));
          // the rest of the other row, 64 entries at a time
          for (uint64_t base = b[k] + 64; !covered && !__all(past) && base < e[k]; base += 64) {
            const uint64_t j = base + lane;
            uint32_t xx = 0xFFFFFFFFu, dd = 0;
            if (j < e[k]) {
              seg_entry(oseg, j, xx, dd);
              ++reads;
            }
            past = xx >= h[k];
            covered = __any(!past && covers(seg, rb, staged, sh, sd, c, d[k], xx, dd));
          }
          if (!covered) kept |= 1ull << (i0 + k);
        }
      }
      if (lane < cn) keep[rb + c0 + lane] = (uint8_t)((kept >> lane) & 1ull);
    }
    __builtin_amdgcn_wave_barrier();
  }
  for (int o = 32; o > 0; o >>= 1) reads += __shfl_xor(reads, o);
  if (lane == 0 && reads) atomicAdd(&ctr[CTR_PRUNE_ENT], reads);
}

// ---- compaction, merge -----------------------------------------------------------

__global__ __launch_bounds__(256) void count_kept_kernel(const uint64_t* off, const uint8_t* keep, uint32_t rows,
                                                         uint32_t* cnt) {
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t r = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) >> 6; r < rows; r += nw) {
    uint32_t c = 0;
    for (uint64_t k = off[r] + lane; k < off[r + 1]; k += 64) c += keep[k];
    for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
    if (lane == 0) cnt[r] = c;
  }
}

__device__ __forceinline__ void put_entry(uint32_t fmt, uint32_t h0, uint32_t* key, uint8_t* lo, uint64_t* ent,
                                          uint64_t k, uint32_t h, uint32_t d) {
  if (fmt == SEG_F8) {
    ent[k] = ((uint64_t)h << 32) | d;
  } else {
    key[k] = ((h - h0) << 8) | (d >> 8);
    lo[k] = (uint8_t)(d & 0xFFu);
  }
}

__global__ __launch_bounds__(256) void copy_kept_kernel(SegView seg, const uint8_t* keep, uint32_t rows,
                                                        const uint64_t* noff, uint32_t* key, uint8_t* lo,
                                                        uint64_t* ent) {
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  const uint64_t below = (1ull << lane) - 1;
  for (uint64_t r = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) >> 6; r < rows; r += nw) {
    uint64_t pos = noff[r];
    const uint64_t e = seg.off[r + 1];
    for (uint64_t base = seg.off[r]; base < e; base += 64) {
      const uint64_t k = base + lane;
      const bool kp = k < e && keep[k];
      const uint64_t mask = __ballot(kp);
      if (kp) {
        const uint64_t at = pos + __popcll(mask & below);
        if (seg.fmt == SEG_F8) {
          ent[at] = seg.ent[k];
        } else {
          key[at] = seg.key[k];
          lo[at] = seg.lo[k];
        }
      }
      pos += __popcll(mask);
    }
  }
}

__global__ void row_total_kernel(const SegView* views, uint32_t k, uint32_t rows, uint32_t* cnt) {
  for (uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; r < rows; r += (uint64_t)gridDim.x * blockDim.x) {
    uint64_t t = 0;
    for (uint32_t s = 0; s < k; ++s) t += views[s].off[r + 1] - views[s].off[r];
    cnt[r] = (uint32_t)t;
  }
}

__global__ __launch_bounds__(256) void concat_rows_kernel(const SegView* views, uint32_t k, uint32_t rows,
                                                          const uint64_t* noff, uint32_t fmt, uint32_t h0,
                                                          uint32_t* key, uint8_t* lo, uint64_t* ent) {
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t r = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) >> 6; r < rows; r += nw) {
    uint64_t pos = noff[r];
    for (uint32_t s = 0; s < k; ++s) {
      const SegView v = views[s];
      const uint64_t b = v.off[r], e = v.off[r + 1];
      for (uint64_t j = b + lane; j < e; j += 64) {
        uint32_t h, d;
        seg_entry(v, j, h, d);
        put_entry(fmt, h0, key, lo, ent, pos + (j - b), h, d);
      }
      pos += e - b;
    }
  }
}

__global__ void max_row_kernel(const SegView* views, uint32_t k, uint32_t rows, unsigned long long* out) {
  unsigned long long mx = 0;
  for (uint64_t r = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; r < rows; r += (uint64_t)gridDim.x * blockDim.x) {
    unsigned long long t = 0;
    for (uint32_t s = 0; s < k; ++s) t += views[s].off[r + 1] - views[s].off[r];
    mx = t > mx ? t : mx;
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

// ---- digest, ranges, decode ----------------------------------------------------

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
__global__ __launch_bounds__(256) void checksum_kernel(SegView seg, const uint32_t* row_of, uint32_t n,
                                                       unsigned long long* out) {
  unsigned long long acc = 0;
  const uint32_t lane = threadIdx.x & 63;
  const uint64_t w0 = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) >> 6;
  const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  for (uint64_t v = w0; v < n; v += nw) {
    const uint32_t r = row_of[v];
    const uint64_t key = (uint64_t)v * 0x9E3779B97F4A7C15ull;
    for (uint64_t k = seg.off[r] + lane; k < seg.off[r + 1]; k += 64) {
      uint32_t h, d;
      seg_entry(seg, k, h, d);
      acc += mix64((((uint64_t)h << 32) | d) ^ key);
    }
  }
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
  if (lane == 0 && acc) atomicAdd(out, acc);
}

__global__ void row_ranges_kernel(const SegView* views, uint32_t k, const uint32_t* row_of, uint32_t n,
                                  uint64_t* rng) {
  const uint64_t total = (uint64_t)n * k;
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < total; i += (uint64_t)gridDim.x * blockDim.x) {
    const uint64_t v = i / k;
    const uint32_t sg = (uint32_t)(i - v * k);
    const uint32_t r = row_of[v];
    const uint64_t* off = views[sg].off;
    reinterpret_cast<ulonglong2*>(rng)[i] = make_ulonglong2(off[r], off[r + 1]);
  }
}

// Entries of a row segment with hub in [hlo, hhi): [lo, hi) by two binary
// searches (the row is hub-ascending).
__device__ __forceinline__ void row_window(const SegView& s, uint64_t b, uint64_t e, uint32_t hlo, uint32_t hhi,
                                           uint64_t& lo, uint64_t& hi) {
  uint64_t a = b, c = e;
  while (a < c) {
    const uint64_t m = (a + c) >> 1;
    uint32_t h, d;
    seg_entry(s, m, h, d);
    if (h < hlo) a = m + 1;
    else c = m;
  }
  lo = a;
  c = e;
  while (a < c) {
    const uint64_t m = (a + c) >> 1;
    uint32_t h, d;
    seg_entry(s, m, h, d);
    if (h < hhi) a = m + 1;
    else c = m;
  }
  hi = a;
}

// cnt[v] = entries of vertex v's row (row row_of[v]) with hub in [hlo, hhi),
// over the k segments; with noff set, those entries are written from noff[v].
__global__ void window_kernel(const SegView* views, uint32_t k, const uint32_t* row_of, uint32_t n, uint32_t hlo,
                              uint32_t hhi, uint32_t* cnt, const uint64_t* noff, uint32_t* hub, int32_t* dist) {
  for (uint64_t v = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; v < n; v += (uint64_t)gridDim.x * blockDim.x) {
    const uint32_t r = row_of[v];
    uint64_t pos = noff ? noff[v] : 0, c = 0;
    for (uint32_t sg = 0; sg < k; ++sg) {
      const SegView s = views[sg];
      uint64_t lo, hi;
      row_window(s, s.off[r], s.off[r + 1], hlo, hhi, lo, hi);
      if (noff)
        for (uint64_t j = lo; j < hi; ++j, ++pos) {
          uint32_t h, d;
          seg_entry(s, j, h, d);
          hub[pos] = h;
          dist[pos] = (int32_t)d;
        }
      c += hi - lo;
    }
    if (!noff) cnt[v] = (uint32_t)c;
  }
}

__global__ void decode_kernel(SegView seg, uint64_t count, uint64_t* out) {
  for (uint64_t k = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; k < count; k += (uint64_t)gridDim.x * blockDim.x) {
    uint32_t h, d;
    seg_entry(seg, k, h, d);
    out[k] = ((uint64_t)h << 32) | d;
  }
}

__global__ void copy_u64_kernel(const uint64_t* src, uint64_t count, uint64_t* dst) {
  for (uint64_t i = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x; i < count; i += (uint64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

}  // namespace

// ---- launches ------------------------------------------------------------------

hipError_t launch_count_rows(const uint32_t* row, uint64_t m, uint32_t* cnt, hipStream_t s) {
  if (!m) return hipSuccess;
  hipLaunchKernelGGL(count_rows_kernel, dim3(grid_for(m, 256, 16384)), dim3(256), 0, s, row, m, cnt);
  return hipGetLastError();
}

size_t scan_tmp_words(uint64_t count) {
  const uint64_t tiles = (count + kScanTile - 1) / kScanTile;
  if (tiles <= 1) return 1;
  return (size_t)(2 * tiles + 1) + scan_tmp_words(tiles);
}

hipError_t launch_scan_u32(const uint32_t* in, uint64_t count, uint64_t* out, uint64_t* tmp, hipStream_t s) {
  return scan_any<uint32_t>(in, count, out, tmp, s);
}

hipError_t launch_scatter_labels(const uint32_t* row, const uint64_t* hd, uint64_t m, uint64_t* cur, uint32_t fmt,
                                 uint32_t h0, uint32_t* key, uint8_t* lo, uint64_t* ent, hipStream_t s) {
  if (!m) return hipSuccess;
  hipLaunchKernelGGL(scatter_labels_kernel, dim3(grid_for(m, 256, 16384)), dim3(256), 0, s, row, hd, m,
                     (unsigned long long*)cur, fmt, h0, key, lo, ent);
  return hipGetLastError();
}

hipError_t launch_sort_rows(const uint64_t* off, uint32_t rows, uint32_t fmt, uint32_t* key, uint8_t* lo,
                            uint64_t* ent, uint32_t* longq, hipStream_t s) {
  if (!rows) return hipSuccess;
  hipError_t e = launch_fill_u32(longq, 0, 1, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(sort_small_rows_kernel, dim3(grid_for(rows, 256, 16384)), dim3(256), 0, s, off, rows, fmt, key,
                     lo, ent, longq);
  hipLaunchKernelGGL(sort_long_rows_kernel, dim3(4096), dim3(256), 0, s, off, fmt, key, lo, ent, longq);
  return hipGetLastError();
}

hipError_t launch_clean(const SegView& seg, const SegView& oseg, uint32_t rows, uint8_t* keep,
                        unsigned long long* ctr, hipStream_t s) {
  if (!rows) return hipSuccess;
  hipLaunchKernelGGL(clean_kernel, dim3(grid_for((uint64_t)rows * 64, 256, 16384)), dim3(256), 0, s, seg, oseg, rows,
                     keep, ctr);
  return hipGetLastError();
}

hipError_t launch_count_kept(const uint64_t* off, const uint8_t* keep, uint32_t rows, uint32_t* cnt, hipStream_t s) {
  if (!rows) return hipSuccess;
  hipLaunchKernelGGL(count_kept_kernel, dim3(grid_for((uint64_t)rows * 64, 256, 16384)), dim3(256), 0, s, off, keep,
                     rows, cnt);
  return hipGetLastError();
}

hipError_t launch_copy_kept(const SegView& seg, const uint8_t* keep, uint32_t rows, const uint64_t* noff,
                            uint32_t* key, uint8_t* lo, uint64_t* ent, hipStream_t s) {
  if (!rows) return hipSuccess;
  hipLaunchKernelGGL(copy_kept_kernel, dim3(grid_for((uint64_t)rows * 64, 256, 16384)), dim3(256), 0, s, seg, keep,
                     rows, noff, key, lo, ent);
  return hipGetLastError();
}

hipError_t launch_row_total(const SegView* views, uint32_t k, uint32_t rows, uint32_t* cnt, hipStream_t s) {
  if (!rows) return hipSuccess;
  hipLaunchKernelGGL(row_total_kernel, dim3(grid_for(rows, 256, 16384)), dim3(256), 0, s, views, k, rows, cnt);
  return hipGetLastError();
}

hipError_t launch_concat_rows(const SegView* views, uint32_t k, uint32_t rows, const uint64_t* noff, uint32_t fmt,
                              uint32_t h0, uint32_t* key, uint8_t* lo, uint64_t* ent, hipStream_t s) {
  if (!rows) return hipSuccess;
  hipLaunchKernelGGL(concat_rows_kernel, dim3(grid_for((uint64_t)rows * 64, 256, 16384)), dim3(256), 0, s, views, k,
                     rows, noff, fmt, h0, key, lo, ent);
  return hipGetLastError();
}

hipError_t launch_max_row(const SegView* views, uint32_t k, uint32_t rows, unsigned long long* out, hipStream_t s) {
  if (!rows || !k) return hipSuccess;
  hipLaunchKernelGGL(max_row_kernel, dim3(grid_for(rows, 256, 4096)), dim3(256), 0, s, views, k, rows, out);
  return hipGetLastError();
}

hipError_t launch_switch_test(const uint32_t* tree_size, const uint32_t* label_count, uint32_t R, double ratio,
                              unsigned long long* flag, hipStream_t s) {
  if (!R) return hipSuccess;
  hipLaunchKernelGGL(switch_test_kernel, dim3(grid_for(R, 256, 4096)), dim3(256), 0, s, tree_size, label_count, R,
                     ratio, flag);
  return hipGetLastError();
}

hipError_t launch_checksum(const SegView& seg, const uint32_t* row_of, uint32_t n, unsigned long long* out,
                           hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(checksum_kernel, dim3(grid_for((uint64_t)n * 64, 256, 16384)), dim3(256), 0, s, seg, row_of, n,
                     out);
  return hipGetLastError();
}

hipError_t launch_row_ranges(const SegView* views, uint32_t k, const uint32_t* row_of, uint32_t n, uint64_t* rng,
                             hipStream_t s) {
  if (!n || !k) return hipSuccess;
  hipLaunchKernelGGL(row_ranges_kernel, dim3(grid_for((uint64_t)n * k, 256, 8192)), dim3(256), 0, s, views, k, row_of,
                     n, rng);
  return hipGetLastError();
}

hipError_t launch_window(const SegView* views, uint32_t k, const uint32_t* row_of, uint32_t n, uint32_t hlo,
                         uint32_t hhi, uint32_t* cnt, const uint64_t* noff, uint32_t* hub, int32_t* dist,
                         hipStream_t s) {
  if (!n || !k) return hipSuccess;
  hipLaunchKernelGGL(window_kernel, dim3(grid_for(n, 256, 16384)), dim3(256), 0, s, views, k, row_of, n, hlo, hhi, cnt,
                     noff, hub, dist);
  return hipGetLastError();
}

hipError_t launch_decode(const SegView& seg, uint64_t count, uint64_t* out, hipStream_t s) {
  if (!count) return hipSuccess;
  hipLaunchKernelGGL(decode_kernel, dim3(grid_for(count, 256, 16384)), dim3(256), 0, s, seg, count, out);
  return hipGetLastError();
}

hipError_t launch_copy_u64(const uint64_t* src, uint64_t count, uint64_t* dst, hipStream_t s) {
  if (!count) return hipSuccess;
  hipLaunchKernelGGL(copy_u64_kernel, dim3(grid_for(count, 256, 16384)), dim3(256), 0, s, src, count, dst);
  return hipGetLastError();
}

}  // namespace chl
