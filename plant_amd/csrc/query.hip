// The distance query as a merge-join kernel over device-resident labels:
// Labeling::query (shared_memory/labeling.hpp:56-65) for a batch of pairs, as
// batchLocalQuery runs it (distributed_memory/query_funcs.hpp:23-30):
// d(u, v) = min over hubs h common to L1(u) and L0(v) of L1(u)[h] + L0(v)[h],
// INT_MAX when they share none.
//
// One wavefront per pair. Lane s fetches the bounds of u's side-1 row and v's
// side-0 row in segment s (one vertex-major table per side, so every segment's
// bounds arrive in one round trip). A segment holds a contiguous hub range, so
// the pair's answer is the minimum over segments of the per-segment merge-join.
// Per segment where both rows are non-empty, u's row is staged into LDS in
// tiles of up to kQT entries; v's row is read in chunks of kQT entries, kQPer
// per lane, and every entry's hub is looked up among the staged hubs by a
// branch-free lower-bound search. Every lane issues all its loads of a tile and
// of a chunk before it uses any of them, and its kQPer searches advance in lock
// step, so their LDS reads overlap.
//
// Both rows are hub-ascending, so a tile ends its scan of v's row at the first
// chunk lying wholly past the tile's last hub, and the next tile starts at the
// first chunk that held an entry past it: v's row is read about once however
// many tiles u's row takes. When no entry of v's row lies past a tile, the
// later tiles cannot match and the segment ends there.
//
// Chunk layout: a SEG_F8 chunk starting at entry c covers [c, c + kQT), entry
// c + lane + 64 i in slot i of a lane. A SEG_F5 chunk covers [c & ~3, that +
// kQT): a lane loads kQPer / 4 aligned 4-entry groups (16 bytes of keys and 4
// bytes of low distance bytes each), entry (c & ~3) + 4 lane + 256 j + t in
// slot 4 j + t. Slots outside [c, end) hold a hub of all ones (past every hub).
// Segment fields come from the kernel argument with a wavefront-uniform index
// (scalar loads).
#include "dev.hpp"

// Tile (and chunk) entries: 512 entries of LDS (4 KB) per wavefront. Against
// 1024 it doubles the workgroups an LDS-limited compute unit holds; on the
// 1M-vertex road-like labels it served 3.4e8 pairs/s against 3.3e8 (one
// segment) and 2.7e8 against 1.8e8 (seven segments, shorter rows per segment).
#ifndef CHL_QUERY_TILE
#define CHL_QUERY_TILE 512
#endif

namespace chl {
namespace {

constexpr uint32_t kQInf = 2147483647u;
constexpr int kQWaves = 4;
constexpr int kQT = CHL_QUERY_TILE;
constexpr int kQPer = kQT / 64;
static_assert(kQPer % 4 == 0, "a lane loads whole 4-entry groups of a SEG_F5 chunk");

__device__ __forceinline__ uint64_t chunk_base(uint32_t fmt, uint64_t c) { return fmt == SEG_F8 ? c : (c & ~3ull); }
__device__ __forceinline__ uint64_t slot_entry(uint32_t fmt, uint64_t base, uint32_t lane, int i) {
  return fmt == SEG_F8 ? base + lane + 64u * i : base + 4u * lane + 256u * (i >> 2) + (i & 3);
}

// The chunk of side S's segment sg starting at entry c, as (hub << 32) | dist
// per slot; all ones outside [c, end).
template <int FMT>
__device__ __forceinline__ void load_chunk(const QuerySide& S, uint32_t sg, uint64_t c, uint64_t end, uint32_t lane,
                                           uint64_t (&e)[kQPer]) {
  const uint32_t fmt = FMT ? (uint32_t)FMT : S.fmt[sg];
  if (fmt == SEG_F8) {
    const uint64_t* ent = S.ent[sg];
#pragma unroll
    for (int i = 0; i < kQPer; ++i) {
      const uint64_t j = c + lane + 64u * i;
      e[i] = j < end ? ent[j] : ~0ull;
    }
  } else {
    const uint4* key4 = reinterpret_cast<const uint4*>(S.key[sg]);
    const uint32_t* lo4 = reinterpret_cast<const uint32_t*>(S.lo[sg]);
    const uint32_t h0 = S.h0[sg];
    const uint64_t g0 = (c >> 2) + lane;
    uint4 kk[kQPer / 4];
    uint32_t ll[kQPer / 4];
#pragma unroll
    for (int j = 0; j < kQPer / 4; ++j) {
      const uint64_t g = g0 + 64u * j;
      // a group that starts before `end` lies within the arrays (kF5Pad spare entries)
      if (4 * g < end) {
        kk[j] = key4[g];
        ll[j] = lo4[g];
      } else {
        kk[j] = make_uint4(0, 0, 0, 0);
        ll[j] = 0;
      }
    }
#pragma unroll
    for (int j = 0; j < kQPer / 4; ++j) {
      const uint64_t g = g0 + 64u * j;
      const uint32_t x[4] = {kk[j].x, kk[j].y, kk[j].z, kk[j].w};
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        const uint64_t idx = 4 * g + t;
        e[4 * j + t] = (idx >= c && idx < end)
                           ? (((uint64_t)(h0 + (x[t] >> 8)) << 32) | ((x[t] & 0xFFu) << 8) | ((ll[j] >> (8 * t)) & 0xFFu))
                           : ~0ull;
      }
    }
  }
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int FMT>
__global__ __launch_bounds__(256) void query_kernel(const QueryView qv, const uint32_t* pairs, uint32_t n, uint64_t q,
                                                    int32_t* out) {
  __shared__ uint32_t s_h[kQWaves][kQT];
  __shared__ uint32_t s_d[kQWaves][kQT];
  const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint64_t nw = ((uint64_t)gridDim.x * blockDim.x) >> 6;
  const uint32_t ks = qv.nseg;
  for (uint64_t k = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) >> 6; k < q; k += nw) {
    const uint32_t pu = pairs[2 * k], pv = pairs[2 * k + 1];
    if (pu >= n || pv >= n) {
      if (lane == 0) out[k] = -1;
      continue;
    }
    ulonglong2 ru = make_ulonglong2(0, 0), rv = make_ulonglong2(0, 0);
    if (lane < ks) {
      ru = reinterpret_cast<const ulonglong2*>(qv.s1.rng)[(uint64_t)pu * ks + lane];
      rv = reinterpret_cast<const ulonglong2*>(qv.s0.rng)[(uint64_t)pv * ks + lane];
    }
    uint64_t segs = __ballot(ru.x < ru.y && rv.x < rv.y);
    uint32_t best = kQInf;
    while (segs) {
      const uint32_t sg = __builtin_amdgcn_readfirstlane((uint32_t)__ffsll((long long)segs) - 1);
      segs &= segs - 1;
      const uint64_t ab = __shfl(ru.x, sg), ae = __shfl(ru.y, sg), bb = __shfl(rv.x, sg), be = __shfl(rv.y, sg);
      const uint32_t f1 = FMT ? (uint32_t)FMT : qv.s1.fmt[sg];
      const uint32_t f0 = FMT ? (uint32_t)FMT : qv.s0.fmt[sg];
      uint64_t vres = bb;  // the first chunk of v's row the next tile needs
      for (uint64_t t0 = ab; t0 < ae && vres < be;) {
        const uint64_t tb = chunk_base(f1, t0);
        const uint64_t tend = min(ae, tb + (uint64_t)kQT);
        const uint32_t tl = (uint32_t)(tend - t0);
        uint64_t eu[kQPer], ev[kQPer];
        load_chunk<FMT>(qv.s1, sg, t0, tend, lane, eu);
        uint64_t c0 = vres;
        load_chunk<FMT>(qv.s0, sg, c0, be, lane, ev);
#pragma unroll
        for (int i = 0; i < kQPer; ++i) {
          const uint64_t idx = slot_entry(f1, tb, lane, i);
          if (idx >= t0 && idx < tend) {
            s_h[wave][idx - t0] = (uint32_t)(eu[i] >> 32);
            s_d[wave][idx - t0] = (uint32_t)eu[i];
          }
        }
        wave_sync();
        const uint32_t lo_h = s_h[wave][0], hi_h = s_h[wave][tl - 1];
        uint64_t mres = be;  // the first chunk holding an entry past this tile
        for (;;) {
          uint32_t base[kQPer];
          bool past = true;   // every slot of this lane lies past the tile
          bool more = false;  // some entry of this lane lies past the tile
#pragma unroll
          for (int i = 0; i < kQPer; ++i) {
            const uint32_t h = (uint32_t)(ev[i] >> 32);
            base[i] = 0;
            past = past && h > hi_h;
            more = more || (h != 0xFFFFFFFFu && h > hi_h);
          }
          if (mres == be && __any(more)) mres = c0;
          // kQPer branch-free lower-bound searches in lock step: every step is one
          // LDS read per search, all independent, so their latencies overlap
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
          const uint64_t next = chunk_base(f0, c0) + kQT;
          if (__all(past) || next >= be) break;
          c0 = next;
          load_chunk<FMT>(qv.s0, sg, c0, be, lane, ev);
        }
        vres = mres;
        wave_sync();
        t0 = tend;
      }
    }
    // wave minimum
    for (int o = 32; o > 0; o >>= 1) {
      const uint32_t other = __shfl_xor(best, o);
      best = other < best ? other : best;
    }
    if (lane == 0) out[k] = (int32_t)best;
  }
}

unsigned query_grid(uint64_t q) {
  const uint64_t b = (q * 64 + 255) / 256;
  return (unsigned)(b == 0 ? 1 : b < 8192 ? b : 8192);
}

}  // namespace

hipError_t launch_query(const QueryView& v, const uint32_t* pairs, uint32_t n, uint64_t q, int32_t* out,
                        hipStream_t s) {
  if (!q) return hipSuccess;
  const dim3 grid(query_grid(q)), block(256);
  if (v.fmts == SEG_F5) hipLaunchKernelGGL(query_kernel<SEG_F5>, grid, block, 0, s, v, pairs, n, q, out);
  else if (v.fmts == SEG_F8) hipLaunchKernelGGL(query_kernel<SEG_F8>, grid, block, 0, s, v, pairs, n, q, out);
  else hipLaunchKernelGGL(query_kernel<0>, grid, block, 0, s, v, pairs, n, q, out);
  return hipGetLastError();
}

}  // namespace chl
