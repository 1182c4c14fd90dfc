// Multi-root shortest-path-tree kernel: the label-producing hot loop of PLaNT
// (distributed_memory/akiba.hpp:52-104) and of GLL's pruned Dijkstra
// (shared_memory/akiba.hpp:65-125), run for many roots per persistent launch.
//
// Vertices carry two ids: the device id (a locality order of the graph, so that
// neighbours share cache lines) and the rank (position in the vertex order,
// rank 0 most important). Hubs, label rows and the flag rule use ranks.
//
// One workgroup owns one root at a time (it pulls the next root index from a
// device cursor) and a private slot of scratch sized n: a 16-byte cell per
// vertex holding the packed state word, a stamp and the vertex's rank, two
// frontier lists, the list of touched vertices and a deferred pile. The search
// is a bucketed label-correcting Dijkstra (delta stepping): each round
// processes the frontier entries whose distance is below the bucket bound B;
// the others wait on the pile (the lane-per-vertex path) or in the next
// frontier (the pruned path) until B advances. A vertex whose state improves is
// queued once per round (stamp) and processed again, so every vertex is finally
// processed at its final state.
//
// State word = (epoch key << 34) | (dist << 2) | tag (state_word in dev.hpp).
// With TREE_FLAGS the tag is clean while every vertex on the best path found so
// far, the vertex itself included, is ranked at or below the root, and covered
// otherwise; a 64-bit atomicMin keeps the shortest distance and, among equal
// distances, the covered tag: the path through a more important vertex. That is
// the tie rule of update_wo_prune (distributed_memory/dijkstra.hpp:56-78): a
// label (root, d) is emitted iff the final tag is clean. Without TREE_FLAGS
// (GLL) every reached vertex is clean and vertices ranked above the root are not
// expanded (akiba.hpp:92).
//
// The epoch key falls by one with every tree a slot starts, so the cells a tree
// leaves behind read as unreached to the next one: the search state is never
// written back after a tree, only read once to emit its labels. Round stamps
// likewise keep growing over a slot's trees, so a stale stamp never equals a
// current one. The slot's two counts (seq) persist from launch to launch; past
// their bounds the slot's cells are reset in place and the counts restart.
//
// With TREE_PRUNE a processed vertex u at distance d is dropped when the
// completed labels (all hubs < prefix) already give d(root, u) <= d: one
// wavefront scans u's label row coalesced and probes each hub in the root's
// label, held in an LDS hash table (akiba.hpp:94-110). A pruned vertex gets the
// pruned tag at that distance, so an equal-distance update cannot revive it.
//
// With TREE_FLAGS the search stops at a bucket boundary once no queued vertex
// is clean: every later vertex is then reached through a more important vertex
// (the queue_left test, distributed_memory/akiba.hpp:95).
#include "dev.hpp"

namespace chl {
namespace {

constexpr int kThreads = 256;
constexpr int kWaves = kThreads / 64;
constexpr uint32_t kLdsHashBits = 12;
constexpr uint32_t kLdsHash = 1u << kLdsHashBits;  // entries of the LDS root-label table
constexpr uint32_t kEmpty = 0xFFFFFFFFu;

// Frontiers of p.sort_min to kFrontSort entries are sorted by device id in LDS
// before they are processed (kFrontSort 0 or sort_min 0: never), so that the
// lanes of a wavefront take vertices that lie close together in the locality
// order and their cell loads share cache lines.
#ifndef CHL_FRONT_SORT
#define CHL_FRONT_SORT 4096
#endif
constexpr uint32_t kFrontSort = CHL_FRONT_SORT;
static_assert(kFrontSort == 0 || (kFrontSort >= 64 && (kFrontSort & (kFrontSort - 1)) == 0),
              "the frontier sort bound is 0 or a power of two of at least 64");

// Orders the LDS accesses of one wavefront's lanes (no workgroup barrier).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ uint32_t hash_slot(uint32_t h, uint32_t shift) { return (h * 0x9E3779B1u) >> shift; }

__device__ __forceinline__ uint32_t table_find(const uint32_t* key, const uint32_t* val, uint32_t shift,
                                               uint32_t mask, uint32_t h) {
  uint32_t s = hash_slot(h, shift);
  for (;;) {
    uint32_t k = key[s];
    if (k == h) return val[s];
    if (k == kEmpty) return kEmpty;
    s = (s + 1) & mask;
  }
}

__device__ __forceinline__ void table_insert(uint32_t* key, uint32_t* val, uint32_t shift, uint32_t mask,
                                             uint32_t h, uint32_t d) {
  uint32_t s = hash_slot(h, shift);
  for (;;) {
    if (atomicCAS(&key[s], kEmpty, h) == kEmpty) {
      val[s] = d;
      return;
    }
    s = (s + 1) & mask;
  }
}

// cell[2 v] = state word; cell[2 v + 1] = round stamp (low half) and the
// vertex's rank (high half). The rank is written once per build when the slots
// are initialised and kept by every reset, so one 16-byte load gives a head's
// state and rank during a relaxation, and processing and emission read no rank
// array; queueing exchanges the stamp half only.
__device__ __forceinline__ uint32_t* stamp_of(uint64_t* cell, uint32_t v) {
  return reinterpret_cast<uint32_t*>(cell + 2 * (size_t)v + 1);
}
__device__ __forceinline__ uint32_t rank_of(const uint64_t* cell, uint32_t v) {
  return reinterpret_cast<const uint32_t*>(cell + 2 * (size_t)v + 1)[1];
}

// Bucket bookkeeping of the next frontier, gathered from every state change
// of the round and from the re-queued entries: below = some queued distance is
// under the bucket bound B, mn = the smallest queued distance. A vertex's final
// state is the smallest of its recorded changes, so both are exact. (Whether a
// queued vertex still has low == root rank needs the final states; it is read
// at bucket boundaries only.)
struct Book {
  uint32_t below = 0, mn = 0xFFFFFFFFu;
  __device__ __forceinline__ void note(uint32_t d, uint32_t B) {
    below |= d < B;
    mn = min(mn, d);
  }
};

// The 16-byte cell of v (state word, stamp and rank).
__device__ __forceinline__ ulonglong2 load_cell(const uint64_t* cell, uint32_t v) {
  return *reinterpret_cast<const ulonglong2*>(cell + 2 * (size_t)v);
}

// Relaxes arc (u -> v, w) from u's distance d and tag, given a recent read
// `seen` of v's cell (v's rank is fixed; a stale state only costs an atomic
// that changes nothing); registers v in vis at its first state of this tree
// (epoch key ekey) and queues it in N once per round. rr: the root's rank.
template <bool FLAGS>
__device__ __forceinline__ void relax_seen(uint64_t* cell, uint32_t* N, uint32_t* vis, uint32_t* s_nsize,
                                           uint32_t* s_nvis, uint32_t nstamp, uint32_t ekey, uint32_t rr,
                                           uint64_t arc, ulonglong2 seen, uint32_t d, uint32_t tag, uint32_t B,
                                           Book& bk) {
  const uint32_t v = (uint32_t)(arc >> 32);
  const uint32_t rv = (uint32_t)(seen.y >> 32);
  const uint32_t nd = d + (uint32_t)arc;
  const uint32_t ntag = FLAGS && rv < rr ? kTagCovered : tag;
  const uint64_t nv = state_word(ekey, nd, ntag);
  uint64_t* sv = cell + 2 * (size_t)v;
  if (nv < seen.x) {
    const uint64_t old = atomicMin((unsigned long long*)sv, (unsigned long long)nv);
    if (nv < old) {
      bk.note(nd, B);
      if (state_ekey(old) != ekey) vis[atomicAdd(s_nvis, 1u)] = v;
      if (atomicExch(stamp_of(cell, v), nstamp) != nstamp) N[atomicAdd(s_nsize, 1u)] = v;
    }
  }
}

template <bool FLAGS>
__device__ __forceinline__ void relax(uint64_t* cell, uint32_t* N, uint32_t* vis, uint32_t* s_nsize,
                                      uint32_t* s_nvis, uint32_t nstamp, uint32_t ekey, uint32_t rr, uint64_t arc,
                                      uint32_t d, uint32_t tag, uint32_t B, Book& bk) {
  relax_seen<FLAGS>(cell, N, vis, s_nsize, s_nvis, nstamp, ekey, rr, arc, load_cell(cell, (uint32_t)(arc >> 32)), d,
                    tag, B, bk);
}

// Arcs of one vertex relaxed per step of the lane-per-vertex loop: their head
// states are read together, so the reads overlap instead of following each
// other's atomics.
constexpr uint32_t kArcStep = 4;

// THREADS = 256: a workgroup of four wavefronts per root; 64: one wavefront per
// root (more roots in flight, no cross-wave barrier cost for small trees); 1024:
// sixteen wavefronts per root (a quarter of the trees in flight, so a quarter of
// the per-tree state competes for each XCD's L2, for wide search fronts).
//
// KEEP (no FLAGS, no PRUNE): workgroup b searches root b only and leaves the
// search in its slot (cells at their final distances, vis[0, touched[b]) the
// reached vertices) for kernels that read the tree afterwards; nothing is
// emitted and nothing is reset.
template <bool FLAGS, bool PRUNE, int THREADS, bool KEEP = false>
__global__ __launch_bounds__(THREADS) void tree_kernel(TreeParams p) {
  constexpr int kThreads = THREADS;
  constexpr int kWaves = THREADS / 64;
  __shared__ uint32_t s_key[PRUNE ? kLdsHash : 1];
  __shared__ uint32_t s_val[PRUNE ? kLdsHash : 1];
  // (wavefront-per-root launches run many workgroups per CU: no sort there)
  constexpr uint32_t kSortCap = THREADS >= 256 ? kFrontSort : 0u;
  __shared__ uint32_t s_sort[kSortCap ? kSortCap : 1];
  __shared__ uint32_t s_ri, s_fsize, s_nsize, s_nvis, s_below, s_clean, s_minnext, s_emit, s_dsize, s_dmin;
  __shared__ unsigned long long s_base;
  __shared__ unsigned long long s_ctr[6];

  const uint32_t tid = threadIdx.x;
  const uint32_t lane = tid & 63;
  const uint32_t wave = tid >> 6;
  const size_t n = p.n;
  const size_t slot = blockIdx.x;
  uint64_t* cell = p.cell + slot * 2 * n;
  uint32_t* vis = p.vis + slot * n;
  uint64_t* D = p.qd + slot * n;  // deferred pile: (vertex << 32) | stamp at deferral

  unsigned long long c_proc = 0, c_arcs = 0, c_prune = 0, c_rounds = 0, c_settled = 0, c_sarcs = 0;
  uint32_t c_maxd = 0;
  // this root's share, added to the launch counters only when its labels fit
  // (a spilled root runs again in a later launch and is counted there)
  unsigned long long r_arcs = 0, r_prune = 0, r_settled = 0, r_sarcs = 0;
  if (tid < 6) s_ctr[tid] = 0;
  // the slot's tree count and next stamp base (the same in every thread)
  uint32_t epoch = p.seq[2 * slot], sbase = p.seq[2 * slot + 1];

  for (;;) {
    if (tid == 0 && KEEP) s_ri = blockIdx.x < p.R ? (uint32_t)blockIdx.x : kEmpty;
    if (tid == 0 && !KEEP) {
      unsigned long long r = atomicAdd(&p.ctr[CTR_NEXT], 1ull);
      // once a root's labels did not fit, later roots are not started: they go
      // back to the host with the spilled ones (no tree is computed in vain)
      while (r < p.R && __hip_atomic_load(&p.ctr[CTR_OVERFLOW], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
        const unsigned long long k = atomicAdd(&p.ctr[CTR_OVERFLOW], 1ull);
        p.spill[k] = (uint32_t)r;
        atomicAdd(&p.ctr[CTR_SKIPPED], 1ull);
        r = atomicAdd(&p.ctr[CTR_NEXT], 1ull);
      }
      s_ri = r < p.R ? (uint32_t)r : kEmpty;
    }
    __syncthreads();
    const uint32_t ri = s_ri;
    if (ri == kEmpty) break;
    if (epoch >= kEpochMax || sbase > kStampMax) {
      // both counts at their bounds: the slot's cells back to unreached, then
      // the counts restart
      for (size_t i = tid; i < n; i += kThreads) {
        cell[2 * i] = kStateInf;
        cell[2 * i + 1] = (uint64_t)p.rank[i] << 32;
      }
      epoch = 0;
      sbase = 0;
      __syncthreads();
    }
    ++epoch;
    const uint32_t ekey = kEpochMax - epoch;
    const uint32_t root = p.roots[ri];  // device id
    const uint32_t rr = p.rank[root];   // its rank = the hub of its labels
    const uint32_t skip = FLAGS ? (PRUNE ? p.prefix : 0u) : rr;  // ranks below are not expanded
    uint32_t* F = p.qa + slot * n;
    uint32_t* N = p.qb + slot * n;

    // Root label (all completed hubs) into the hash table.
    uint32_t* tkey = s_key;
    uint32_t* tval = s_val;
    uint32_t tshift = 32 - kLdsHashBits, tmask = kLdsHash - 1;
    if (PRUNE) {
      // the root's row over all segments of the root side
      uint64_t rlen = 0;
      for (uint32_t sg = 0; sg < p.root_nseg; ++sg) rlen += p.root_view[sg].off[rr + 1] - p.root_view[sg].off[rr];
      if (2 * rlen > kLdsHash) {
        tkey = p.gkey + slot * p.gcap;
        tval = p.gval + slot * p.gcap;
        tmask = p.gcap - 1;
        tshift = 32 - (31 - __clz(p.gcap));
      }
      for (uint32_t i = tid; i <= tmask; i += kThreads) tkey[i] = kEmpty;
      __syncthreads();
      for (uint32_t sg = 0; sg < p.root_nseg; ++sg) {
        const SegView rv = p.root_view[sg];
        const uint64_t rb = rv.off[rr], re = rv.off[rr + 1];
        for (uint64_t k = rb + tid; k < re; k += kThreads) {
          uint32_t h, d;
          seg_entry(rv, k, h, d);
          table_insert(tkey, tval, tshift, tmask, h, d);
        }
      }
    }
    if (tid == 0) {
      cell[2 * (size_t)root] = state_word(ekey, 0, kTagClean);
      cell[2 * (size_t)root + 1] = ((uint64_t)rr << 32) | (sbase + 1);  // rank, the stamp before round 1
      F[0] = root;
      vis[0] = root;
      s_fsize = 1;
      s_nvis = 1;
      s_below = 1;
      s_clean = 1;
      s_minnext = 0;
      s_dsize = 0;
      s_dmin = kEmpty;
    }
    uint32_t B = p.delta;
    uint32_t bfin = kEmpty;  // the bucket bound where the search stopped (all ones: it ran out)
    uint32_t round = 1;
    __syncthreads();

    for (;;) {
      const uint32_t fsize = s_fsize;
      const uint32_t dsize = min(s_dsize, (uint32_t)n);
      if (fsize == 0 && dsize == 0) break;
      // a bucket advance (nothing queued below B) also scans the deferred pile
      const bool adv = s_below == 0;
      if (adv) {
        // bucket [.., B) is final; every queued vertex is at >= B
        if (FLAGS && s_clean == 0) {
          bfin = B;
          break;
        }
        B = min(s_minnext, s_dmin) + p.delta;
      }
      const uint32_t dscan = adv ? dsize : 0u;
      __syncthreads();
      if (tid == 0) {
        s_nsize = 0;
        s_below = 0;
        s_clean = 0;
        s_minnext = kEmpty;
        if (adv) {
          s_dsize = 0;
          s_dmin = kEmpty;
        }
      }
      __syncthreads();
      // stamps of this round: queued for the next round, or put on the pile
      // (distinct, so that an improvement of a vertex put on the pile in the same
      // round still queues it for the next round)
      const uint32_t nstamp = sbase + 2 * round;
      const uint32_t pstamp = nstamp + 1;
      Book bk;
      uint32_t dmin = kEmpty;

      if (!PRUNE) {
        // one lane per entry of the frontier and, on a bucket advance, of the
        // deferred pile. Entries at or beyond B wait in the pile until the next
        // advance instead of being carried from round to round; a pile entry is
        // live only while the vertex's stamp is the one it was deferred with
        // (a later queueing supersedes it).
        for (uint32_t k = tid; k < fsize + dscan; k += kThreads) {
          uint32_t u;
          ulonglong2 c;
          if (k < fsize) {
            u = F[k];
            c = load_cell(cell, u);
          } else {
            const uint64_t e = D[k - fsize];
            u = (uint32_t)(e >> 32);
            c = load_cell(cell, u);
            if ((uint32_t)c.y != (uint32_t)e) continue;
          }
          const uint64_t x = c.x;
          const uint32_t ab = p.off[u], ae = p.off[u + 1];  // read beside the state
          const uint32_t d = state_dist(x);
          const uint32_t tag = state_tag(x);
          if (d >= B) {
            uint32_t* st = stamp_of(cell, u);
            bool to_next = adv;
            if (!adv) {
              // onto the pile unless queued already in this round
              uint32_t o = __hip_atomic_load(st, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
              while (o != nstamp && o != pstamp) {
                const uint32_t prev = atomicCAS(st, o, pstamp);
                if (prev == o) {
                  const uint32_t j = atomicAdd(&s_dsize, 1u);
                  if (j < n) {
                    D[j] = ((uint64_t)u << 32) | pstamp;
                    dmin = min(dmin, d);
                  } else {
                    to_next = true;  // the pile is full: the next frontier takes it
                  }
                  break;
                }
                o = prev;
              }
            }
            if (to_next) {
              bk.note(d, B);
              if (atomicExch(st, nstamp) != nstamp) N[atomicAdd(&s_nsize, 1u)] = u;
            }
            continue;
          }
          if (!FLAGS && rank_of(cell, u) < skip) continue;
          ++c_proc;
          r_arcs += ae - ab;
          for (uint32_t a0 = ab; a0 < ae; a0 += kArcStep) {
            uint64_t av[kArcStep];
            ulonglong2 seen[kArcStep];
#pragma unroll
            for (uint32_t j = 0; j < kArcStep; ++j)
              if (a0 + j < ae) av[j] = p.arc[a0 + j];
#pragma unroll
            for (uint32_t j = 0; j < kArcStep; ++j)
              if (a0 + j < ae) seen[j] = load_cell(cell, (uint32_t)(av[j] >> 32));
#pragma unroll
            for (uint32_t j = 0; j < kArcStep; ++j)
              if (a0 + j < ae)
                relax_seen<FLAGS>(cell, N, vis, &s_nsize, &s_nvis, nstamp, ekey, rr, av[j], seen[j], d, tag, B,
                                  bk);
          }
        }
      } else {
        // one wavefront per frontier entry: prune check, then relax
        for (uint32_t k = wave; k < fsize; k += kWaves) {
          const uint32_t u = F[k];
          const uint64_t x = cell[2 * (size_t)u];
          const uint32_t d = state_dist(x);
          const uint32_t tag = state_tag(x);
          if (d >= B) {
            if (lane == 0) {
              bk.note(d, B);
              if (atomicExch(stamp_of(cell, u), nstamp) != nstamp) N[atomicAdd(&s_nsize, 1u)] = u;
            }
            continue;
          }
          const uint32_t ru = rank_of(cell, u);
          if (ru < skip || tag == kTagPruned) continue;  // pruned at this distance already
          if (lane == 0) ++c_proc;
          bool pruned = false;
          uint64_t scanned = 0;
          for (uint32_t sg = 0; sg < p.scan_nseg && !pruned; ++sg) {
            const SegView sv = p.scan_view[sg];
            const uint64_t rb = sv.off[ru], re = sv.off[ru + 1];
            uint64_t base = rb;
            for (; base < re; base += 64) {
              const uint64_t kk = base + lane;
              bool hit = false;
              if (kk < re) {
                uint32_t h, ed;
                seg_entry(sv, kk, h, ed);
                const uint32_t rd = table_find(tkey, tval, tshift, tmask, h);
                hit = rd != kEmpty && rd + ed <= d;
              }
              if (__any(hit)) {
                pruned = true;
                break;
              }
            }
            scanned += (pruned ? min(base + 64, re) : re) - rb;
          }
          if (lane == 0) r_prune += scanned;
          if (pruned) {
            if (lane == 0)
              atomicMin((unsigned long long*)&cell[2 * (size_t)u], (unsigned long long)state_word(ekey, d, kTagPruned));
            continue;
          }
          const uint32_t ab = p.off[u], ae = p.off[u + 1];
          if (lane == 0) r_arcs += ae - ab;
          for (uint32_t a = ab + lane; a < ae; a += 64)
            relax<FLAGS>(cell, N, vis, &s_nsize, &s_nvis, nstamp, ekey, rr, p.arc[a], d, tag, B, bk);
        }
      }
      // bookkeeping of the next frontier (wavefront reductions, then LDS)
      const bool any_below = __any(bk.below != 0);
      uint32_t mn = bk.mn;
      for (int off = 32; off > 0; off >>= 1) mn = min(mn, (uint32_t)__shfl_xor((int)mn, off));
      for (int off = 32; off > 0; off >>= 1) dmin = min(dmin, (uint32_t)__shfl_xor((int)dmin, off));
      if (lane == 0) {
        if (any_below) s_below = 1;
        if (mn != kEmpty) atomicMin(&s_minnext, mn);
        if (dmin != kEmpty) atomicMin(&s_dmin, dmin);
      }
      __syncthreads();
      const uint32_t nsize = s_nsize;
      if (FLAGS && s_below == 0) {
        // bucket boundary: is some queued vertex (next frontier or live pile
        // entry) still reached through the root alone? (stops at the first one
        // found)
        const uint32_t pile = min(s_dsize, (uint32_t)n);
        for (uint32_t k = tid; k < nsize + pile; k += kThreads) {
          if (s_clean) break;
          if (k < nsize) {
            if (state_tag(cell[2 * (size_t)N[k]]) == kTagClean) s_clean = 1;
          } else {
            const uint64_t e = D[k - nsize];
            const ulonglong2 c = load_cell(cell, (uint32_t)(e >> 32));
            if ((uint32_t)c.y == (uint32_t)e && state_tag(c.x) == kTagClean) s_clean = 1;
          }
        }
        __syncthreads();
      }
      if (kSortCap && p.sort_min && nsize > 1 && nsize >= p.sort_min && nsize <= kSortCap) {
        // bitonic sort of the next frontier by device id, padded with all ones to
        // a power of two. Thread t owns the pair (2t, 2t + 1) at loading and at
        // write-back and, at strides up to 64, only pairs inside its wavefront's
        // 128-entry block: a stage needs a workgroup barrier only when it or the
        // stage before it crosses those blocks, else a wavefront sync. (The
        // round's closing barrier orders the write-back before the next round
        // reads it.)
        uint32_t P = 64;
        while (P < nsize) P <<= 1;
        for (uint32_t t = tid; t < P / 2; t += kThreads) {
          s_sort[2 * t] = 2 * t < nsize ? N[2 * t] : kEmpty;
          s_sort[2 * t + 1] = 2 * t + 1 < nsize ? N[2 * t + 1] : kEmpty;
        }
        bool crossed = false;  // the previous stage crossed wavefront blocks
        for (uint32_t size = 2; size <= P; size <<= 1)
          for (uint32_t stride = size >> 1; stride > 0; stride >>= 1) {
            if (crossed || stride > 64)
              __syncthreads();
            else
              wave_sync();
            crossed = stride > 64;
            for (uint32_t t = tid; t < P / 2; t += kThreads) {
              const uint32_t i = 2 * t - (t & (stride - 1)), j = i + stride;
              const uint32_t a = s_sort[i], b = s_sort[j];
              if ((a > b) == ((i & size) == 0)) {
                s_sort[i] = b;
                s_sort[j] = a;
              }
            }
          }
        for (uint32_t t = tid; t < P / 2; t += kThreads) {
          if (2 * t < nsize) N[2 * t] = s_sort[2 * t];
          if (2 * t + 1 < nsize) N[2 * t + 1] = s_sort[2 * t + 1];
        }
      }
      if (tid == 0) s_fsize = nsize;
      uint32_t* t = F;
      F = N;
      N = t;
      ++round;
      ++c_rounds;
      __syncthreads();
    }
    sbase += 2 * round + 2;  // above every stamp of this tree

    if (KEEP) {
      if (tid == 0) p.touched[ri] = s_nvis;
      break;
    }

    // Emit (root rank, d) at every touched vertex whose final tag is clean. One
    // pass reads each touched cell once (and writes none: the next tree's epoch
    // key retires them) and keeps the label's (rank, d) in the two frontier lists
    // (free once the search has ended); then the root's output range is reserved
    // and the kept labels are written out, coalesced.
    const uint32_t nvis = s_nvis;
    if (tid == 0) s_emit = 0;
    __syncthreads();
    for (uint32_t k = tid; k < nvis; k += kThreads) {
      const uint32_t v = vis[k];
      const ulonglong2 c = load_cell(cell, v);
      const uint32_t rv = (uint32_t)(c.y >> 32);
      const uint32_t dv = state_dist(c.x), tv = state_tag(c.x);
      if (tv == kTagClean && rv >= rr) {
        const uint32_t j = atomicAdd(&s_emit, 1u);
        F[j] = rv;
        N[j] = dv;
        c_maxd = max(c_maxd, dv);
      }
      if (dv < bfin) {
        // settled: processed at its final distance; expanded unless pruned
        // (low 0) or ranked above a GLL root
        ++r_settled;
        if ((!PRUNE || tv != kTagPruned) && (FLAGS || rv >= skip)) r_sarcs += p.off[v + 1] - p.off[v];
      }
    }
    __syncthreads();
    const uint32_t emit = s_emit;
    if (tid == 0) {
      // reserve the root's output range (wait-free); a range that does not fit
      // is not written: the root is recorded for a second launch, and the
      // smallest such start marks the end of the contiguous written prefix
      unsigned long long b = atomicAdd(&p.ctr[CTR_LABELS], (unsigned long long)emit);
      if (b + emit > p.out_cap) {
        const unsigned long long k = atomicAdd(&p.ctr[CTR_OVERFLOW], 1ull);
        p.spill[k] = ri;
        atomicAdd(&p.ctr[CTR_SPILL], (unsigned long long)emit);
        atomicMin(&p.ctr[CTR_SPILL_START], b);
        b = ~0ull;
      }
      s_base = b;
      p.label_count[ri] = emit;
      if (p.touched) p.touched[ri] = nvis;
    }
    __syncthreads();
    const unsigned long long obase = s_base;
    if (obase != ~0ull)
      for (uint32_t k = tid; k < emit; k += kThreads) {
        p.out_row[obase + k] = F[k];
        p.out_hd[obase + k] = ((uint64_t)rr << 32) | N[k];
      }
    // tree size = processed entries of this root
    if (c_proc) atomicAdd(&s_ctr[0], c_proc);
    if (obase != ~0ull) {
      c_arcs += r_arcs;
      c_prune += r_prune;
      c_settled += r_settled;
      c_sarcs += r_sarcs;
    }
    r_arcs = r_prune = r_settled = r_sarcs = 0;
    __syncthreads();
    if (tid == 0) {
      p.tree_size[ri] = (uint32_t)s_ctr[0];
      if (obase != ~0ull) s_ctr[1] += s_ctr[0];
      s_ctr[0] = 0;
    }
    c_proc = 0;
  }
  if (c_arcs) atomicAdd(&s_ctr[2], c_arcs);
  if (c_prune) atomicAdd(&s_ctr[3], c_prune);
  if (c_settled) atomicAdd(&s_ctr[4], c_settled);
  if (c_sarcs) atomicAdd(&s_ctr[5], c_sarcs);
  for (int o = 32; o > 0; o >>= 1) c_maxd = max(c_maxd, (uint32_t)__shfl_xor((int)c_maxd, o));
  if (lane == 0 && c_maxd) atomicMax(&p.ctr[CTR_MAXD], (unsigned long long)c_maxd);
  __syncthreads();
  if (tid == 0) {
    p.seq[2 * slot] = epoch;
    p.seq[2 * slot + 1] = sbase;
    if (s_ctr[1]) atomicAdd(&p.ctr[CTR_PROCESSED], s_ctr[1]);
    if (s_ctr[2]) atomicAdd(&p.ctr[CTR_ARCS], s_ctr[2]);
    if (s_ctr[3]) atomicAdd(&p.ctr[CTR_PRUNE_ENT], s_ctr[3]);
    if (s_ctr[4]) atomicAdd(&p.ctr[CTR_SETTLED], s_ctr[4]);
    if (s_ctr[5]) atomicAdd(&p.ctr[CTR_SETTLED_ARCS], s_ctr[5]);
    if (c_rounds) atomicAdd(&p.ctr[CTR_ROUNDS], c_rounds);
  }
}

// Cells start unreached: state all ones, stamp 0, and the vertex's rank (cell i
// of a slot belongs to device id i mod n).
__global__ void init_cells_kernel(uint64_t* cell, size_t count, const uint32_t* rank, uint32_t n) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x) {
    cell[2 * i] = kStateInf;
    cell[2 * i + 1] = (uint64_t)rank[i % n] << 32;
  }
}
__global__ void fill_u64_kernel(uint64_t* p, uint64_t v, size_t count) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x)
    p[i] = v;
}
__global__ void fill_u32_kernel(uint32_t* p, uint32_t v, size_t count) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x)
    p[i] = v;
}

__global__ void scatter_u32_kernel(const uint32_t* src, const uint32_t* idx, uint32_t m, uint32_t* dst) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < m; i += (size_t)gridDim.x * blockDim.x)
    dst[idx[i]] = src[i];
}

static unsigned fill_blocks(size_t count) {
  size_t b = (count + 255) / 256;
  return (unsigned)(b < 8192 ? (b ? b : 1) : 8192);
}

}  // namespace

template <int T>
static void launch_tree_t(const TreeParams& p, uint32_t flags, uint32_t slots, hipStream_t s) {
  dim3 grid(slots), block(T);
  switch (flags & (TREE_FLAGS | TREE_PRUNE | TREE_KEEP)) {
    case TREE_FLAGS: hipLaunchKernelGGL((tree_kernel<true, false, T>), grid, block, 0, s, p); break;
    case TREE_FLAGS | TREE_PRUNE: hipLaunchKernelGGL((tree_kernel<true, true, T>), grid, block, 0, s, p); break;
    case TREE_PRUNE: hipLaunchKernelGGL((tree_kernel<false, true, T>), grid, block, 0, s, p); break;
    case TREE_KEEP: hipLaunchKernelGGL((tree_kernel<false, false, T, true>), grid, block, 0, s, p); break;
    default: hipLaunchKernelGGL((tree_kernel<false, false, T>), grid, block, 0, s, p); break;
  }
}

hipError_t launch_tree(const TreeParams& p, uint32_t flags, uint32_t slots, int threads_per_root, hipStream_t s) {
  if (threads_per_root == 64) launch_tree_t<64>(p, flags, slots, s);
  else if (threads_per_root == 1024) launch_tree_t<1024>(p, flags, slots, s);
  else launch_tree_t<kThreads>(p, flags, slots, s);
  return hipGetLastError();
}

hipError_t launch_init_cells(uint64_t* cell, size_t count, const uint32_t* rank, uint32_t n, hipStream_t s) {
  if (!count) return hipSuccess;
  hipLaunchKernelGGL(init_cells_kernel, dim3(fill_blocks(count)), dim3(256), 0, s, cell, count, rank, n);
  return hipGetLastError();
}

hipError_t launch_fill_u64(uint64_t* p, uint64_t v, size_t count, hipStream_t s) {
  if (!count) return hipSuccess;
  hipLaunchKernelGGL(fill_u64_kernel, dim3(fill_blocks(count)), dim3(256), 0, s, p, v, count);
  return hipGetLastError();
}
hipError_t launch_fill_u32(uint32_t* p, uint32_t v, size_t count, hipStream_t s) {
  if (!count) return hipSuccess;
  hipLaunchKernelGGL(fill_u32_kernel, dim3(fill_blocks(count)), dim3(256), 0, s, p, v, count);
  return hipGetLastError();
}
hipError_t launch_scatter_u32(const uint32_t* src, const uint32_t* idx, uint32_t m, uint32_t* dst, hipStream_t s) {
  if (!m) return hipSuccess;
  hipLaunchKernelGGL(scatter_u32_kernel, dim3(fill_blocks(m)), dim3(256), 0, s, src, idx, m, dst);
  return hipGetLastError();
}

}  // namespace chl
