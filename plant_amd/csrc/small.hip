// PLaNT trees of small and medium roots (distributed_memory/akiba.hpp:52-104)
// with the search state in a hash table per slot instead of the tree kernel's
// n-sized cell array: small trees one wavefront per root (tables of tens of
// thousands of entries), medium trees one workgroup of 256 threads per root
// (tables of hundreds of thousands).
//
// Most roots of a large graph grow small trees (on the 16M-vertex road-like
// graph the median tree below rank 1.7M processes 42 vertices, and the trees of
// ranks 300K-2M about 10^5), but a dense slot costs 36 bytes per vertex of the
// graph, so a memory-limited build runs only a few hundred of them at once. A
// hashed slot costs 28 bytes per table entry: thousands run at once. The tables
// are emptied entry by entry as each tree ends, so a slot's next tree starts on
// an empty table at no cost beyond the tree's own size.
//
// The search is the tree kernel's PLaNT search (tree.hip): bucketed
// label-correcting Dijkstra with the packed (dist, tag) state and a 64-bit
// atomicMin, a vertex queued once per round by a stamp exchange, and the stop at
// a bucket boundary once no queued vertex is clean. Entries at or past the
// bucket bound stay in the next frontier (no deferred pile). The frontier lists
// hold table indices. A root marked big by the caller is not started, and a
// tree whose inserted entries reach half the table is abandoned (its entries
// emptied): both go to the spill list, unstarted for the launch's counters, and
// the caller runs them with the tree kernel. Labels do not depend on which
// kernel built a tree.
#include "dev.hpp"

namespace chl {
namespace {

constexpr uint32_t kNone = 0xFFFFFFFFu;
// the epoch key of every small-tree state: tables are emptied after each tree
constexpr uint32_t kSmallKey = kEpochMax - 1;

__device__ __forceinline__ uint32_t home_slot(uint32_t v, uint32_t shift) { return (v * 0x9E3779B1u) >> shift; }

// Table index of vertex v (stored as v + 1; 0 is empty), inserted when absent
// (*ins then set); kNone when a whole sweep of the table found no room. The
// probe starts at slot s, whose key was read as `first` (a stale read is safe:
// keys are never removed during a tree, and an empty one is claimed by CAS).
__device__ __forceinline__ uint32_t find_or_insert_at(uint32_t* key, uint32_t mask, uint32_t v, uint32_t s,
                                                      uint32_t first, bool& ins) {
  const uint32_t k = v + 1;
  ins = false;
  for (uint32_t probe = 0; probe <= mask; ++probe) {
    uint32_t cur = probe ? __hip_atomic_load(&key[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : first;
    if (cur == 0) {
      cur = atomicCAS(&key[s], 0u, k);
      if (cur == 0) {
        ins = true;
        return s;
      }
    }
    if (cur == k) return s;
    s = (s + 1) & mask;
  }
  return kNone;
}

__device__ __forceinline__ uint32_t find_or_insert(uint32_t* key, uint32_t shift, uint32_t mask, uint32_t v,
                                                   bool& ins) {
  const uint32_t s = home_slot(v, shift);
  return find_or_insert_at(key, mask, v, s, __hip_atomic_load(&key[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                           ins);
}

// Arcs of one vertex relaxed per step: their arc words, head ranks and first
// probes are read together, so the reads overlap instead of following each
// other's atomics.
constexpr uint32_t kHArcStep = 4;

__device__ __forceinline__ uint32_t wave_min(uint32_t x) {
  for (int o = 32; o > 0; o >>= 1) x = min(x, (uint32_t)__shfl_xor((int)x, o));
  return x;
}

// THREADS = 64: one wavefront per root (small tables); 256: a workgroup of four
// wavefronts per root (medium tables).
template <int THREADS>
__global__ __launch_bounds__(THREADS) void small_tree_kernel(TreeParams p) {
  __shared__ uint32_t s_ri, s_fsize, s_nsize, s_nvis, s_below, s_clean, s_minnext, s_emit, s_over;
  __shared__ unsigned long long s_base, s_tsize;
  const uint32_t tid = threadIdx.x;
  const uint32_t lane = tid & 63;
  const size_t slot = blockIdx.x;
  const uint32_t C = p.scap, mask = C - 1, limit = C / 2;
  const uint32_t shift = (uint32_t)__clz((int)C) + 1;  // (v * phi) >> shift: log2(C) bits
  uint64_t* state = p.sstate + slot * C;
  uint32_t* key = p.skey + slot * C;
  uint32_t* stamp = p.sstamp + slot * C;
  uint32_t* L0 = p.slist + slot * 3 * (size_t)C;
  uint32_t* L1 = L0 + C;
  uint32_t* vis = L1 + C;
  unsigned long long c_proc_fit = 0, c_arcs = 0, c_settled = 0, c_sarcs = 0, c_rounds = 0;
  uint32_t c_maxd = 0;

  for (;;) {
    if (tid == 0) {
      unsigned long long r = atomicAdd(&p.ctr[CTR_NEXT], 1ull);
      // roots predicted big, and every root once some root's labels did not fit
      // the output (a spill start was recorded), go to the spill list unstarted
      while (r < p.R && ((p.big && p.big[r]) || __hip_atomic_load(&p.ctr[CTR_SPILL_START], __ATOMIC_RELAXED,
                                                                  __HIP_MEMORY_SCOPE_AGENT) != ~0ull)) {
        const unsigned long long k = atomicAdd(&p.ctr[CTR_OVERFLOW], 1ull);
        p.spill[k] = (uint32_t)r;
        atomicAdd(&p.ctr[CTR_SKIPPED], 1ull);
        r = atomicAdd(&p.ctr[CTR_NEXT], 1ull);
      }
      s_ri = r < p.R ? (uint32_t)r : kNone;
    }
    __syncthreads();
    const uint32_t ri = s_ri;
    if (ri == kNone) break;
    const uint32_t root = p.roots[ri];  // device id
    const uint32_t rr = p.rank[root];   // the hub of its labels
    if (tid == 0) {
      bool ins;
      const uint32_t s0 = find_or_insert(key, shift, mask, root, ins);  // the table is empty
      state[s0] = state_word(kSmallKey, 0, kTagClean);
      stamp[s0] = 1;
      L0[0] = s0;
      vis[0] = s0;
      s_fsize = 1;
      s_nvis = 1;
      s_below = 1;
      s_clean = 1;
      s_minnext = 0;
      s_over = 0;
      s_tsize = 0;
    }
    uint32_t* F = L0;
    uint32_t* N = L1;
    uint32_t B = p.delta;
    uint32_t bfin = kNone;  // the bucket bound where the search stopped (all ones: it ran out)
    uint32_t round = 1;
    unsigned long long c_proc = 0, r_arcs = 0;
    __syncthreads();

    for (;;) {
      const uint32_t fsize = s_fsize;
      if (fsize == 0) break;
      if (s_below == 0) {
        // bucket [.., B) is final; every queued vertex is at >= B
        if (s_clean == 0) {
          bfin = B;
          break;
        }
        B = s_minnext + p.delta;
      }
      __syncthreads();
      if (tid == 0) {
        s_nsize = 0;
        s_below = 0;
        s_clean = 0;
        s_minnext = kNone;
      }
      __syncthreads();
      const uint32_t nstamp = 2 * round;  // the root's first stamp is 1; emptied entries hold 0
      uint32_t below = 0, mn = kNone;
      for (uint32_t k = tid; k < fsize; k += THREADS) {
        const uint32_t si = F[k];
        const uint64_t x = state[si];
        const uint32_t d = state_dist(x);
        const uint32_t tag = state_tag(x);
        if (d >= B) {
          // waits in the next frontier until the bucket reaches it
          mn = min(mn, d);
          if (atomicExch(&stamp[si], nstamp) != nstamp) N[atomicAdd(&s_nsize, 1u)] = si;
          continue;
        }
        const uint32_t u = key[si] - 1;
        ++c_proc;
        const uint32_t ab = p.off[u], ae = p.off[u + 1];
        r_arcs += ae - ab;
        for (uint32_t a0 = ab; a0 < ae; a0 += kHArcStep) {
          if (__hip_atomic_load(&s_over, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) break;
          uint64_t av[kHArcStep];
          uint32_t rv[kHArcStep], hs[kHArcStep], hk[kHArcStep];
#pragma unroll
          for (uint32_t j = 0; j < kHArcStep; ++j)
            if (a0 + j < ae) av[j] = p.arc[a0 + j];
#pragma unroll
          for (uint32_t j = 0; j < kHArcStep; ++j)
            if (a0 + j < ae) {
              const uint32_t v = (uint32_t)(av[j] >> 32);
              rv[j] = p.rank[v];
              hs[j] = home_slot(v, shift);
              hk[j] = __hip_atomic_load(&key[hs[j]], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
          bool full = false;
#pragma unroll
          for (uint32_t j = 0; j < kHArcStep; ++j)
            if (a0 + j < ae && !full) {
              const uint32_t v = (uint32_t)(av[j] >> 32);
              const uint32_t nd = d + (uint32_t)av[j];
              const uint32_t ntag = rv[j] < rr ? kTagCovered : tag;
              const uint64_t nv = state_word(kSmallKey, nd, ntag);
              bool ins;
              const uint32_t sv = find_or_insert_at(key, mask, v, hs[j], hk[j], ins);
              if (sv == kNone) {
                s_over = 1;
                full = true;
                continue;
              }
              if (ins) {
                const uint32_t q = atomicAdd(&s_nvis, 1u);
                vis[q] = sv;
                if (q + 1 >= limit) s_over = 1;
              }
              const uint64_t old = atomicMin((unsigned long long*)&state[sv], (unsigned long long)nv);
              if (nv < old) {
                below |= nd < B;
                mn = min(mn, nd);
                if (atomicExch(&stamp[sv], nstamp) != nstamp) N[atomicAdd(&s_nsize, 1u)] = sv;
              }
            }
          if (full) break;
        }
      }
      // bookkeeping of the next frontier (wavefront reductions, then LDS)
      const bool any_below = __any(below != 0);
      mn = wave_min(mn);
      if (lane == 0) {
        if (any_below) s_below = 1;
        if (mn != kNone) atomicMin(&s_minnext, mn);
      }
      __syncthreads();
      if (s_over) break;
      const uint32_t nsize = s_nsize;
      if (s_below == 0) {
        // bucket boundary: is some queued vertex still reached through the root alone?
        bool cl = false;
        for (uint32_t k = tid; k < nsize && !cl; k += THREADS) cl = state_tag(state[N[k]]) == kTagClean;
        const bool any_clean = __any(cl);
        if (lane == 0 && any_clean) s_clean = 1;
      }
      if (tid == 0) s_fsize = nsize;
      uint32_t* t = F;
      F = N;
      N = t;
      ++round;
      ++c_rounds;
      __syncthreads();
    }

    const uint32_t nvis = min(s_nvis, C);
    if (s_over) {
      // abandoned: the entries go back to empty and the root to the tree kernel
      for (uint32_t k = tid; k < nvis; k += THREADS) {
        const uint32_t si = vis[k];
        state[si] = kStateInf;
        stamp[si] = 0;
        key[si] = 0;
      }
      if (tid == 0) {
        const unsigned long long k = atomicAdd(&p.ctr[CTR_OVERFLOW], 1ull);
        p.spill[k] = ri;
        atomicAdd(&p.ctr[CTR_SKIPPED], 1ull);
      }
      __syncthreads();
      continue;
    }

    // Emit (root rank, d) at every inserted vertex whose final tag is clean,
    // keeping the labels in the idle frontier lists, and empty each entry.
    {
      // tree size = processed entries of this root
      unsigned long long w = c_proc;
      for (int o = 32; o > 0; o >>= 1) w += __shfl_xor(w, o);
      if (lane == 0 && w) atomicAdd(&s_tsize, w);
    }
    if (tid == 0) s_emit = 0;
    __syncthreads();
    unsigned long long r_settled = 0, r_sarcs = 0;
    for (uint32_t k = tid; k < nvis; k += THREADS) {
      const uint32_t si = vis[k];
      const uint32_t v = key[si] - 1;
      const uint64_t x = state[si];
      const uint32_t rv = p.rank[v];
      const uint32_t dv = state_dist(x), tv = state_tag(x);
      if (tv == kTagClean && rv >= rr) {
        const uint32_t j = atomicAdd(&s_emit, 1u);
        F[j] = rv;
        N[j] = dv;
        c_maxd = max(c_maxd, dv);
      }
      if (dv < bfin) {
        ++r_settled;
        r_sarcs += p.off[v + 1] - p.off[v];
      }
      state[si] = kStateInf;
      stamp[si] = 0;
      key[si] = 0;
    }
    __syncthreads();
    const uint32_t emit = s_emit;
    if (tid == 0) {
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
      p.tree_size[ri] = (uint32_t)s_tsize;
      if (p.touched) p.touched[ri] = nvis;
    }
    __syncthreads();
    const unsigned long long obase = s_base;
    if (obase != ~0ull)
      for (uint32_t k = tid; k < emit; k += THREADS) {
        p.out_row[obase + k] = F[k];
        p.out_hd[obase + k] = ((uint64_t)rr << 32) | N[k];
      }
    if (obase != ~0ull) {
      c_proc_fit += c_proc;
      c_arcs += r_arcs;
      c_settled += r_settled;
      c_sarcs += r_sarcs;
    }
    __syncthreads();
  }
  for (int o = 32; o > 0; o >>= 1) {
    c_proc_fit += __shfl_xor(c_proc_fit, o);
    c_arcs += __shfl_xor(c_arcs, o);
    c_settled += __shfl_xor(c_settled, o);
    c_sarcs += __shfl_xor(c_sarcs, o);
    c_maxd = max(c_maxd, (uint32_t)__shfl_xor((int)c_maxd, o));
  }
  if (lane == 0) {
    if (c_proc_fit) atomicAdd(&p.ctr[CTR_PROCESSED], c_proc_fit);
    if (c_arcs) atomicAdd(&p.ctr[CTR_ARCS], c_arcs);
    if (c_settled) atomicAdd(&p.ctr[CTR_SETTLED], c_settled);
    if (c_sarcs) atomicAdd(&p.ctr[CTR_SETTLED_ARCS], c_sarcs);
    if (c_rounds) atomicAdd(&p.ctr[CTR_ROUNDS], c_rounds);
    if (c_maxd) atomicMax(&p.ctr[CTR_MAXD], (unsigned long long)c_maxd);
  }
}

__global__ void init_small_kernel(uint64_t* state, uint32_t* key, uint32_t* stamp, size_t count) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x) {
    state[i] = kStateInf;
    key[i] = 0;
    stamp[i] = 0;
  }
}

}  // namespace

hipError_t launch_small_trees(const TreeParams& p, uint32_t slots, int threads, hipStream_t s) {
  if (!p.R || !slots) return hipSuccess;
  if (threads == 256)
    hipLaunchKernelGGL(small_tree_kernel<256>, dim3(slots), dim3(256), 0, s, p);
  else
    hipLaunchKernelGGL(small_tree_kernel<64>, dim3(slots), dim3(64), 0, s, p);
  return hipGetLastError();
}

hipError_t launch_init_small(const TreeParams& p, uint32_t slots, hipStream_t s) {
  const size_t count = (size_t)slots * p.scap;
  if (!count) return hipSuccess;
  size_t b = (count + 255) / 256;
  hipLaunchKernelGGL(init_small_kernel, dim3((unsigned)(b < 8192 ? b : 8192)), dim3(256), 0, s, p.sstate, p.skey,
                     p.sstamp, count);
  return hipGetLastError();
}

}  // namespace chl
