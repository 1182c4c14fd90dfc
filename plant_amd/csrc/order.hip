// Sampled-betweenness vertex order on the device. The reference takes the order
// as an input file (shared_memory/hhl.cpp:59-62, made by an external tool,
// README.md:24); this computes the same order as the host restatement
// (gen.cpp order_betweenness): from each of R seeded roots a shortest-path tree
// whose parent of v is the smallest-id u with dist(u) + len(u, v) == dist(v);
// a vertex scores its proper descendants summed over the trees; the order is
// score descending, then vertex id ascending.
//
// One workgroup owns one root at a time (a device cursor hands out roots) and a
// slot of scratch sized n. Per root:
//  1. distances by bucketed label-correcting Dijkstra (delta stepping), a
//     32-bit atomicMin per relaxation and a round stamp that queues a vertex
//     once per round;
//  2. every reached vertex v != root scans its in-arcs for the parent and adds
//     one child to the parent's accumulator word (children << 32 | subtree);
//  3. leaves walk up: each child adds (subtree - 2^32) to its parent's word in
//     one 64-bit atomic, so the child that brings the children count to zero
//     learns the parent's final subtree size from the returned word and
//     continues upward. Every non-leaf is finished by exactly one child, which
//     adds subtree - 1 to the vertex's score. No fence orders the counts against
//     the sizes: they are the same word.
//  4. the touched cells are restored to their unreached state.
#include "dev.hpp"

namespace chl {
namespace {

constexpr int kBcThreads = 256;
constexpr uint32_t kNone = 0xFFFFFFFFu;

// Distances read after other lanes' atomics: loads at device scope.
__device__ __forceinline__ uint32_t load_dist(const uint32_t* cell, uint32_t v) {
  return __hip_atomic_load(cell + 2 * (size_t)v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ __launch_bounds__(kBcThreads) void betweenness_kernel(BcParams p) {
  __shared__ uint32_t s_ri, s_fsize, s_nsize, s_nvis, s_below, s_minnext;
  __shared__ unsigned long long s_arcs;

  const uint32_t tid = threadIdx.x;
  const uint32_t lane = tid & 63;
  const size_t n = p.n;
  const size_t slot = blockIdx.x;
  uint32_t* cell = p.cell + slot * 2 * n;  // [2 v] distance, [2 v + 1] round stamp, then the parent
  unsigned long long* acc = p.acc + slot * n;
  uint32_t* vis = p.vis + slot * n;

  for (;;) {
    if (tid == 0) {
      const unsigned long long r = atomicAdd(&p.ctr[BC_NEXT], 1ull);
      s_ri = r < p.R ? (uint32_t)r : kNone;
      s_arcs = 0;
    }
    __syncthreads();
    const uint32_t ri = s_ri;
    if (ri == kNone) break;
    const uint32_t root = p.roots[ri];
    uint32_t* F = p.qa + slot * n;
    uint32_t* N = p.qb + slot * n;
    if (tid == 0) {
      cell[2 * (size_t)root] = 0;
      cell[2 * (size_t)root + 1] = 1;
      F[0] = root;
      vis[0] = root;
      s_fsize = 1;
      s_nvis = 1;
      s_below = 1;
      s_minnext = 0;
    }
    uint32_t B = p.delta;
    uint32_t round = 1;
    unsigned long long arcs = 0, rounds = 0;
    __syncthreads();

    // 1. distances
    for (;;) {
      const uint32_t fsize = s_fsize;
      if (fsize == 0) break;
      if (s_below == 0) B = s_minnext + p.delta;  // every queued distance is >= B
      __syncthreads();
      if (tid == 0) {
        s_nsize = 0;
        s_below = 0;
        s_minnext = kNone;
      }
      __syncthreads();
      const uint32_t nstamp = round + 1;
      uint32_t below = 0, mn = kNone;
      for (uint32_t k = tid; k < fsize; k += kBcThreads) {
        const uint32_t u = F[k];
        const uint32_t d = load_dist(cell, u);
        if (d >= B) {
          mn = min(mn, d);
          if (atomicExch(&cell[2 * (size_t)u + 1], nstamp) != nstamp) N[atomicAdd(&s_nsize, 1u)] = u;
          continue;
        }
        const uint32_t ab = p.off[u], ae = p.off[u + 1];
        arcs += ae - ab;
        for (uint32_t a = ab; a < ae; ++a) {
          const uint64_t arc = p.arc[a];
          const uint32_t v = (uint32_t)(arc >> 32);
          const uint32_t nd = d + (uint32_t)arc;
          uint32_t* dv = &cell[2 * (size_t)v];
          if (nd < *dv) {
            const uint32_t old = atomicMin(dv, nd);
            if (nd < old) {
              below |= nd < B;
              mn = min(mn, nd);
              if (old == kNone) vis[atomicAdd(&s_nvis, 1u)] = v;
              if (atomicExch(&cell[2 * (size_t)v + 1], nstamp) != nstamp) N[atomicAdd(&s_nsize, 1u)] = v;
            }
          }
        }
      }
      // the next frontier's bucket bookkeeping: every state change of the round
      // and every re-queued entry was noted, so both figures are exact
      const bool any_below = __any(below != 0);
      for (int off = 32; off > 0; off >>= 1) mn = min(mn, (uint32_t)__shfl_xor((int)mn, off));
      if (lane == 0) {
        if (any_below) s_below = 1;
        if (mn != kNone) atomicMin(&s_minnext, mn);
      }
      __syncthreads();
      if (tid == 0) s_fsize = s_nsize;
      uint32_t* t = F;
      F = N;
      N = t;
      ++round;
      ++rounds;
      __syncthreads();
    }

    // 2. parents (smallest vertex id among the shortest-path predecessors) and
    // children counts
    const uint32_t nvis = s_nvis;
    for (uint32_t k = tid; k < nvis; k += kBcThreads) {
      const uint32_t v = vis[k];
      if (v == root) continue;
      const uint32_t dv = load_dist(cell, v);
      uint32_t best = kNone, bp = kNone;
      const uint32_t ab = p.ioff[v], ae = p.ioff[v + 1];
      for (uint32_t a = ab; a < ae; ++a) {
        const uint64_t arc = p.iarc[a];
        const uint32_t u = (uint32_t)(arc >> 32);
        const uint32_t du = load_dist(cell, u);
        if (du != kNone && du + (uint32_t)arc == dv) {
          const uint32_t iu = p.inv[u];
          if (iu < best) {
            best = iu;
            bp = u;
          }
        }
      }
      cell[2 * (size_t)v + 1] = bp;
      if (bp != kNone) atomicAdd(&acc[bp], 1ull << 32);
    }
    __syncthreads();
    // leaves (s_fsize is 0 since the search ended)
    for (uint32_t k = tid; k < nvis; k += kBcThreads) {
      const uint32_t v = vis[k];
      if ((__hip_atomic_load(&acc[v], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >> 32) == 0)
        F[atomicAdd(&s_fsize, 1u)] = v;
    }
    __syncthreads();
    // 3. subtree sizes, leaves upward
    const uint32_t nleaf = s_fsize;
    for (uint32_t k = tid; k < nleaf; k += kBcThreads) {
      uint32_t v = F[k];
      uint32_t s = 1;
      while (v != root) {
        const uint32_t pv = __hip_atomic_load(&cell[2 * (size_t)v + 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (pv == kNone) break;
        const unsigned long long old = atomicAdd(&acc[pv], (unsigned long long)s - (1ull << 32));
        if ((old >> 32) != 1) break;  // a sibling is still to come
        s += (uint32_t)old;
        atomicAdd(&p.score[pv], (unsigned long long)(s - 1));
        v = pv;
      }
    }
    __syncthreads();
    // 4. restore the slot
    for (uint32_t k = tid; k < nvis; k += kBcThreads) {
      const uint32_t v = vis[k];
      cell[2 * (size_t)v] = kNone;
      cell[2 * (size_t)v + 1] = 0;
      acc[v] = 1;
    }
    for (int off = 32; off > 0; off >>= 1) arcs += (unsigned long long)__shfl_xor((long long)arcs, off);
    if (lane == 0) atomicAdd(&s_arcs, arcs);
    __syncthreads();
    if (tid == 0) {
      atomicAdd(&p.ctr[BC_REACHED], (unsigned long long)nvis);
      atomicAdd(&p.ctr[BC_ARCS], s_arcs);
      atomicAdd(&p.ctr[BC_ROUNDS], rounds);
    }
  }
}

// key[v] = ~score, val[v] = v for vertex ids v (an ascending radix sort of the
// keys, stable, then gives score descending with ids ascending among ties).
__global__ void score_keys_kernel(const unsigned long long* score, const uint32_t* inv, uint32_t n, uint64_t* key,
                                  uint64_t* val) {
  for (size_t x = blockIdx.x * (size_t)blockDim.x + threadIdx.x; x < n; x += (size_t)gridDim.x * blockDim.x) {
    const uint32_t v = inv[x];
    key[v] = ~(uint64_t)score[x];
    val[v] = v;
  }
}

// Degree of vertex v: out-arcs plus in-arcs (graph.hpp:324).
__device__ __forceinline__ uint32_t degree_of(const uint64_t* off1, const uint64_t* off0, uint32_t v) {
  return (uint32_t)((off1[v + 1] - off1[v]) + (off0[v + 1] - off0[v]));
}

// The SplitMix64 generator step (Steele, Lea and Flood, OOPSLA 2014): the
// seeded tie key of the degree order.
__device__ __forceinline__ uint64_t splitmix_step(uint64_t s) {
  uint64_t z = s + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// First sort keys of the degree order, val[v] = v. reference_ties: one key
// (degree descending, then id descending, distributed_memory/degree.cpp:62-67);
// otherwise the seeded tie key, which a second, stable sort by degree keeps.
__global__ void degree_keys_kernel(const uint64_t* off1, const uint64_t* off0, uint32_t n, uint64_t seed,
                                   int reference_ties, uint64_t* key, uint64_t* val) {
  for (size_t x = blockIdx.x * (size_t)blockDim.x + threadIdx.x; x < n; x += (size_t)gridDim.x * blockDim.x) {
    const uint32_t v = (uint32_t)x;
    key[v] = reference_ties ? ((uint64_t)(0xFFFFFFFFu - degree_of(off1, off0, v)) << 32) | (0xFFFFFFFFu - v)
                            : splitmix_step(seed ^ ((uint64_t)v * 0xD1B54A32D192ED03ull));
    val[v] = v;
  }
}

// key[i] = the degree key (32 bits, descending) of vertex val[i].
__global__ void degree_rekey_kernel(const uint64_t* off1, const uint64_t* off0, const uint64_t* val, uint32_t n,
                                    uint64_t* key) {
  for (size_t x = blockIdx.x * (size_t)blockDim.x + threadIdx.x; x < n; x += (size_t)gridDim.x * blockDim.x)
    key[x] = 0xFFFFFFFFu - degree_of(off1, off0, (uint32_t)val[x]);
}

unsigned grid_for(size_t count) {
  const size_t b = (count + 255) / 256;
  return (unsigned)(b < 8192 ? (b ? b : 1) : 8192);
}

}  // namespace

hipError_t launch_betweenness(const BcParams& p, uint32_t slots, hipStream_t s) {
  if (!p.R || !slots) return hipSuccess;
  hipLaunchKernelGGL(betweenness_kernel, dim3(slots), dim3(kBcThreads), 0, s, p);
  return hipGetLastError();
}

hipError_t launch_score_keys(const unsigned long long* score, const uint32_t* inv, uint32_t n, uint64_t* key,
                             uint64_t* val, hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(score_keys_kernel, dim3(grid_for(n)), dim3(256), 0, s, score, inv, n, key, val);
  return hipGetLastError();
}

hipError_t launch_degree_keys(const uint64_t* off1, const uint64_t* off0, uint32_t n, uint64_t seed, int reference_ties,
                              uint64_t* key, uint64_t* val, hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(degree_keys_kernel, dim3(grid_for(n)), dim3(256), 0, s, off1, off0, n, seed, reference_ties, key,
                     val);
  return hipGetLastError();
}

hipError_t launch_degree_rekey(const uint64_t* off1, const uint64_t* off0, const uint64_t* val, uint32_t n,
                               uint64_t* key, hipStream_t s) {
  if (!n) return hipSuccess;
  hipLaunchKernelGGL(degree_rekey_kernel, dim3(grid_for(n)), dim3(256), 0, s, off1, off0, val, n, key);
  return hipGetLastError();
}

}  // namespace chl
