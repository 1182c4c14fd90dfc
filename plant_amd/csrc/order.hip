// Vertex orders on the device.
//
// Sampled betweenness. The reference takes the order as an input file
// (shared_memory/hhl.cpp:59-62, made by an external tool, README.md:24); this
// computes the same order as the host restatement (gen.cpp order_betweenness):
// from each of R seeded roots a shortest-path tree whose parent of v is the
// smallest-id u with dist(u) + len(u, v) == dist(v); a vertex scores its proper
// descendants summed over the trees; the order is score descending, then vertex
// id ascending. Per group of roots, one per search slot:
//  1. the tree kernel (tree.hip, TREE_KEEP) runs each root's search in its own
//     slot and leaves the final distances in the slot's cells and the reached
//     vertices in its vis list;
//  2. bc_init: every reached vertex's accumulator word (children << 32 |
//     subtree) is set to (0, 1);
//  3. bc_parent: every reached vertex other than the root scans its in-arcs for
//     its parent and adds one child to the parent's word;
//  4. bc_leaves: the reached vertices without children are listed (a pass of
//     its own, so no vertex is taken for a leaf once its children have come);
//  5. bc_walk: each leaf walks up: a child adds (subtree - 2^32) to its
//     parent's word in one 64-bit atomic, so the child that brings the count to
//     zero reads the parent's final subtree size from the returned word, adds
//     subtree - 1 to the parent's score and continues upward; the others stop.
//     Counts and sizes share one word, so no fence orders them.
// Each kernel runs over (reached vertex, slot) on a two-dimensional grid.
#include <algorithm>

#include "dev.hpp"

namespace chl {
namespace {

constexpr uint32_t kNone = 0xFFFFFFFFu;

// distance of v in the slot's search (all ones when unreached: the slots hold
// one search each, on cells initialised before it)
__device__ __forceinline__ uint32_t tree_dist(const uint64_t* cell, uint32_t v) {
  return state_dist(cell[2 * (size_t)v]);
}

__global__ void bc_init_kernel(BcTrees t) {
  const uint32_t s = blockIdx.y;
  const size_t n = t.n;
  const uint32_t cnt = t.touched[s];
  const uint32_t* vis = t.vis + s * n;
  unsigned long long* acc = t.acc + s * n;
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < cnt; k += gridDim.x * blockDim.x) acc[vis[k]] = 1;
}

__global__ void bc_parent_kernel(BcTrees t) {
  const uint32_t s = blockIdx.y;
  const size_t n = t.n;
  const uint32_t cnt = t.touched[s], root = t.roots[s];
  const uint32_t* vis = t.vis + s * n;
  const uint64_t* cell = t.cell + 2 * s * n;
  uint32_t* parent = t.parent + s * n;
  unsigned long long* acc = t.acc + s * n;
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < cnt; k += gridDim.x * blockDim.x) {
    const uint32_t v = vis[k];
    if (v == root) {
      parent[v] = kNone;
      continue;
    }
    const uint32_t dv = tree_dist(cell, v);
    uint32_t best = kNone, bp = kNone;
    for (uint32_t a = t.ioff[v]; a < t.ioff[v + 1]; ++a) {
      const uint64_t arc = t.iarc[a];
      const uint32_t u = (uint32_t)(arc >> 32);
      const uint32_t du = tree_dist(cell, u);
      if (du != kNone && du + (uint32_t)arc == dv && t.inv[u] < best) {
        best = t.inv[u];
        bp = u;
      }
    }
    parent[v] = bp;
    if (bp != kNone) atomicAdd(&acc[bp], 1ull << 32);
  }
}

__global__ void bc_leaves_kernel(BcTrees t) {
  const uint32_t s = blockIdx.y;
  const size_t n = t.n;
  const uint32_t cnt = t.touched[s];
  const uint32_t* vis = t.vis + s * n;
  const unsigned long long* acc = t.acc + s * n;
  uint32_t* leaf = t.leaf + s * n;
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < cnt; k += gridDim.x * blockDim.x) {
    const uint32_t v = vis[k];
    if ((acc[v] >> 32) == 0) leaf[atomicAdd(&t.nleaf[s], 1u)] = v;
  }
}

__global__ void bc_walk_kernel(BcTrees t) {
  const uint32_t s = blockIdx.y;
  const size_t n = t.n;
  const uint32_t cnt = t.nleaf[s], root = t.roots[s];
  const uint32_t* leaf = t.leaf + s * n;
  const uint32_t* parent = t.parent + s * n;
  unsigned long long* acc = t.acc + s * n;
  for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < cnt; k += gridDim.x * blockDim.x) {
    uint32_t v = leaf[k];
    uint32_t size = 1;
    while (v != root) {
      const uint32_t pv = parent[v];
      if (pv == kNone) break;
      const unsigned long long old = atomicAdd(&acc[pv], (unsigned long long)size - (1ull << 32));
      if ((old >> 32) != 1) break;  // a sibling is still to come
      size += (uint32_t)old;
      atomicAdd(&t.score[pv], (unsigned long long)(size - 1));
      v = pv;
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

hipError_t launch_bc_trees(const BcTrees& t, unsigned max_touched, hipStream_t s) {
  if (!t.S) return hipSuccess;
  const dim3 grid(std::min<unsigned>(1024, std::max<unsigned>(1, (max_touched + 255) / 256)), t.S), block(256);
  hipLaunchKernelGGL(bc_init_kernel, grid, block, 0, s, t);
  hipLaunchKernelGGL(bc_parent_kernel, grid, block, 0, s, t);
  hipLaunchKernelGGL(bc_leaves_kernel, grid, block, 0, s, t);
  hipLaunchKernelGGL(bc_walk_kernel, grid, block, 0, s, t);
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
