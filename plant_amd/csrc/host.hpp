// Host-side data structures of the CHL engine: graph CSR, vertex order, label
// store, and the seeded generators used for the benchmark workloads.
#pragma once
#include <atomic>
#include <cstdint>
#include <string>
#include <thread>
#include <vector>

namespace chl {

typedef uint32_t Vertex;
typedef int32_t Distance;
static const Distance kInf = 2147483647;   // INT_MAX, as shared_memory/graph.hpp:47
static const Vertex kNone = 0xFFFFFFFFu;   // shared_memory/graph.hpp:46

// Directed graph as two CSRs. range 1 = out-arcs (the reference's "forward" arc
// range, graph.hpp:319 with forward=true), range 0 = in-arcs. Each row holds one
// arc per neighbour with the minimum length over parallel arcs, which is what the
// reference's dedup + merge (graph.hpp:103-153) leaves visible to Dijkstra.
struct Graph {
  int64_t n = 0;
  int64_t m_input = 0;                 // arcs counted as the reference counts them
  std::vector<uint64_t> off[2];
  std::vector<uint32_t> head[2];
  std::vector<int32_t> len[2];
  bool symmetric = false;              // every arc has an equal-length reverse arc
};

struct Arc3 {
  uint32_t u, v;
  int32_t w;
};

// Builds both ranges from directed arcs (0-based). Returns 0 or a CHL_E_* code.
int graph_from_arcs(Graph& g, int64_t n, std::vector<Arc3>& arcs, bool undirected);
// Reads DIMACS / METIS / SNAP / edgelist with autodetection in that order
// (graph.hpp:327-335). Returns the detected format 1..4, or a negative CHL_E_*.
int graph_read(Graph& g, const char* path, bool undirected);
int graph_write_dimacs(const Graph& g, const char* path);

// Seeded generators. All use splitmix64 streams so that results depend only on
// the arguments.
uint64_t splitmix64(uint64_t& s);
void gen_grid(int64_t rows, int64_t cols, int32_t wmin, int32_t wmax, uint64_t seed, std::vector<Arc3>& arcs);
void gen_knn(int64_t n, int k, uint64_t seed, std::vector<Arc3>& arcs, std::vector<double>* xy);
void gen_rmat(int scale, int64_t samples, double a, double b, double c, int32_t wmax, uint64_t seed,
              std::vector<Arc3>& arcs);

// Vertex orders: order[rank] = vertex, rank 0 most important.
// The distinct seeded roots of the sampled-betweenness order, in draw order.
void sample_roots(int64_t n, int64_t roots, uint64_t seed, std::vector<uint32_t>& rs);
void order_betweenness(const Graph& g, int64_t roots, uint64_t seed, int threads, std::vector<uint32_t>& order,
                       std::vector<uint64_t>* score);
void order_degree(const Graph& g, uint64_t seed, bool reference_ties, std::vector<uint32_t>& order);
int order_read(const char* path, std::vector<uint32_t>& order);
int order_write(const char* path, const std::vector<uint32_t>& order);

// Host-resident labels in per-side CSR form, rows sorted by hub rank.
struct Labels {
  int64_t n = 0;
  std::vector<uint64_t> off[2];
  std::vector<uint64_t> ent[2];  // (hub << 32) | (uint32)dist
};
int labels_write(const Labels& L, const char* path);
int labels_read(Labels& L, const char* path);

// Runs f(i) for i in [0, count) on up to `threads` std::threads in contiguous
// chunks of `grain`.
template <class F>
void parallel_for(int64_t count, int threads, int64_t grain, F f) {
  if (threads <= 1 || count <= grain) {
    for (int64_t i = 0; i < count; ++i) f(i);
    return;
  }
  std::vector<std::thread> ts;
  std::atomic_int64_t next(0);
  for (int t = 0; t < threads; ++t)
    ts.emplace_back([&]() {
      while (true) {
        int64_t s = next.fetch_add(grain);
        if (s >= count) break;
        int64_t e = s + grain < count ? s + grain : count;
        for (int64_t i = s; i < e; ++i) f(i);
      }
    });
  for (auto& t : ts) t.join();
}

int host_threads();

}  // namespace chl
