// Seeded workload generators (grid, k-nearest-neighbour road-like graphs, R-MAT)
// and vertex orders (sampled betweenness, degree). The reference takes its
// inputs from external tools (README.md:24-25) and its degree order from
// distributed_memory/degree.cpp:40-70; everything here is specified by its
// arguments alone, so the same call gives the same graph on any machine.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <queue>

#include "../../include/chl.h"
#include "host.hpp"

namespace chl {

uint64_t splitmix64(uint64_t& s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

static inline double u01(uint64_t& s) { return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }

// rows x cols 4-neighbour grid, id = r*cols + c; edges in row-major order, the
// right edge before the down edge; weight = wmin + draw % (wmax - wmin + 1).
void gen_grid(int64_t rows, int64_t cols, int32_t wmin, int32_t wmax, uint64_t seed, std::vector<Arc3>& arcs) {
  uint64_t s = seed;
  uint64_t span = (uint64_t)(wmax - wmin) + 1;
  arcs.clear();
  for (int64_t r = 0; r < rows; ++r)
    for (int64_t c = 0; c < cols; ++c) {
      uint32_t id = (uint32_t)(r * cols + c);
      if (c + 1 < cols) arcs.push_back({id, id + 1, (int32_t)(wmin + splitmix64(s) % span)});
      if (r + 1 < rows) arcs.push_back({id, (uint32_t)(id + cols), (int32_t)(wmin + splitmix64(s) % span)});
    }
}

// n points uniform in [0,1)^2 (x then y per point from one stream); each point is
// joined to its k nearest other points (squared distance, ties by index), the
// edge set is symmetrised and deduplicated, and an edge of euclidean length L
// weighs floor(1e4*L + 0.5) + 1. Arcs are emitted once per undirected edge
// (u < v), sorted.
void gen_knn(int64_t n, int k, uint64_t seed, std::vector<Arc3>& arcs, std::vector<double>* xy_out) {
  uint64_t s = seed;
  std::vector<double> xy(2 * n);
  for (int64_t i = 0; i < n; ++i) { xy[2 * i] = u01(s); xy[2 * i + 1] = u01(s); }
  const int64_t G = std::max<int64_t>(1, (int64_t)std::sqrt((double)n / 2.0));
  std::vector<uint32_t> cell_start(G * G + 1, 0), cell_pts(n);
  auto cell_of = [&](int64_t i) {
    int64_t cx = std::min<int64_t>(G - 1, (int64_t)(xy[2 * i] * G));
    int64_t cy = std::min<int64_t>(G - 1, (int64_t)(xy[2 * i + 1] * G));
    return cy * G + cx;
  };
  for (int64_t i = 0; i < n; ++i) cell_start[cell_of(i) + 1]++;
  for (int64_t c = 0; c < G * G; ++c) cell_start[c + 1] += cell_start[c];
  {
    std::vector<uint32_t> pos(cell_start.begin(), cell_start.end() - 1);
    for (int64_t i = 0; i < n; ++i) cell_pts[pos[cell_of(i)]++] = (uint32_t)i;
  }
  std::vector<uint32_t> nbr((size_t)n * k, 0xFFFFFFFFu);
  parallel_for(n, host_threads(), 1024, [&](int64_t i) {
    const double px = xy[2 * i], py = xy[2 * i + 1];
    std::vector<std::pair<double, uint32_t>> best;  // up to k, sorted
    int64_t cx = std::min<int64_t>(G - 1, (int64_t)(px * G)), cy = std::min<int64_t>(G - 1, (int64_t)(py * G));
    for (int64_t ring = 0;; ++ring) {
      for (int64_t yy = cy - ring; yy <= cy + ring; ++yy)
        for (int64_t xx = cx - ring; xx <= cx + ring; ++xx) {
          if (std::max(std::llabs(yy - cy), std::llabs(xx - cx)) != ring) continue;
          if (xx < 0 || yy < 0 || xx >= G || yy >= G) continue;
          int64_t c = yy * G + xx;
          for (uint32_t t = cell_start[c]; t < cell_start[c + 1]; ++t) {
            uint32_t j = cell_pts[t];
            if (j == (uint32_t)i) continue;
            double dx = xy[2 * j] - px, dy = xy[2 * j + 1] - py;
            std::pair<double, uint32_t> cand(dx * dx + dy * dy, j);
            if ((int)best.size() < k || cand < best.back()) {
              best.insert(std::upper_bound(best.begin(), best.end(), cand), cand);
              if ((int)best.size() > k) best.pop_back();
            }
          }
        }
      // every point outside the searched square is at least ring/G away
      double reach = (double)ring / (double)G;
      if ((int)best.size() == k && best.back().first < reach * reach) break;
      if (ring > 2 * G) break;
    }
    for (size_t t = 0; t < best.size(); ++t) nbr[(size_t)i * k + t] = best[t].second;
  });
  std::vector<uint64_t> e;
  e.reserve((size_t)n * k);
  for (int64_t i = 0; i < n; ++i)
    for (int t = 0; t < k; ++t) {
      uint32_t j = nbr[(size_t)i * k + t];
      if (j == 0xFFFFFFFFu) continue;
      uint32_t a = std::min<uint32_t>((uint32_t)i, j), b = std::max<uint32_t>((uint32_t)i, j);
      e.push_back(((uint64_t)a << 32) | b);
    }
  std::sort(e.begin(), e.end());
  e.erase(std::unique(e.begin(), e.end()), e.end());
  arcs.clear();
  arcs.reserve(e.size());
  for (uint64_t x : e) {
    uint32_t a = (uint32_t)(x >> 32), b = (uint32_t)x;
    double dx = xy[2 * a] - xy[2 * b], dy = xy[2 * a + 1] - xy[2 * b + 1];
    double L = std::sqrt(dx * dx + dy * dy);
    arcs.push_back({a, b, (int32_t)std::floor(1e4 * L + 0.5) + 1});
  }
  if (xy_out) xy_out->swap(xy);
}

// R-MAT: `samples` edges, each choosing one of four quadrants per level with
// probabilities (a, b, c, 1-a-b-c); symmetrised and deduplicated on the
// unordered pair, self-loops dropped. Weights are 1 when wmax <= 1, else
// U[1, wmax] drawn per unordered edge in sorted edge order from a second stream.
void gen_rmat(int scale, int64_t samples, double a, double b, double c, int32_t wmax, uint64_t seed,
              std::vector<Arc3>& arcs) {
  uint64_t s = seed;
  std::vector<uint64_t> e;
  e.reserve((size_t)samples);
  for (int64_t t = 0; t < samples; ++t) {
    uint32_t u = 0, v = 0;
    for (int l = 0; l < scale; ++l) {
      double r = u01(s);
      uint32_t bu = 0, bv = 0;
      if (r < a) {
      } else if (r < a + b) {
        bv = 1;
      } else if (r < a + b + c) {
        bu = 1;
      } else {
        bu = 1;
        bv = 1;
      }
      u = (u << 1) | bu;
      v = (v << 1) | bv;
    }
    if (u == v) continue;
    uint32_t lo = std::min(u, v), hi = std::max(u, v);
    e.push_back(((uint64_t)lo << 32) | hi);
  }
  std::sort(e.begin(), e.end());
  e.erase(std::unique(e.begin(), e.end()), e.end());
  uint64_t ws = seed ^ 0x5DEECE66DULL;
  arcs.clear();
  arcs.reserve(e.size());
  for (uint64_t x : e) {
    int32_t w = wmax <= 1 ? 1 : (int32_t)(1 + splitmix64(ws) % (uint64_t)wmax);
    arcs.push_back({(uint32_t)(x >> 32), (uint32_t)x, w});
  }
}

// Sampled betweenness: `roots` distinct seeded roots; from each, a Dijkstra
// shortest-path tree over out-arcs whose parent of v is the smallest-id u with
// dist(u) + len(u,v) == dist(v); a vertex scores the number of its proper
// descendants, summed over the trees. Order: score descending, then id ascending.
void sample_roots(int64_t n, int64_t roots, uint64_t seed, std::vector<uint32_t>& rs) {
  roots = std::min<int64_t>(roots, n);
  rs.clear();
  std::vector<uint8_t> used(n, 0);
  uint64_t s = seed;
  while ((int64_t)rs.size() < roots) {
    uint32_t r = (uint32_t)(splitmix64(s) % (uint64_t)n);
    if (!used[r]) { used[r] = 1; rs.push_back(r); }
  }
}

void order_betweenness(const Graph& g, int64_t roots, uint64_t seed, int threads, std::vector<uint32_t>& order,
                       std::vector<uint64_t>* score_out) {
  const int64_t n = g.n;
  roots = std::min<int64_t>(roots, n);
  std::vector<uint32_t> rs;
  sample_roots(n, roots, seed, rs);
  if (threads <= 0) threads = host_threads();
  int T = (int)std::min<int64_t>(threads, std::max<int64_t>(1, roots));
  std::vector<std::vector<uint64_t>> part(T, std::vector<uint64_t>(n, 0));
  std::atomic_int64_t next(0);
  std::vector<std::thread> ts;
  for (int t = 0; t < T; ++t)
    ts.emplace_back([&, t]() {
      std::vector<int32_t> dist(n, kInf);
      std::vector<uint32_t> sub(n, 0), settled;
      settled.reserve(n);
      typedef std::pair<int32_t, uint32_t> QE;
      while (true) {
        int64_t k = next.fetch_add(1);
        if (k >= roots) break;
        uint32_t src = rs[k];
        std::priority_queue<QE, std::vector<QE>, std::greater<QE>> pq;
        settled.clear();
        dist[src] = 0;
        pq.push({0, src});
        while (!pq.empty()) {
          QE q = pq.top();
          pq.pop();
          if (q.first != dist[q.second]) continue;
          uint32_t u = q.second;
          settled.push_back(u);
          for (uint64_t a = g.off[1][u]; a < g.off[1][u + 1]; ++a) {
            int32_t nd = q.first + g.len[1][a];
            uint32_t h = g.head[1][a];
            if (nd < dist[h]) { dist[h] = nd; pq.push({nd, h}); }
          }
        }
        // descendants accumulate child -> parent in reverse settle order
        for (uint32_t v : settled) sub[v] = 1;
        for (int64_t idx = (int64_t)settled.size() - 1; idx >= 1; --idx) {
          uint32_t v = settled[idx];
          uint32_t p = kNone;
          for (uint64_t a = g.off[0][v]; a < g.off[0][v + 1]; ++a) {
            uint32_t u = g.head[0][a];
            if (dist[u] != kInf && (int64_t)dist[u] + g.len[0][a] == dist[v] && u < p) p = u;
          }
          sub[p] += sub[v];
          part[t][v] += sub[v] - 1;
        }
        part[t][src] += sub[src] - 1;
        for (uint32_t v : settled) { dist[v] = kInf; sub[v] = 0; }
      }
    });
  for (auto& th : ts) th.join();
  std::vector<uint64_t> score(n, 0);
  for (int t = 0; t < T; ++t)
    for (int64_t v = 0; v < n; ++v) score[v] += part[t][v];
  order.resize(n);
  for (int64_t v = 0; v < n; ++v) order[v] = (uint32_t)v;
  std::sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) {
    if (score[x] != score[y]) return score[x] > score[y];
    return x < y;
  });
  if (score_out) score_out->swap(score);
}

// Degree order: out-degree + in-degree descending (graph.hpp:324). With
// reference_ties the order is exactly distributed_memory/degree.cpp:62-67
// (ties: higher id first); otherwise ties break by a seeded key, then id.
void order_degree(const Graph& g, uint64_t seed, bool reference_ties, std::vector<uint32_t>& order) {
  const int64_t n = g.n;
  std::vector<uint64_t> deg(n), key(n);
  for (int64_t v = 0; v < n; ++v) {
    deg[v] = (g.off[1][v + 1] - g.off[1][v]) + (g.off[0][v + 1] - g.off[0][v]);
    uint64_t s = seed ^ ((uint64_t)v * 0xD1B54A32D192ED03ULL);
    key[v] = splitmix64(s);
  }
  order.resize(n);
  for (int64_t v = 0; v < n; ++v) order[v] = (uint32_t)v;
  std::sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) {
    if (deg[x] != deg[y]) return deg[x] > deg[y];
    if (reference_ties) return x > y;
    if (key[x] != key[y]) return key[x] < key[y];
    return x < y;
  });
}

}  // namespace chl
