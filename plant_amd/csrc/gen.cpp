// Seeded workload generators (grid, k-nearest-neighbour road-like graphs, R-MAT)
// and vertex orders (sampled betweenness, degree). The reference takes its
// inputs from external tools (README.md:24-25) and its degree order from
// distributed_memory/degree.cpp:40-70; everything here is specified by its
// arguments alone, so the same call gives the same graph on any machine.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <queue>
#include <thread>

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

// The trees run on a copy of the graph renumbered in breadth-first order from a
// vertex of maximum degree (over both arc ranges), so that a search front
// touches few cache lines; the parent rule still compares original ids. With
// arc lengths below kDialMax the queue is a circular array of one bucket per
// distance (Dial's algorithm: every vertex is popped at its final distance in
// O(arcs + max distance) per tree); longer arcs use a binary heap.
namespace {
constexpr int32_t kDialMax = 1 << 16;

struct LocalCsr {
  std::vector<uint64_t> off;
  std::vector<uint32_t> head;
  std::vector<int32_t> len;
};

void local_csr(const Graph& g, int range, const std::vector<uint32_t>& inv, const std::vector<uint32_t>& pos,
               LocalCsr& c) {
  const int64_t n = g.n;
  c.off.assign(n + 1, 0);
  for (int64_t i = 0; i < n; ++i) c.off[i + 1] = c.off[i] + (g.off[range][inv[i] + 1] - g.off[range][inv[i]]);
  c.head.resize(c.off[n]);
  c.len.resize(c.off[n]);
  parallel_for(n, host_threads(), 4096, [&](int64_t i) {
    uint64_t o = c.off[i];
    for (uint64_t a = g.off[range][inv[i]]; a < g.off[range][inv[i] + 1]; ++a, ++o) {
      c.head[o] = pos[g.head[range][a]];
      c.len[o] = g.len[range][a];
    }
  });
}
}  // namespace

void order_betweenness(const Graph& g, int64_t roots, uint64_t seed, int threads, std::vector<uint32_t>& order,
                       std::vector<uint64_t>* score_out) {
  const int64_t n = g.n;
  roots = std::min<int64_t>(roots, n);
  std::vector<uint32_t> rs;
  sample_roots(n, roots, seed, rs);
  if (threads <= 0) threads = host_threads();
  // breadth-first renumbering (inv: new id -> vertex, pos: vertex -> new id)
  std::vector<uint32_t> inv, pos(n, kNone);
  inv.reserve(n);
  {
    int64_t best = 0;
    uint64_t bdeg = 0;
    for (int64_t v = 0; v < n; ++v) {
      const uint64_t d = (g.off[1][v + 1] - g.off[1][v]) + (g.off[0][v + 1] - g.off[0][v]);
      if (d > bdeg) {
        bdeg = d;
        best = v;
      }
    }
    for (int64_t k = -1; k < n; ++k) {
      const uint32_t s = (uint32_t)(k < 0 ? best : k);
      if (pos[s] != kNone) continue;
      size_t head = inv.size();
      pos[s] = (uint32_t)inv.size();
      inv.push_back(s);
      while (head < inv.size()) {
        const uint32_t u = inv[head++];
        for (int rg = 1; rg >= 0; --rg)
          for (uint64_t a = g.off[rg][u]; a < g.off[rg][u + 1]; ++a) {
            const uint32_t v = g.head[rg][a];
            if (pos[v] == kNone) {
              pos[v] = (uint32_t)inv.size();
              inv.push_back(v);
            }
          }
      }
    }
  }
  LocalCsr out, in;
  local_csr(g, 1, inv, pos, out);
  local_csr(g, 0, inv, pos, in);
  int32_t maxw = 0;
  for (int32_t w : out.len) maxw = std::max(maxw, w);
  const bool dial = maxw < kDialMax;
  const size_t nb = dial ? (size_t)maxw + 1 : 1;

  int T = (int)std::min<int64_t>(threads, std::max<int64_t>(1, roots));
  std::vector<std::vector<uint64_t>> part(T);
  std::atomic_int64_t next(0);
  std::vector<std::thread> ts;
  for (int t = 0; t < T; ++t)
    ts.emplace_back([&, t]() {
      part[t].assign(n, 0);
      std::vector<int32_t> dist(n, kInf);
      std::vector<uint32_t> sub(n, 0), settled;
      settled.reserve(n);
      std::vector<std::vector<uint32_t>> bucket(nb);
      typedef std::pair<int32_t, uint32_t> QE;
      while (true) {
        const int64_t k = next.fetch_add(1);
        if (k >= roots) break;
        const uint32_t src = pos[rs[k]];
        settled.clear();
        dist[src] = 0;
        if (dial) {
          // every queued distance lies in [cur, cur + maxw]: bucket d % nb
          size_t queued = 1;
          bucket[0].push_back(src);
          for (int64_t cur = 0; queued; ++cur) {
            std::vector<uint32_t>& B = bucket[(size_t)(cur % (int64_t)nb)];
            for (size_t i = 0; i < B.size(); ++i) {  // B may grow only through zero-length arcs, which are absent
              const uint32_t u = B[i];
              --queued;
              if (dist[u] != cur) continue;
              settled.push_back(u);
              for (uint64_t a = out.off[u]; a < out.off[u + 1]; ++a) {
                const int32_t nd = (int32_t)cur + out.len[a];
                const uint32_t h = out.head[a];
                if (nd < dist[h]) {
                  dist[h] = nd;
                  bucket[(size_t)(nd % (int64_t)nb)].push_back(h);
                  ++queued;
                }
              }
            }
            B.clear();
          }
        } else {
          std::priority_queue<QE, std::vector<QE>, std::greater<QE>> pq;
          pq.push({0, src});
          while (!pq.empty()) {
            QE q = pq.top();
            pq.pop();
            if (q.first != dist[q.second]) continue;
            const uint32_t u = q.second;
            settled.push_back(u);
            for (uint64_t a = out.off[u]; a < out.off[u + 1]; ++a) {
              const int32_t nd = q.first + out.len[a];
              const uint32_t h = out.head[a];
              if (nd < dist[h]) {
                dist[h] = nd;
                pq.push({nd, h});
              }
            }
          }
        }
        // descendants accumulate child -> parent in reverse settle order; the
        // parent of v is its shortest-path predecessor of smallest original id
        for (uint32_t v : settled) sub[v] = 1;
        for (int64_t idx = (int64_t)settled.size() - 1; idx >= 1; --idx) {
          const uint32_t v = settled[idx];
          uint32_t p = kNone, pid = kNone;
          for (uint64_t a = in.off[v]; a < in.off[v + 1]; ++a) {
            const uint32_t u = in.head[a];
            if (dist[u] != kInf && (int64_t)dist[u] + in.len[a] == dist[v] && inv[u] < pid) {
              pid = inv[u];
              p = u;
            }
          }
          sub[p] += sub[v];
          part[t][v] += sub[v] - 1;
        }
        part[t][src] += sub[src] - 1;
        for (uint32_t v : settled) {
          dist[v] = kInf;
          sub[v] = 0;
        }
      }
    });
  for (auto& th : ts) th.join();
  // per-thread scores are kept by new id (written in search order, so the
  // accumulation stays local); summed and mapped back to vertex ids here
  std::vector<uint64_t> score(n, 0);
  parallel_for(n, host_threads(), 1 << 16, [&](int64_t i) {
    uint64_t s = 0;
    for (int t = 0; t < T; ++t) s += part[t][i];
    score[inv[i]] = s;
  });
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
