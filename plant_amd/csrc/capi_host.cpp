// C ABI: graph, generator, order and error entry points (host only).
#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "../../include/chl.h"
#include "host.hpp"
#include "internal.hpp"

using namespace chl;

namespace chl {
static thread_local std::string g_last_error;
int set_error(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}
}  // namespace chl

extern "C" {

const char* chl_strerror(int code) {
  switch (code) {
    case CHL_OK: return "ok";
    case CHL_E_INVAL: return "invalid argument";
    case CHL_E_IO: return "file cannot be opened or written";
    case CHL_E_FORMAT: return "unrecognised or malformed file";
    case CHL_E_WEIGHT: return "arc length must be positive";
    case CHL_E_ORDER: return "order is not a permutation of the vertices";
    case CHL_E_NOMEM: return "out of memory";
    case CHL_E_DEVICE: return "HIP device error";
    case CHL_E_RANGE: return "value out of supported range";
    case CHL_E_BUSY: return "admission refused";
    case CHL_E_STATE: return "invalid state";
    default: return "unknown error";
  }
}

const char* chl_last_error(void) { return g_last_error.c_str(); }

int chl_graph_load(const char* path, int undirected, chl_graph** out, int* format) {
  if (!path || !out) return set_error(CHL_E_INVAL, "null argument");
  chl_graph* g = new chl_graph();
  int r = graph_read(g->g, path, undirected != 0);
  if (r < 0) {
    delete g;
    *out = nullptr;
    return set_error(r, "cannot read graph from %s: %s", path, chl_strerror(r));
  }
  if (format) *format = r;
  *out = g;
  return CHL_OK;
}

int chl_graph_from_arcs(int64_t n, int64_t m, const uint32_t* src, const uint32_t* dst, const int32_t* len,
                        int undirected, chl_graph** out) {
  if (!out || n <= 0 || m < 0 || (m > 0 && (!src || !dst || !len))) return set_error(CHL_E_INVAL, "bad arguments");
  std::vector<Arc3> arcs((size_t)m);
  for (int64_t k = 0; k < m; ++k) arcs[k] = {src[k], dst[k], len[k]};
  chl_graph* g = new chl_graph();
  int r = graph_from_arcs(g->g, n, arcs, undirected != 0);
  if (r != CHL_OK) {
    delete g;
    *out = nullptr;
    return set_error(r, "graph_from_arcs: %s", chl_strerror(r));
  }
  *out = g;
  return CHL_OK;
}

int chl_graph_write(const chl_graph* g, const char* path) {
  if (!g || !path) return set_error(CHL_E_INVAL, "null argument");
  int r = graph_write_dimacs(g->g, path);
  return r == CHL_OK ? r : set_error(r, "cannot write %s", path);
}

int64_t chl_graph_n(const chl_graph* g) { return g ? g->g.n : -1; }
int64_t chl_graph_m(const chl_graph* g) { return g ? g->g.m_input : -1; }
int64_t chl_graph_arcs(const chl_graph* g, int range) {
  if (!g || range < 0 || range > 1) return -1;
  return (int64_t)g->g.head[range].size();
}
int chl_graph_symmetric(const chl_graph* g) { return g && g->g.symmetric ? 1 : 0; }

int chl_graph_csr(const chl_graph* g, int range, uint64_t* off, uint32_t* head, int32_t* len) {
  if (!g || range < 0 || range > 1 || !off) return set_error(CHL_E_INVAL, "bad arguments");
  const Graph& G = g->g;
  memcpy(off, G.off[range].data(), sizeof(uint64_t) * (G.n + 1));
  if (head) memcpy(head, G.head[range].data(), sizeof(uint32_t) * G.head[range].size());
  if (len) memcpy(len, G.len[range].data(), sizeof(int32_t) * G.len[range].size());
  return CHL_OK;
}

void chl_graph_free(chl_graph* g) { delete g; }

static int finish_gen(int64_t n, std::vector<Arc3>& arcs, chl_graph** out) {
  chl_graph* g = new chl_graph();
  int r = graph_from_arcs(g->g, n, arcs, true);
  if (r != CHL_OK) { delete g; *out = nullptr; return set_error(r, "generator: %s", chl_strerror(r)); }
  *out = g;
  return CHL_OK;
}

int chl_gen_grid(int64_t rows, int64_t cols, int32_t wmin, int32_t wmax, uint64_t seed, chl_graph** out) {
  if (!out || rows <= 0 || cols <= 0 || wmin <= 0 || wmax < wmin) return set_error(CHL_E_INVAL, "bad grid arguments");
  std::vector<Arc3> arcs;
  gen_grid(rows, cols, wmin, wmax, seed, arcs);
  return finish_gen(rows * cols, arcs, out);
}

int chl_gen_knn(int64_t n, int k, uint64_t seed, chl_graph** out) {
  if (!out || n < 2 || k < 1 || k >= n) return set_error(CHL_E_INVAL, "bad knn arguments");
  std::vector<Arc3> arcs;
  gen_knn(n, k, seed, arcs, nullptr);
  return finish_gen(n, arcs, out);
}

int chl_gen_rmat(int scale, int64_t samples, double a, double b, double c, int32_t wmax, uint64_t seed,
                 chl_graph** out) {
  if (!out || scale < 1 || scale > 29 || samples < 0 || a < 0 || b < 0 || c < 0 || a + b + c > 1.0)
    return set_error(CHL_E_INVAL, "bad rmat arguments");
  std::vector<Arc3> arcs;
  gen_rmat(scale, samples, a, b, c, wmax, seed, arcs);
  return finish_gen((int64_t)1 << scale, arcs, out);
}

int chl_order_load(const char* path, uint32_t* order, int64_t cap, int64_t* n) {
  if (!path) return set_error(CHL_E_INVAL, "null path");
  std::vector<uint32_t> o;
  int r = order_read(path, o);
  if (r != CHL_OK) return set_error(r, "cannot read order %s", path);
  if (n) *n = (int64_t)o.size();
  if (order)
    for (int64_t i = 0; i < (int64_t)o.size() && i < cap; ++i) order[i] = o[i];
  return CHL_OK;
}

int chl_order_write(const char* path, const uint32_t* order, int64_t n) {
  if (!path || (n > 0 && !order) || n < 0) return set_error(CHL_E_INVAL, "bad arguments");
  std::vector<uint32_t> o(order, order + n);
  int r = order_write(path, o);
  return r == CHL_OK ? r : set_error(r, "cannot write order %s", path);
}

int chl_order_betweenness_scored(const chl_graph* g, int64_t roots, uint64_t seed, uint32_t* order, float* score) {
  if (!g || !order || roots <= 0) return set_error(CHL_E_INVAL, "bad arguments");
  std::vector<uint32_t> o;
  std::vector<uint64_t> sc;
  order_betweenness(g->g, roots, seed, host_threads(), o, score ? &sc : nullptr);
  memcpy(order, o.data(), sizeof(uint32_t) * o.size());
  if (score)
    for (size_t v = 0; v < sc.size(); ++v) score[v] = (float)sc[v];
  return CHL_OK;
}

int chl_order_betweenness(const chl_graph* g, int64_t roots, uint64_t seed, uint32_t* order) {
  return chl_order_betweenness_scored(g, roots, seed, order, nullptr);
}

int chl_order_degree(const chl_graph* g, uint64_t seed, int reference_ties, uint32_t* order) {
  if (!g || !order) return set_error(CHL_E_INVAL, "bad arguments");
  std::vector<uint32_t> o;
  order_degree(g->g, seed, reference_ties != 0, o);
  memcpy(order, o.data(), sizeof(uint32_t) * o.size());
  return CHL_OK;
}

int chl_gen_local_pairs(const chl_graph* g, int64_t q, int hops, uint64_t seed, uint32_t* pairs) {
  if (!g || q < 0 || hops < 0 || (q > 0 && !pairs) || g->g.n == 0) return set_error(CHL_E_INVAL, "bad arguments");
  const Graph& G = g->g;
  const int64_t n = G.n;
  const int threads = (int)std::max<int64_t>(1, std::min<int64_t>(host_threads(), (q + 255) / 256));
  // one breadth-first ball per pair; every thread has its own visit stamps and
  // takes the pairs k = t, t + threads, ...
  auto work = [&](int t) {
    std::vector<uint32_t> st(n, 0), b;
    for (int64_t k = t; k < q; k += threads) {
      uint64_t s = seed ^ (0x9E3779B97F4A7C15ull * (uint64_t)(k + 1));
      const uint32_t u = (uint32_t)(splitmix64(s) % (uint64_t)n);
      const uint32_t mark = (uint32_t)(k / threads + 1);
      b.clear();
      b.push_back(u);
      st[u] = mark;
      size_t lo = 0;
      for (int h = 0; h < hops && lo < b.size(); ++h) {
        const size_t hi = b.size();
        for (size_t i = lo; i < hi; ++i)
          for (uint64_t a = G.off[1][b[i]]; a < G.off[1][b[i] + 1]; ++a) {
            const uint32_t w = G.head[1][a];
            if (st[w] != mark) {
              st[w] = mark;
              b.push_back(w);
            }
          }
        lo = hi;
      }
      pairs[2 * k] = u;
      pairs[2 * k + 1] = b[splitmix64(s) % b.size()];
    }
  };
  std::vector<std::thread> ts;
  for (int t = 1; t < threads; ++t) ts.emplace_back(work, t);
  work(0);
  for (auto& t : ts) t.join();
  return CHL_OK;
}

// ---- query sources and results (distributed_memory/query.hpp:18-51) ---------

int chl_query_generate(int64_t q, uint32_t range, uint32_t* pairs) {
  if (q < 0 || (q > 0 && (!pairs || range == 0))) return set_error(CHL_E_INVAL, "bad arguments");
  // the C library's additive generator in its own state, seeded 1: the stream
  // that an unseeded rand() gives
  char statebuf[128];
  struct random_data rd;
  memset(&rd, 0, sizeof(rd));
  if (initstate_r(1u, statebuf, sizeof(statebuf), &rd) != 0) return set_error(CHL_E_INVAL, "generator state");
  for (int64_t i = 0; i < 2 * q; ++i) {
    int32_t r = 0;
    random_r(&rd, &r);
    pairs[i] = (uint32_t)r % range;
  }
  return CHL_OK;
}

int chl_query_load(const char* path, uint32_t* pairs, int64_t cap, int64_t* q) {
  if (!path || !q) return set_error(CHL_E_INVAL, "bad arguments");
  FILE* f = fopen(path, "r");
  if (!f) return set_error(CHL_E_IO, "cannot open %s", path);
  char buf[512];
  int64_t k = 0;
  int rc = CHL_OK;
  while (fgets(buf, sizeof(buf), f)) {
    const size_t len = strlen(buf);
    long long u = 0, v = 0;
    if ((len == 0 || buf[len - 1] != '\n') && !feof(f)) {
      rc = set_error(CHL_E_FORMAT, "%s: line %lld too long", path, (long long)k + 1);
      break;
    }
    if (sscanf(buf, "%lld %lld", &u, &v) != 2 || u < 0 || v < 0 || u > 0xFFFFFFFFll || v > 0xFFFFFFFFll) {
      rc = set_error(CHL_E_FORMAT, "%s: line %lld is not a pair of vertex ids", path, (long long)k + 1);
      break;
    }
    if (pairs && k < cap) {
      pairs[2 * k] = (uint32_t)u;
      pairs[2 * k + 1] = (uint32_t)v;
    }
    ++k;
  }
  fclose(f);
  if (rc == CHL_OK) *q = k;
  return rc;
}

int chl_query_write(const char* path, const int32_t* out, int64_t q) {
  if (!path || q < 0 || (q > 0 && !out)) return set_error(CHL_E_INVAL, "bad arguments");
  FILE* f = fopen(path, "w");
  if (!f) return set_error(CHL_E_IO, "cannot open %s", path);
  for (int64_t i = 0; i < q; ++i) fprintf(f, "%d\n", out[i]);
  const bool ok = !ferror(f);
  if (fclose(f) != 0 || !ok) return set_error(CHL_E_IO, "cannot write %s", path);
  return CHL_OK;
}

}  // extern "C"
