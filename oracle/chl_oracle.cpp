// CPU ORACLE for canonical hub labeling — TEST INFRASTRUCTURE ONLY.
//
// This file restates, in plain sequential / OpenMP C++, the reference algorithms
// of qingdo/plant that the GPU engine must reproduce bit-exactly. It is linked
// only by tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg,
// always as the checker, never as the thing measured or shipped. The product
// library (plant_amd/) never links or calls it.
//
// Every function cites the reference file:line it follows. Pinning (what checks
// each function against something other than itself):
//   * orc_graph_*           : hand-written DIMACS/METIS/SNAP/edgelist fixtures with known
//                             arc sets (tests/test_host.py) and the reference `hhl`
//                             binary's labels on the same files.
//   * orc_dijkstra          : scipy.sparse.csgraph.dijkstra golden vectors and
//                             Floyd-Warshall on tiny graphs.
//   * orc_gll               : the reference shared_memory `hhl` built from source
//                             (oracle/_ref/hhl) — byte-identical label files — and the
//                             brute-force CHL definition (orc_brute_chl).
//   * orc_plant_range       : equality with orc_gll and with the brute-force CHL
//                             (tests/test_oracle.py); closed forms (path / star graphs).
//   * orc_hybrid            : equality with orc_gll (CHL is unique for a graph and order);
//                             the distributed reference needs MPI and is not built here.
//   * orc_clean_cover       : runs inside orc_gll's cleaning, so the byte-identical
//                             labels against the reference binary pin it.
//   * orc_query             : Dijkstra on 1e5 pairs and all pairs on tiny graphs.
//   * orc_brute_chl         : closed forms (path with midpoint order, star) and the
//                             2-hop cover property against Floyd-Warshall.
#include <algorithm>
#include <atomic>
#include <cassert>
#include <climits>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctype.h>
#include <fstream>
#include <omp.h>
#include <set>
#include <string>
#include <utility>
#include <vector>

namespace orc {

typedef uint32_t Vertex;
typedef int32_t Distance;
static const Distance INF = INT_MAX;  // shared_memory/graph.hpp:47
static const Vertex NONE = 0xFFFFFFFFu; // shared_memory/graph.hpp:46

// ---------------------------------------------------------------------------
// Graph: in range [0] (arcs u<-x, "reverse" flag) and out range [1] (arcs u->x,
// "forward" flag), as in shared_memory/graph.hpp:61-67,319-320. Storage is two
// CSRs; a merged bidirectional arc of the reference appears in both, exactly as a
// reference Arc with both flags is visible in both ranges.
// ---------------------------------------------------------------------------
struct TmpArc {
  Vertex tail, head;
  Distance len;
  uint8_t fwd, rev;
};

struct Graph {
  int64_t n = 0;
  int64_t m = 0;  // reference's arc counter: += 1 + undirected per add (graph.hpp:160)
  std::vector<uint64_t> off[2];
  std::vector<Vertex> head[2];
  std::vector<Distance> len[2];
  std::vector<TmpArc> tmp;
};

// graph.hpp:156-162
static bool add_tmp_arc(Graph& g, Vertex u, Vertex v, Distance w, bool undirected) {
  if ((int64_t)u >= g.n || (int64_t)v >= g.n) return false;
  g.tmp.push_back({u, v, w, 1, (uint8_t)undirected});
  g.tmp.push_back({v, u, w, (uint8_t)undirected, 1});
  g.m += 1 + (undirected ? 1 : 0);
  return true;
}

// graph.hpp:103-153. Step 1: among arcs with equal (tail, head, fwd, rev) keep the
// shortest. Step 2: arcs with equal (tail, head, length) merge their flags. The
// range an arc lands in is decided by its flags, so after both steps the out
// range of u holds, per head, the shortest u->head arc, and the in range holds the
// shortest head->u arc.
static void finalize(Graph& g) {
  std::vector<TmpArc>& a = g.tmp;
  std::sort(a.begin(), a.end(), [](const TmpArc& x, const TmpArc& y) {
    if (x.tail != y.tail) return x.tail < y.tail;
    if (x.rev != y.rev) return x.rev > y.rev;
    if (x.fwd != y.fwd) return x.fwd < y.fwd;
    if (x.head != y.head) return x.head < y.head;
    return x.len < y.len;
  });
  size_t k = 0;
  for (size_t j = 1; j < a.size(); ++j) {
    if (a[k].tail == a[j].tail && a[k].head == a[j].head && a[k].fwd == a[j].fwd &&
        a[k].rev == a[j].rev)
      continue;
    a[++k] = a[j];
  }
  if (!a.empty()) a.resize(k + 1);
  std::sort(a.begin(), a.end(), [](const TmpArc& x, const TmpArc& y) {
    if (x.tail != y.tail) return x.tail < y.tail;
    if (x.head != y.head) return x.head < y.head;
    return x.len < y.len;
  });
  k = 0;
  for (size_t j = 1; j < a.size(); ++j) {
    if (a[k].tail == a[j].tail && a[k].head == a[j].head && a[k].len == a[j].len) {
      a[k].fwd |= a[j].fwd;
      a[k].rev |= a[j].rev;
      continue;
    }
    a[++k] = a[j];
  }
  if (!a.empty()) a.resize(k + 1);
  for (int s = 0; s < 2; ++s) {
    g.off[s].assign(g.n + 1, 0);
    g.head[s].clear();
    g.len[s].clear();
  }
  // range 0 = arcs with the reverse flag, range 1 = arcs with the forward flag
  for (const TmpArc& x : a) {
    if (x.rev) g.off[0][x.tail + 1]++;
    if (x.fwd) g.off[1][x.tail + 1]++;
  }
  for (int s = 0; s < 2; ++s) {
    for (int64_t v = 0; v < g.n; ++v) g.off[s][v + 1] += g.off[s][v];
    g.head[s].resize(g.off[s][g.n]);
    g.len[s].resize(g.off[s][g.n]);
  }
  std::vector<uint64_t> pos0(g.off[0].begin(), g.off[0].end() - 1);
  std::vector<uint64_t> pos1(g.off[1].begin(), g.off[1].end() - 1);
  for (const TmpArc& x : a) {
    if (x.rev) { g.head[0][pos0[x.tail]] = x.head; g.len[0][pos0[x.tail]++] = x.len; }
    if (x.fwd) { g.head[1][pos1[x.tail]] = x.head; g.len[1][pos1[x.tail]++] = x.len; }
  }
  a.clear();
  a.shrink_to_fit();
}

static void reset(Graph& g) {
  g.n = 0; g.m = 0; g.tmp.clear();
  for (int s = 0; s < 2; ++s) { g.off[s].clear(); g.head[s].clear(); g.len[s].clear(); }
}

// graph.hpp:168-189
static bool read_dimacs(Graph& g, FILE* f, bool undirected) {
  char buf[512];
  long long u, v, w, m;
  bool inited = false;
  while (fgets(buf, sizeof(buf), f) != NULL) {
    size_t L = strlen(buf);
    if (L == 0 || buf[L - 1] != '\n') {
      if (buf[0] != 'c') return false;
      while (!feof(f) && fgetc(f) != '\n') {}
    } else if (buf[0] == 'p') {
      if (inited) return false;
      inited = true;
      long long nn;
      if (sscanf(buf, "p sp %lld %lld", &nn, &m) != 2) return false;
      g.n = nn;
    } else if (buf[0] == 'a') {
      if (sscanf(buf, "a %lld %lld %lld", &u, &v, &w) != 3) return false;
      if (!add_tmp_arc(g, (Vertex)(u - 1), (Vertex)(v - 1), (Distance)w, undirected)) return false;
    } else if (buf[0] == 'c') {
      continue;
    } else {
      return false;
    }
  }
  finalize(g);
  return true;
}

// graph.hpp:211-256
static bool read_metis(Graph& g, FILE* f, bool undirected) {
  int c, i = 0, fmt = 0, ncon = 0, skip = 0;
  long long elem = 0;
  Vertex v = 0, head = NONE;
  bool newline = true, inelem = false;
  do {
    c = fgetc(f);
    if (newline && c == '%') {
      while (c != '\n' && c != EOF) c = fgetc(f);
    } else if (isdigit(c)) {
      inelem = true;
      elem = elem * 10 + c - '0';
    } else if (isspace(c) || c == EOF) {
      if (inelem) {
        if (v == 0) {
          if (i == 0) g.n = elem;
          else if (i == 1) { /* m ignored */ }
          else if (i == 2) {
            if ((elem % 10 > 1) || ((elem / 10) % 10 > 1) || elem > 111) return false;
            fmt = (int)elem;
            skip = (fmt >= 100) + (fmt % 100 >= 10);
          } else if (i == 3) {
            if (fmt % 100 < 10) return false;
            ncon = (int)elem;
            skip = (fmt >= 100) + ncon;
          } else return false;
        } else if (i >= skip) {
          if (fmt % 10 == 1) {
            if ((i - skip) % 2 == 0) head = (Vertex)(elem - 1);
            else if (!add_tmp_arc(g, v - 1, head, (Distance)elem, undirected)) return false;
          } else if (!add_tmp_arc(g, v - 1, (Vertex)(elem - 1), 1, undirected)) {
            return false;
          }
        }
        ++i; elem = 0; inelem = false;
      }
      if (c == '\n' || (!newline && c == EOF)) {
        if (v > 0 && (i < skip || (fmt % 10 == 1 && (i - skip) % 2 == 1))) return false;
        ++v; i = 0; newline = true;
      }
    } else return false;
  } while (c != EOF);
  finalize(g);
  return true;
}

// graph.hpp:262-282
static bool read_snap(Graph& g, FILE* f, bool undirected) {
  char buf[512];
  long long u, v, m;
  bool inited = false;
  while (fgets(buf, sizeof(buf), f)) {
    size_t L = strlen(buf);
    if (buf[0] == '#') {
      long long nn;
      if (sscanf(buf, "# Nodes: %lld Edges: %lld", &nn, &m) == 2) {
        if (inited) return false;
        inited = true;
        g.n = nn;
      }
      for (int c = buf[L - 1]; c != '\n' && c != EOF; c = fgetc(f)) {}
    } else {
      if (buf[L - 1] != '\n' && !feof(f)) return false;
      if (sscanf(buf, "%lld %lld", &u, &v) != 2) return false;
      if (!add_tmp_arc(g, (Vertex)u, (Vertex)v, 1, undirected)) return false;
    }
  }
  finalize(g);
  return true;
}

// graph.hpp:284-308. The reference counts distinct ids as n and does no bounds
// check; ids outside [0, n) are undefined behaviour there and rejected here.
static bool read_edgelist(Graph& g, FILE* f, bool undirected) {
  char buf[512];
  long long u, v, w;
  std::set<long long> nodes;
  std::vector<std::pair<std::pair<long long, long long>, long long>> e;
  while (fgets(buf, sizeof(buf), f)) {
    size_t L = strlen(buf);
    if (buf[L - 1] != '\n' && !feof(f)) return false;
    if (sscanf(buf, "%lld %lld %lld", &u, &v, &w) != 3) return false;
    e.push_back({{u, v}, w});
    nodes.insert(u);
    nodes.insert(v);
  }
  g.n = (int64_t)nodes.size();
  for (auto& x : e)
    if (!add_tmp_arc(g, (Vertex)x.first.first, (Vertex)x.first.second, (Distance)x.second, undirected))
      return false;
  finalize(g);
  return true;
}

// graph.hpp:327-335
static int read_graph(Graph& g, const char* path, bool undirected) {
  FILE* f = fopen(path, "r");
  if (!f) return -1;
  reset(g);
  if (read_dimacs(g, f, undirected)) { fclose(f); return 1; }
  rewind(f); reset(g);
  if (read_metis(g, f, undirected)) { fclose(f); return 2; }
  rewind(f); reset(g);
  if (read_snap(g, f, undirected)) { fclose(f); return 3; }
  rewind(f); reset(g);
  if (read_edgelist(g, f, undirected)) { fclose(f); return 4; }
  fclose(f); reset(g);
  return -2;
}

// ---------------------------------------------------------------------------
// 4-ary indexed heap, shared_memory/kheap.hpp:33-89.
// ---------------------------------------------------------------------------
struct KHeap {
  size_t size = 0;
  std::vector<Vertex> heap;
  std::vector<size_t> pos;
  std::vector<Distance> key;
  static const size_t INV = (size_t)-1;
  explicit KHeap(size_t n = 0) : heap(n), pos(n, INV), key(n) {}
  void swp(size_t i, size_t j) { std::swap(pos[heap[i]], pos[heap[j]]); std::swap(heap[i], heap[j]); }
  size_t kid(size_t i) const {
    size_t c0 = i * 4 + 1, m = 0;
    for (size_t j = 1; j < 4 && c0 + j < size; ++j)
      if (key[heap[c0 + j]] < key[heap[c0 + m]]) m = j;
    return c0 + m;
  }
  void fixup(size_t i) {
    size_t c;
    while (i * 4 + 1 < size && key[heap[i]] > key[heap[(c = kid(i))]]) { swp(i, c); i = c; }
    while (i > 0 && key[heap[i]] < key[heap[(i - 1) / 4]]) { swp(i, (i - 1) / 4); i = (i - 1) / 4; }
  }
  bool empty() const { return size == 0; }
  void update(Vertex v, Distance k) {
    if (pos[v] == INV) { pos[v] = size++; heap[pos[v]] = v; }
    key[v] = k;
    fixup(pos[v]);
  }
  void extract(Vertex v) {
    if (pos[v] == INV) return;
    if (pos[v] < --size) { size_t p = pos[v]; swp(p, size); fixup(p); }
    pos[v] = INV;
  }
  Vertex pop() { Vertex v = heap[0]; extract(v); return v; }
  void clear() { for (size_t i = 0; i < size; ++i) pos[heap[i]] = INV; size = 0; }
};

// ---------------------------------------------------------------------------
// Labeling store, shared_memory/labeling.hpp:43-54 (rows per vertex and side).
// ---------------------------------------------------------------------------
struct Labeling {
  std::vector<std::vector<Vertex>> hv[2];
  std::vector<std::vector<Distance>> hd[2];
  int64_t n = 0;
  explicit Labeling(int64_t nn = 0) : n(nn) {
    for (int s = 0; s < 2; ++s) { hv[s].assign(nn, {}); hd[s].assign(nn, {}); }
  }
  void clear_rows() {
#pragma omp parallel for
    for (int64_t v = 0; v < n; ++v)
      for (int s = 0; s < 2; ++s) { hv[s][v].clear(); hd[s][v].clear(); }
  }
  // labeling.hpp:117-120
  void add(Vertex u, int side, Vertex h, Distance d) { hv[side][u].push_back(h); hd[side][u].push_back(d); }
  // labeling.hpp:56-65
  Distance query(Vertex u, Vertex v, int f = 1) const {
    Distance r = INF;
    const auto& a = hv[f][u]; const auto& b = hv[!f][v];
    const auto& ad = hd[f][u]; const auto& bd = hd[!f][v];
    for (size_t i = 0, j = 0; i < a.size() && j < b.size();) {
      if (a[i] == b[j]) { r = std::min(r, ad[i++] + bd[j++]); }
      else if (a[i] < b[j]) ++i;
      else ++j;
    }
    return r;
  }
  // labeling.hpp:78-88 (shared) / distributed_memory/labeling.hpp:61-70
  bool clean_cover(Vertex u, Vertex v, int f, Distance d, size_t hub_order) const {
    const auto& a = hv[f][u]; const auto& b = hv[!f][v];
    const auto& ad = hd[f][u]; const auto& bd = hd[!f][v];
    for (size_t i = 0, j = 0; i < a.size() && j < b.size();) {
      if (a[i] >= hub_order || b[j] >= hub_order) return false;
      if (a[i] == b[j]) { if (d >= ad[i++] + bd[j++]) return true; }
      else if (a[i] < b[j]) ++i;
      else ++j;
    }
    return false;
  }
  // labeling.hpp:222-256
  void sort_rows() {
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t v = 0; v < n; ++v)
      for (int s = 0; s < 2; ++s) {
        std::vector<std::pair<Vertex, Distance>> t(hv[s][v].size());
        for (size_t i = 0; i < t.size(); ++i) t[i] = {hv[s][v][i], hd[s][v][i]};
        std::sort(t.begin(), t.end());
        for (size_t i = 0; i < t.size(); ++i) { hv[s][v][i] = t[i].first; hd[s][v][i] = t[i].second; }
      }
  }
  int64_t total() const {
    int64_t t = 0;
    for (int64_t v = 0; v < n; ++v) t += hv[0][v].size() + hv[1][v].size();
    return t;
  }
};

// ---------------------------------------------------------------------------
// Per-thread search scratch: shared_memory/dijkstra.hpp:34-63 (distance, parent,
// dirty list, heap) plus the dense root-label table of shared_memory/akiba.hpp:46-63.
// ---------------------------------------------------------------------------
struct Search {
  const Graph* g;
  KHeap q;
  std::vector<Distance> dist;
  std::vector<Vertex> parent;
  std::vector<uint8_t> dirty_flag;
  std::vector<Vertex> dirty;
  std::vector<Distance> ltmp;
  std::vector<uint8_t> ltmp_flag;
  std::vector<Vertex> ltmp_dirty;
  explicit Search(const Graph& gg)
      : g(&gg), q(gg.n), dist(gg.n, INF), parent(gg.n, NONE), dirty_flag(gg.n, 0),
        ltmp(gg.n, INF), ltmp_flag(gg.n, 0) {}
  // dijkstra.hpp:45-50
  void update(Vertex v, Distance d, Vertex p = NONE) {
    dist[v] = d; parent[v] = p; q.update(v, d);
    if (!dirty_flag[v]) { dirty.push_back(v); dirty_flag[v] = 1; }
  }
  // distributed_memory/dijkstra.hpp:56-78: the PLaNT flag rule. parent==1 means
  // "no more important vertex on the path so far", 0 means "via a more important
  // vertex"; on a tie 0 wins. Returns the change in the number of queued clean
  // vertices.
  int update_wo_prune(Vertex v, Distance d, Vertex p) {
    Distance pd = dist[v];
    Vertex pp = parent[v];
    int r = 0;
    if (d < pd) {
      dist[v] = d; parent[v] = p; q.update(v, d);
      if (!dirty_flag[v]) { dirty.push_back(v); dirty_flag[v] = 1; }
      if (pp != 1 && p == 1) r = 1;
      else if (pp == 1 && p == 0) r = -1;
    } else if (d == pd) {
      if (pp == 1 && p == 0) { parent[v] = p; r = -1; }
    }
    return r;
  }
  // dijkstra.hpp:53-63
  void clear() {
    q.clear();
    for (Vertex v : dirty) { parent[v] = NONE; dist[v] = INF; dirty_flag[v] = 0; }
    dirty.clear();
  }
  // akiba.hpp:50-63
  void update_tmp(Vertex h, Distance d) {
    ltmp[h] = d;
    if (!ltmp_flag[h]) { ltmp_dirty.push_back(h); ltmp_flag[h] = 1; }
  }
  void clear_tmp() {
    for (Vertex h : ltmp_dirty) { ltmp[h] = INF; ltmp_flag[h] = 0; }
    ltmp_dirty.clear();
  }
};

// shared_memory/dijkstra.hpp:83-95 (plain SSSP; positive weights are asserted there)
static void dijkstra(const Graph& g, Vertex src, int forward, Distance* out) {
  Search s(g);
  s.update(src, 0);
  const int r = forward ? 1 : 0;
  while (!s.q.empty()) {
    Vertex u = s.q.pop();
    Distance d = s.dist[u];
    for (uint64_t a = g.off[r][u]; a < g.off[r][u + 1]; ++a) {
      Distance dd = d + g.len[r][a];
      if (!(dd > d && dd < INF)) abort();  // dijkstra.hpp:91
      if (dd < s.dist[g.head[r][a]]) s.update(g.head[r][a], dd, u);
    }
  }
  for (int64_t v = 0; v < g.n; ++v) out[v] = s.dist[v];
}

// Spin lock per (vertex, side), shared_memory/akiba.hpp:73,102.
struct Locks {
  std::vector<std::atomic<uint8_t>> l;
  explicit Locks(size_t k) : l(k) { for (auto& x : l) x.store(0); }
  void lock(size_t i) { uint8_t z; do { z = 0; } while (!l[i].compare_exchange_weak(z, 1, std::memory_order_acquire)); }
  void unlock(size_t i) { l[i].store(0, std::memory_order_release); }
};

// shared_memory/akiba.hpp:65-125: GLL pruned Dijkstra from root order[i].
// forward=true reads out-arcs and fills side 0; forward=false reads in-arcs and
// fills side 1 (akiba.hpp:112,116).
static unsigned iteration_gll(Search& S, size_t i, bool forward, Locks& lck,
                              const std::vector<Vertex>& order, const std::vector<Vertex>& rev,
                              const Labeling& glob, Labeling& loc) {
  const Graph& g = *S.g;
  S.clear();
  S.clear_tmp();
  Vertex v = order[i];
  size_t lk = (forward ? 1 : 0) + (size_t)v * 2;
  unsigned count = 0;
  S.update(v, 0);
  const int fs = forward ? 1 : 0;  // side index used for the root's own labels
  lck.lock(lk);
  {
    size_t sz = loc.hv[fs][v].size();
    for (size_t a = 0; a < sz; ++a) {
      if (loc.hv[fs][v][a] > i) continue;
      S.update_tmp(loc.hv[fs][v][a], loc.hd[fs][v][a]);
    }
  }
  lck.unlock(lk);
  for (size_t a = 0; a < glob.hv[fs][v].size(); ++a) S.update_tmp(glob.hv[fs][v][a], glob.hd[fs][v][a]);
  const int os = forward ? 0 : 1;  // side filled in visited vertices (= !forward)
  const int r = forward ? 1 : 0;   // arc range
  while (!S.q.empty()) {
    Vertex u = S.q.pop();
    Distance d = S.dist[u];
    if (rev[u] < i) continue;
    bool pruned = false;
    for (size_t a = 0; a < glob.hv[os][u].size(); ++a) {
      Vertex h = glob.hv[os][u][a];
      if (h > i) continue;
      if (S.ltmp[h] != INF && S.ltmp[h] + glob.hd[os][u][a] <= d) { pruned = true; break; }
    }
    if (pruned) continue;
    size_t lk2 = (forward ? 0 : 1) + (size_t)u * 2;
    lck.lock(lk2);
    for (size_t a = 0; a < loc.hv[os][u].size(); ++a) {
      Vertex h = loc.hv[os][u][a];
      if (h > i) continue;
      if (S.ltmp[h] != INF && S.ltmp[h] + loc.hd[os][u][a] <= d) { pruned = true; break; }
    }
    if (pruned) { lck.unlock(lk2); continue; }
    loc.add(u, os, (Vertex)i, d);
    lck.unlock(lk2);
    count++;
    for (uint64_t a = g.off[r][u]; a < g.off[r][u + 1]; ++a) {
      Distance dd = d + g.len[r][a];
      if (!(dd > d && dd < INF)) abort();  // akiba.hpp:118
      if (dd < S.dist[g.head[r][a]]) S.update(g.head[r][a], dd);
    }
  }
  return count;
}

struct GllTimes { double lt, st, ct; int64_t supersteps; int64_t roots_done; };

// shared_memory/akiba.hpp:132-217: GLL supersteps.
// max_seconds > 0 stops after the first superstep at which lt + st + ct passes it
// (a bounded sample of the build: the completed rank prefix and its labels).
static void run_gll(const Graph& g, const std::vector<Vertex>& order, const std::vector<Vertex>& rev,
                    Labeling& glob, int P, size_t label_limit, GllTimes* tm, double max_seconds = 0) {
  const int64_t N = (int64_t)order.size();
  Labeling loc(N);
  Locks lck(2 * N + 2);
  std::vector<Search*> ak;
  for (int t = 0; t < P; ++t) ak.push_back(new Search(g));
  double lt = 0, st = 0, ct = 0;
  std::atomic<uint64_t> cnt(0);
  int64_t steps = 0;
  while ((int64_t)cnt.load() < 2 * N) {
    std::atomic<uint64_t> label_sum(0);
    loc.clear_rows();  // akiba.hpp:156 (outside the timers)
    double ls = omp_get_wtime();
#pragma omp parallel for num_threads(P) schedule(static, 1)
    for (int th = 0; th < P; ++th) {
      uint64_t mine = 0;
      while (true) {
        uint64_t rr = cnt.fetch_add(2);
        if ((int64_t)rr >= 2 * N) { label_sum.fetch_add(mine); break; }
        size_t root = rr / 2;
        bool side = (rr % 2) == 1;
        mine += iteration_gll(*ak[th], root, side, lck, order, rev, glob, loc);
        side = (rr % 2) == 0;
        mine += iteration_gll(*ak[th], root, side, lck, order, rev, glob, loc);
        if (mine > (uint64_t)(N / 32)) {  // akiba.hpp:180-183
          if (label_sum.fetch_add(mine) >= label_limit - mine) break;
          mine = 0;
        }
      }
    }
    double le = omp_get_wtime();
    lt += le - ls;
    loc.sort_rows();  // akiba.hpp:189
    double se = omp_get_wtime();
    st += se - le;
    // akiba.hpp:197-208: keep a superstep label unless an in-superstep hub ranked
    // above it covers the same pair at no greater distance.
#pragma omp parallel for num_threads(P) schedule(dynamic, 64)
    for (int64_t v = 0; v < N; ++v)
      for (int s = 0; s < 2; ++s)
        for (size_t k = 0; k < loc.hv[s][v].size(); ++k) {
          size_t ho = loc.hv[s][v][k];
          Vertex hub = order[ho];
          Distance hd = loc.hd[s][v][k];
          if (!loc.clean_cover((Vertex)v, hub, s, hd, ho)) glob.add((Vertex)v, s, (Vertex)ho, hd);
        }
    double ce = omp_get_wtime();
    ct += ce - se;
    steps++;
    if (max_seconds > 0 && lt + st + ct > max_seconds) break;
  }
  for (auto p : ak) delete p;
  if (tm) {
    tm->lt = lt; tm->st = st; tm->ct = ct; tm->supersteps = steps;
    tm->roots_done = std::min<int64_t>(N, (int64_t)cnt.load() / 2);
  }
}

// distributed_memory/akiba.hpp:52-104: PLaNT tree from root order[i]. `common`
// (may be null) holds a complete rank prefix of final labels used only to prune.
// Returns (label_count, tree_size).
static std::pair<unsigned, unsigned> plant(Search& S, size_t i, bool forward, const std::vector<Vertex>& order,
                                           const std::vector<Vertex>& rev, Labeling& out,
                                           const Labeling* common, Locks* lck = nullptr) {
  const Graph& g = *S.g;
  S.clear();
  S.clear_tmp();
  Vertex v = order[i];
  S.update(v, 0, 1);
  unsigned count = 0, tree = 0;
  int64_t queue_left = 1;
  const int fs = forward ? 1 : 0, os = forward ? 0 : 1, r = forward ? 1 : 0;
  if (common)
    for (size_t a = 0; a < common->hv[fs][v].size(); ++a) S.update_tmp(common->hv[fs][v][a], common->hd[fs][v][a]);
  while (!S.q.empty()) {
    Vertex u = S.q.pop();
    tree++;
    bool via = (S.parent[u] == 0) || (rev[u] < i);
    if (!via) queue_left--;
    Distance d = S.dist[u];
    bool pruned = false;
    if (common) {
      for (size_t a = 0; a < common->hv[os][u].size(); ++a) {
        Vertex h = common->hv[os][u][a];
        if (h > i) break;
        if (S.ltmp[h] != INF && S.ltmp[h] + common->hd[os][u][a] <= d) { pruned = true; break; }
      }
    }
    if (!pruned) {
      if (!via) {
        // Labeling::add takes the (vertex, side) lock (distributed_memory/labeling.hpp:73-78)
        if (lck) lck->lock((size_t)u * 2 + os);
        out.add(u, os, (Vertex)i, d);
        if (lck) lck->unlock((size_t)u * 2 + os);
        count++;
      }
      for (uint64_t a = g.off[r][u]; a < g.off[r][u + 1]; ++a) {
        Distance dd = d + g.len[r][a];
        if (!(dd > d && dd < INF)) abort();  // akiba.hpp:88
        Vertex hd = g.head[r][a];
        Vertex p = (via || rev[hd] < i) ? 0 : 1;
        if (dd <= S.dist[hd]) queue_left += S.update_wo_prune(hd, dd, p);
      }
    }
    if (queue_left == 0) break;
  }
  return {count, tree};
}

// distributed_memory/akiba.hpp:107-180: pruned Dijkstra of the post-switch phase,
// pruning against common, global and in-superstep labels.
static unsigned iteration_buffer(Search& S, size_t i, bool forward, const std::vector<Vertex>& order,
                                 const std::vector<Vertex>& rev, const Labeling& glob, Labeling& loc,
                                 const Labeling& common, Locks& lck) {
  const Graph& g = *S.g;
  S.clear();
  S.clear_tmp();
  Vertex v = order[i];
  size_t lk = (forward ? 1 : 0) + (size_t)v * 2;
  S.update(v, 0, 1);
  unsigned count = 0;
  const int fs = forward ? 1 : 0, os = forward ? 0 : 1, r = forward ? 1 : 0;
  lck.lock(lk);
  for (size_t a = 0; a < loc.hv[fs][v].size(); ++a) {
    if (loc.hv[fs][v][a] > i) continue;
    S.update_tmp(loc.hv[fs][v][a], loc.hd[fs][v][a]);
  }
  lck.unlock(lk);
  for (size_t a = 0; a < glob.hv[fs][v].size(); ++a) S.update_tmp(glob.hv[fs][v][a], glob.hd[fs][v][a]);
  for (size_t a = 0; a < common.hv[fs][v].size(); ++a) S.update_tmp(common.hv[fs][v][a], common.hd[fs][v][a]);
  auto covered = [&](const Labeling& L, Vertex u, Distance d) {
    for (size_t a = 0; a < L.hv[os][u].size(); ++a) {
      Vertex h = L.hv[os][u][a];
      if (h > i) continue;
      if (S.ltmp[h] != INF && S.ltmp[h] + L.hd[os][u][a] <= d) return true;
    }
    return false;
  };
  while (!S.q.empty()) {
    Vertex u = S.q.pop();
    if (rev[u] < i) continue;
    Distance d = S.dist[u];
    if (covered(common, u, d) || covered(glob, u, d)) continue;
    size_t lk2 = (forward ? 0 : 1) + (size_t)u * 2;
    lck.lock(lk2);
    if (covered(loc, u, d)) { lck.unlock(lk2); continue; }
    loc.add(u, os, (Vertex)i, d);
    count++;
    lck.unlock(lk2);
    for (uint64_t a = g.off[r][u]; a < g.off[r][u + 1]; ++a) {
      Distance dd = d + g.len[r][a];
      if (!(dd > d && dd < INF)) abort();  // akiba.hpp:169
      if (dd < S.dist[g.head[r][a]]) S.update(g.head[r][a], dd, 0);
    }
  }
  return count;
}

struct HybridStats { double total_s; int64_t switch_root; int64_t supersteps; };

// distributed_memory/akiba.hpp:190-510 at one rank (world_size 1): every MPI call is
// an identity, the label exchange degenerates into moving in-superstep labels into
// the common (rank prefix within budget) or global labeling, and the pruned-phase
// filter is computeLocalMask (utils.hpp:249-276) with its AND-reduction a no-op.
// cb and ps are floats, as run's parameters are (akiba.hpp:190): the budget is a
// float product (akiba.hpp:204) and the ratio test compares with ps widened to
// double (akiba.hpp:206, 301).
static void run_hybrid(const Graph& g, const std::vector<Vertex>& order, const std::vector<Vertex>& rev,
                       Labeling& out, int P, float cb, float ps, HybridStats* hs) {
  const int64_t N = (int64_t)order.size();
  Labeling loc(N), common(N), glob(N);
  Locks lck(2 * N + 2);
  std::vector<Search*> ak;
  for (int t = 0; t < P; ++t) ak.push_back(new Search(g));
  uint64_t common_budget = std::min<uint64_t>((uint64_t)(unsigned)(N * 2 * 16), (uint64_t)(unsigned)(1000000 * cb));
  bool do_plant = true, switch_compute = true, comm_done = false;
  int64_t sync_thres = P, prev_sync = 0;
  std::atomic<int64_t> cnt(0);
  int64_t switch_root = -1, steps = 0;
  double t0 = omp_get_wtime();
  while (cnt.load() < N) {
    std::vector<int64_t> per_hub(sync_thres - prev_sync, 0);
    std::atomic<int> phase_flag(0);
    // Threads claim roots of [prev_sync, sync_thres) in rank order. In the PLaNT
    // phase a root whose tree/label ratio passes ps raises the switch flag; the
    // roots already claimed are finished and the superstep ends there
    // (akiba.hpp:290-376: the goto build_trees pass completes claimed trees).
#pragma omp parallel for num_threads(P)
    for (int th = 0; th < P; ++th) {
      for (;;) {
        if (do_plant && switch_compute && phase_flag.load()) break;  // akiba.hpp:323-327
        int64_t lc = cnt.fetch_add(1);
        if (lc >= sync_thres || lc >= N) break;
        if (do_plant) {
          auto s1 = plant(*ak[th], lc, true, order, rev, loc, &common, &lck);
          auto s2 = plant(*ak[th], lc, false, order, rev, loc, &common, &lck);
          unsigned ls = s1.first + s2.first, ts = s1.second + s2.second;
          per_hub[lc - prev_sync] += ls;
          double ratio = (double)ts / (double)ls;  // akiba.hpp:297
          if (switch_compute && ratio > (double)ps) phase_flag.store(1);
        } else {
          per_hub[lc - prev_sync] += iteration_buffer(*ak[th], lc, true, order, rev, glob, loc, common, lck);
          per_hub[lc - prev_sync] += iteration_buffer(*ak[th], lc, false, order, rev, glob, loc, common, lck);
        }
      }
    }
    // claims at or past the range end were not processed
    const int64_t done = std::min<int64_t>(cnt.load(), std::min<int64_t>(sync_thres, N));
    cnt.store(done);
    if (do_plant && switch_compute && phase_flag.load()) {
      sync_thres = done;
      switch_root = done;
      switch_compute = false;
    }
    per_hub.resize(sync_thres - prev_sync);
    loc.sort_rows();
    // common-label cut: the longest prefix of this superstep's hubs whose
    // cumulative label count fits the remaining budget (comm_funcs.hpp:45-75)
    int64_t lastCV = sync_thres;
    {
      std::vector<uint64_t> cum(per_hub.size());
      uint64_t acc = 0;
      for (size_t k = 0; k < per_hub.size(); ++k) { acc += per_hub[k]; cum[k] = acc; }
      if (cum.empty() || cum[0] > common_budget || common_budget == 0) {
        common_budget = 0;
        lastCV = prev_sync;
      } else {
        for (size_t k = 0; k + 1 < cum.size(); ++k)
          if (cum[k] <= common_budget && cum[k + 1] > common_budget) { lastCV = prev_sync + k + 1; break; }
        if (cum.back() >= common_budget) common_budget = 0; else common_budget -= cum.back();
      }
    }
    if (do_plant) {
      // PLaNT labels are canonical: no filter (akiba.hpp:399-438)
      for (int64_t v = 0; v < N; ++v)
        for (int s = 0; s < 2; ++s)
          for (size_t k = 0; k < loc.hv[s][v].size(); ++k) {
            Vertex h = loc.hv[s][v][k];
            if ((int64_t)h < lastCV) common.add((Vertex)v, s, h, loc.hd[s][v][k]);
            else glob.add((Vertex)v, s, h, loc.hd[s][v][k]);
          }
      if (lastCV < sync_thres) comm_done = true;
    } else {
      // computeLocalMask (utils.hpp:249-276): drop a candidate covered through
      // global (HP) or in-superstep (LP) labels of hubs ranked above it.
      std::vector<std::vector<uint8_t>> keep[2];
      for (int s = 0; s < 2; ++s) keep[s].resize(N);
#pragma omp parallel for num_threads(P) schedule(dynamic, 64)
      for (int64_t v = 0; v < N; ++v)
        for (int s = 0; s < 2; ++s) {
          keep[s][v].assign(loc.hv[s][v].size(), 1);
          for (size_t k = 0; k < loc.hv[s][v].size(); ++k) {
            size_t ho = loc.hv[s][v][k];
            Vertex hub = order[ho];
            Distance d = loc.hd[s][v][k];
            if (glob.clean_cover((Vertex)v, hub, s, d, ho) || loc.clean_cover((Vertex)v, hub, s, d, ho))
              keep[s][v][k] = 0;
          }
        }
      for (int64_t v = 0; v < N; ++v)
        for (int s = 0; s < 2; ++s)
          for (size_t k = 0; k < loc.hv[s][v].size(); ++k) {
            if (!keep[s][v][k]) continue;
            Vertex h = loc.hv[s][v][k];
            if ((int64_t)h < lastCV) common.add((Vertex)v, s, h, loc.hd[s][v][k]);
            else glob.add((Vertex)v, s, h, loc.hd[s][v][k]);
          }
      comm_done = comm_done || lastCV < sync_thres;
    }
    glob.sort_rows();
    common.sort_rows();
    if (!comm_done) comm_done = (common_budget == 0);
    prev_sync = sync_thres;
    do_plant = do_plant && switch_compute;
    sync_thres = std::min<int64_t>(sync_thres * 2, N);
    loc.clear_rows();
    steps++;
  }
  double t1 = omp_get_wtime();
  // remove_common at one rank keeps every common label (utils.hpp:292-304)
  for (int64_t v = 0; v < N; ++v)
    for (int s = 0; s < 2; ++s) {
      for (size_t k = 0; k < common.hv[s][v].size(); ++k) out.add((Vertex)v, s, common.hv[s][v][k], common.hd[s][v][k]);
      for (size_t k = 0; k < glob.hv[s][v].size(); ++k) out.add((Vertex)v, s, glob.hv[s][v][k], glob.hd[s][v][k]);
    }
  out.sort_rows();
  for (auto p : ak) delete p;
  if (hs) { hs->total_s = t1 - t0; hs->switch_root = switch_root; hs->supersteps = steps; }
}

}  // namespace orc

using namespace orc;

// ---------------------------------------------------------------------------
// C ABI for the tests (ctypes). Handles are opaque pointers.
// ---------------------------------------------------------------------------
struct OrcLabels {
  int64_t n;
  std::vector<uint64_t> off[2];
  std::vector<uint32_t> hub[2];
  std::vector<int32_t> dist[2];
};

static OrcLabels* to_flat(const Labeling& L) {
  OrcLabels* o = new OrcLabels();
  o->n = L.n;
  for (int s = 0; s < 2; ++s) {
    o->off[s].assign(L.n + 1, 0);
    for (int64_t v = 0; v < L.n; ++v) o->off[s][v + 1] = o->off[s][v] + L.hv[s][v].size();
    o->hub[s].resize(o->off[s][L.n]);
    o->dist[s].resize(o->off[s][L.n]);
    for (int64_t v = 0; v < L.n; ++v) {
      std::copy(L.hv[s][v].begin(), L.hv[s][v].end(), o->hub[s].begin() + o->off[s][v]);
      std::copy(L.hd[s][v].begin(), L.hd[s][v].end(), o->dist[s].begin() + o->off[s][v]);
    }
  }
  return o;
}

static Labeling from_flat(const OrcLabels* o) {
  Labeling L(o->n);
  for (int s = 0; s < 2; ++s)
    for (int64_t v = 0; v < o->n; ++v)
      for (uint64_t k = o->off[s][v]; k < o->off[s][v + 1]; ++k) L.add((Vertex)v, s, o->hub[s][k], o->dist[s][k]);
  return L;
}

static bool make_order(const Graph& g, const uint32_t* order, int64_t n, std::vector<Vertex>& ord,
                       std::vector<Vertex>& rev) {
  if (n != g.n) return false;
  ord.assign(order, order + n);
  rev.assign(n, NONE);
  for (int64_t i = 0; i < n; ++i) {
    if (ord[i] >= (Vertex)n || rev[ord[i]] != NONE) return false;
    rev[ord[i]] = (Vertex)i;
  }
  return true;
}

extern "C" {

// Build a graph from directed arcs as the DIMACS reader would after `a u v w` lines
// (0-based ids); undirected=1 adds each arc in both directions (graph.hpp:156-162).
void* orc_graph_from_arcs(int64_t n, int64_t m, const uint32_t* src, const uint32_t* dst, const int32_t* w,
                          int undirected) {
  Graph* g = new Graph();
  g->n = n;
  for (int64_t k = 0; k < m; ++k)
    if (!add_tmp_arc(*g, src[k], dst[k], w[k], undirected != 0)) { delete g; return nullptr; }
  finalize(*g);
  return g;
}

// Returns the detected format (1 DIMACS, 2 METIS, 3 SNAP, 4 edgelist) or <0.
int orc_graph_read(const char* path, int undirected, void** out) {
  Graph* g = new Graph();
  int r = read_graph(*g, path, undirected != 0);
  if (r < 0) { delete g; *out = nullptr; return r; }
  *out = g;
  return r;
}

void orc_graph_free(void* g) { delete (Graph*)g; }
int64_t orc_graph_n(void* g) { return ((Graph*)g)->n; }
int64_t orc_graph_m(void* g) { return ((Graph*)g)->m; }
int64_t orc_graph_arcs(void* g, int range) { return (int64_t)((Graph*)g)->head[range].size(); }

// Copy one range as CSR (range 1 = out-arcs, 0 = in-arcs), rows sorted by head.
void orc_graph_csr(void* gp, int range, uint64_t* off, uint32_t* head, int32_t* len) {
  Graph* g = (Graph*)gp;
  for (int64_t v = 0; v <= g->n; ++v) off[v] = g->off[range][v];
  for (int64_t v = 0; v < g->n; ++v) {
    std::vector<std::pair<Vertex, Distance>> row;
    for (uint64_t a = g->off[range][v]; a < g->off[range][v + 1]; ++a) row.push_back({g->head[range][a], g->len[range][a]});
    std::sort(row.begin(), row.end());
    for (size_t k = 0; k < row.size(); ++k) { head[off[v] + k] = row[k].first; len[off[v] + k] = row[k].second; }
  }
}

void orc_dijkstra(void* g, uint32_t src, int forward, int32_t* out) { dijkstra(*(Graph*)g, src, forward, out); }

// Multi-source Dijkstra (OpenMP over sources): out is [k][n].
void orc_dijkstra_multi(void* gp, const uint32_t* srcs, int64_t k, int forward, int32_t* out, int threads) {
  Graph* g = (Graph*)gp;
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1)
  for (int64_t i = 0; i < k; ++i) dijkstra(*g, srcs[i], forward, out + i * g->n);
}

// Order file (ordering.hpp:52-67): returns N, fills up to cap ids; -1 if unreadable.
int64_t orc_order_read(const char* path, uint32_t* out, int64_t cap) {
  std::ifstream f(path);
  if (!f.is_open()) return -1;
  size_t s = 0;
  f >> s;
  std::vector<uint32_t> o(s);
  for (size_t i = 0; i < s; ++i) f >> o[i];
  for (int64_t i = 0; i < (int64_t)s && i < cap; ++i) out[i] = o[i];
  return (int64_t)s;
}

// GLL build (shared_memory/akiba.hpp:132-217) with P threads and sf (hhl.cpp:25,65).
// Bounded GLL sample: supersteps until max_seconds of lt+st+ct; stats5 =
// lt, st, ct, supersteps, roots claimed (a root is claimed by the superstep that
// builds both its trees, so the labels cover ranks [0, roots)).
void* orc_gll_bounded(void* gp, const uint32_t* order, int64_t n, int P, int sf, double max_seconds, double* stats5) {
  Graph* g = (Graph*)gp;
  std::vector<Vertex> ord, rev;
  if (!make_order(*g, order, n, ord, rev)) return nullptr;
  Labeling L(n);
  GllTimes tm;
  run_gll(*g, ord, rev, L, P, (size_t)sf * (size_t)n, &tm, max_seconds);
  if (stats5) {
    stats5[0] = tm.lt; stats5[1] = tm.st; stats5[2] = tm.ct; stats5[3] = (double)tm.supersteps;
    stats5[4] = (double)tm.roots_done;
  }
  return to_flat(L);
}

void* orc_gll(void* gp, const uint32_t* order, int64_t n, int P, int sf, double* times4) {
  Graph* g = (Graph*)gp;
  std::vector<Vertex> ord, rev;
  if (!make_order(*g, order, n, ord, rev)) return nullptr;
  Labeling L(n);
  GllTimes tm;
  run_gll(*g, ord, rev, L, P, (size_t)sf * (size_t)n, &tm);
  if (times4) { times4[0] = tm.lt; times4[1] = tm.st; times4[2] = tm.ct; times4[3] = (double)tm.supersteps; }
  return to_flat(L);
}

// Hybrid PLaNT -> pruned at one rank (distributed_memory/akiba.hpp:190-510).
void* orc_hybrid(void* gp, const uint32_t* order, int64_t n, int P, double cb, double ps, double* stats3) {
  Graph* g = (Graph*)gp;
  std::vector<Vertex> ord, rev;
  if (!make_order(*g, order, n, ord, rev)) return nullptr;
  Labeling L(n);
  HybridStats hs;
  run_hybrid(*g, ord, rev, L, P, (float)cb, (float)ps, &hs);
  if (stats3) { stats3[0] = hs.total_s; stats3[1] = (double)hs.switch_root; stats3[2] = (double)hs.supersteps; }
  return to_flat(L);
}

// PLaNT trees for ranks [r0, r1), both directions, no pruning (akiba.hpp:52-104).
// Labels are emitted into a fresh labeling, rows sorted; tree_sizes[2*(i-r0)+f]
// receives the tree size of direction f (f=1 forward).
void* orc_plant_range(void* gp, const uint32_t* order, int64_t n, int64_t r0, int64_t r1, void* common_labels,
                      uint32_t* tree_sizes, uint32_t* label_counts, int P) {
  Graph* g = (Graph*)gp;
  std::vector<Vertex> ord, rev;
  if (!make_order(*g, order, n, ord, rev)) return nullptr;
  Labeling common_l;
  const Labeling* cp = nullptr;
  if (common_labels) { common_l = from_flat((OrcLabels*)common_labels); cp = &common_l; }
  std::vector<Labeling> outs(P, Labeling(n));
  std::vector<Search*> ak;
  for (int t = 0; t < P; ++t) ak.push_back(new Search(*g));
#pragma omp parallel for num_threads(P) schedule(dynamic, 1)
  for (int64_t i = r0; i < r1; ++i) {
    int t = omp_get_thread_num();
    for (int f = 1; f >= 0; --f) {
      auto s = plant(*ak[t], (size_t)i, f == 1, ord, rev, outs[t], cp);
      if (tree_sizes) tree_sizes[2 * (i - r0) + f] = s.second;
      if (label_counts) label_counts[2 * (i - r0) + f] = s.first;
    }
  }
  Labeling L(n);
  for (int t = 0; t < P; ++t)
    for (int64_t v = 0; v < n; ++v)
      for (int s = 0; s < 2; ++s)
        for (size_t k = 0; k < outs[t].hv[s][v].size(); ++k) L.add((Vertex)v, s, outs[t].hv[s][v][k], outs[t].hd[s][v][k]);
  L.sort_rows();
  for (auto p : ak) delete p;
  return to_flat(L);
}

// Brute-force CHL from the definition: (h, d) is in L_0(v) iff h is the smallest
// rank on the union of shortest paths order[h] -> v, and in L_1(v) for paths
// v -> order[h]. Intended for n up to a few thousand.
void* orc_brute_chl(void* gp, const uint32_t* order, int64_t n, int P) {
  Graph* g = (Graph*)gp;
  std::vector<Vertex> ord, rev;
  if (!make_order(*g, order, n, ord, rev)) return nullptr;
  std::vector<Distance> D((size_t)n * n);  // D[s*n+t] = d(s -> t)
#pragma omp parallel for num_threads(P) schedule(dynamic, 1)
  for (int64_t s = 0; s < n; ++s) dijkstra(*g, (Vertex)s, 1, &D[(size_t)s * n]);
  // best[s*n+t] = min rank over vertices on shortest s->t paths
  std::vector<Vertex> best((size_t)n * n, NONE);
#pragma omp parallel for num_threads(P) schedule(dynamic, 1)
  for (int64_t s = 0; s < n; ++s)
    for (int64_t t = 0; t < n; ++t) {
      Distance dst = D[(size_t)s * n + t];
      if (dst == INF) continue;
      Vertex b = NONE;
      for (int64_t w = 0; w < n; ++w) {
        Distance a = D[(size_t)s * n + w], c = D[(size_t)w * n + t];
        if (a == INF || c == INF) continue;
        if ((int64_t)a + c == dst && rev[w] < b) b = rev[w];
      }
      best[(size_t)s * n + t] = b;
    }
  Labeling L(n);
  for (int64_t v = 0; v < n; ++v)
    for (int64_t h = 0; h < n; ++h) {
      Vertex x = ord[h];
      if (best[(size_t)x * n + v] == (Vertex)h) L.add((Vertex)v, 0, (Vertex)h, D[(size_t)x * n + v]);
      if (best[(size_t)v * n + x] == (Vertex)h) L.add((Vertex)v, 1, (Vertex)h, D[(size_t)v * n + x]);
    }
  return to_flat(L);
}

void orc_labels_free(void* l) { delete (OrcLabels*)l; }
int64_t orc_labels_n(void* l) { return ((OrcLabels*)l)->n; }
int64_t orc_labels_total(void* l, int side) { return (int64_t)((OrcLabels*)l)->hub[side].size(); }
void orc_labels_copy(void* lp, int side, uint64_t* off, uint32_t* hub, int32_t* dist) {
  OrcLabels* l = (OrcLabels*)lp;
  std::copy(l->off[side].begin(), l->off[side].end(), off);
  std::copy(l->hub[side].begin(), l->hub[side].end(), hub);
  std::copy(l->dist[side].begin(), l->dist[side].end(), dist);
}
// Build labels from flat arrays (for checking product output with the oracle's query).
void* orc_labels_from(int64_t n, const uint64_t* off0, const uint32_t* hub0, const int32_t* d0, const uint64_t* off1,
                      const uint32_t* hub1, const int32_t* d1) {
  OrcLabels* o = new OrcLabels();
  o->n = n;
  const uint64_t* offs[2] = {off0, off1};
  const uint32_t* hubs[2] = {hub0, hub1};
  const int32_t* ds[2] = {d0, d1};
  for (int s = 0; s < 2; ++s) {
    o->off[s].assign(offs[s], offs[s] + n + 1);
    o->hub[s].assign(hubs[s], hubs[s] + offs[s][n]);
    o->dist[s].assign(ds[s], ds[s] + offs[s][n]);
  }
  return o;
}

// labeling.hpp:167-183 label file writer.
int orc_labels_write(void* lp, const char* path) {
  OrcLabels* l = (OrcLabels*)lp;
  FILE* f = fopen(path, "w");
  if (!f) return -1;
  fprintf(f, "%lld\n", (long long)l->n);
  for (int64_t v = 0; v < l->n; ++v)
    for (int s = 0; s < 2; ++s) {
      fprintf(f, "%llu", (unsigned long long)(l->off[s][v + 1] - l->off[s][v]));
      for (uint64_t k = l->off[s][v]; k < l->off[s][v + 1]; ++k) fprintf(f, " %u %d", l->hub[s][k], l->dist[s][k]);
      fputc('\n', f);
    }
  return fclose(f) == 0 ? 0 : -1;
}

// labeling.hpp:186-208 label file reader.
void* orc_labels_read(const char* path) {
  std::ifstream f(path);
  if (!f.is_open()) return nullptr;
  int64_t n;
  f >> n;
  if (!f) return nullptr;
  Labeling L(n);
  for (int64_t v = 0; v < n; ++v)
    for (int s = 0; s < 2; ++s) {
      size_t k;
      f >> k;
      L.hv[s][v].resize(k);
      L.hd[s][v].resize(k);
      for (size_t i = 0; i < k; ++i) f >> L.hv[s][v][i] >> L.hd[s][v][i];
    }
  if (f.fail()) return nullptr;
  return to_flat(L);
}

// Labeling::query (labeling.hpp:56-65) batched like batchLocalQuery
// (distributed_memory/query_funcs.hpp:23-30).
void orc_query_batch(void* lp, const uint32_t* pairs, int64_t q, int32_t* out, int P) {
  OrcLabels* l = (OrcLabels*)lp;
#pragma omp parallel for num_threads(P) schedule(dynamic, 256)
  for (int64_t k = 0; k < q; ++k) {
    Vertex u = pairs[2 * k], v = pairs[2 * k + 1];
    Distance r = INF;
    uint64_t i = l->off[1][u], ie = l->off[1][u + 1], j = l->off[0][v], je = l->off[0][v + 1];
    while (i < ie && j < je) {
      uint32_t a = l->hub[1][i], b = l->hub[0][j];
      if (a == b) { r = std::min(r, l->dist[1][i++] + l->dist[0][j++]); }
      else if (a < b) ++i;
      else ++j;
    }
    out[k] = r;
  }
}

// clean_cover (labeling.hpp:78-88) on a flat labeling: 1 if L_f(u) and L_!f(v)
// share a hub below hub_order whose distances sum to at most d.
int orc_clean_cover(void* lp, uint32_t u, uint32_t v, int f, int32_t d, uint64_t hub_order) {
  OrcLabels* l = (OrcLabels*)lp;
  uint64_t i = l->off[f][u], ie = l->off[f][u + 1], j = l->off[!f][v], je = l->off[!f][v + 1];
  while (i < ie && j < je) {
    uint32_t a = l->hub[f][i], b = l->hub[!f][j];
    if (a >= hub_order || b >= hub_order) return 0;
    if (a == b) { if (d >= l->dist[f][i++] + l->dist[!f][j++]) return 1; }
    else if (a < b) ++i;
    else ++j;
  }
  return 0;
}

}  // extern "C"
