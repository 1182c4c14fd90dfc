// Graph ingest (DIMACS, METIS, SNAP, edgelist with autodetection), CSR build
// and order files for the CHL engine. Files of any of the four formats in its
// plain layout are parsed on all host threads; every other file takes the
// sequential readers.
//
// Semantics follow the reference readers (shared_memory/graph.hpp:164-335):
// the same formats are tried in the same order, ids are 1-based for DIMACS/METIS
// and 0-based for SNAP/edgelist, out-of-range ids make a reader fail, and parallel
// arcs keep their minimum length (graph.hpp:107-114). Where the reference would
// abort later on a non-positive length (assert at shared_memory/dijkstra.hpp:91)
// the engine rejects the graph at load time with CHL_E_WEIGHT.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cctype>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <set>

#include "../../include/chl.h"
#include "host.hpp"

namespace chl {

int host_threads() {
  const char* e = getenv("CHL_HOST_THREADS");
  if (e && atoi(e) > 0) return atoi(e);
  unsigned h = std::thread::hardware_concurrency();
  if (h == 0) h = 1;
  return h > 16 ? 16 : (int)h;
}

// One range: rows of (head, min length) sorted by head. The scatter into rows is
// sequential (a row's order before sorting does not matter); sorting, the
// per-row deduplication and the final copy run on all host threads.
static void build_range(int64_t n, std::vector<Arc3>& a, bool by_tail, std::vector<uint64_t>& off,
                        std::vector<uint32_t>& head, std::vector<int32_t>& len) {
  // key (row, col, w) with row = tail for out-arcs, head for in-arcs
  std::vector<uint64_t> cnt(n + 1, 0);
  for (const Arc3& x : a) cnt[(by_tail ? x.u : x.v) + 1]++;
  for (int64_t v = 0; v < n; ++v) cnt[v + 1] += cnt[v];
  std::vector<uint64_t> pos(cnt.begin(), cnt.end() - 1);
  std::vector<std::pair<uint32_t, int32_t>> tmp(a.size());
  for (const Arc3& x : a) {
    uint32_t r = by_tail ? x.u : x.v, c = by_tail ? x.v : x.u;
    tmp[pos[r]++] = {c, x.w};
  }
  const int T = host_threads();
  // sort each row; count its distinct heads (the first of equal heads has the
  // minimum length)
  std::vector<uint64_t> keep(n + 1, 0);
  parallel_for(n, T, 4096, [&](int64_t v) {
    auto b = tmp.begin() + cnt[v], e = tmp.begin() + cnt[v + 1];
    std::sort(b, e);
    uint64_t k = 0;
    for (auto it = b; it != e; ++it)
      if (it == b || it->first != (it - 1)->first) ++k;
    keep[v + 1] = k;
  });
  off.assign(n + 1, 0);
  for (int64_t v = 0; v < n; ++v) off[v + 1] = off[v] + keep[v + 1];
  head.resize(off[n]);
  len.resize(off[n]);
  parallel_for(n, T, 4096, [&](int64_t v) {
    auto b = tmp.begin() + cnt[v], e = tmp.begin() + cnt[v + 1];
    uint64_t o = off[v];
    for (auto it = b; it != e; ++it)
      if (it == b || it->first != (it - 1)->first) {
        head[o] = it->first;
        len[o] = it->second;
        ++o;
      }
  });
}

int graph_from_arcs(Graph& g, int64_t n, std::vector<Arc3>& arcs, bool undirected) {
  if (n <= 0 || n >= (int64_t)1 << 30) return CHL_E_RANGE;  // Arc::head is 30 bits (graph.hpp:51)
  g.n = n;
  g.m_input = 0;
  for (const Arc3& x : arcs) {
    if ((int64_t)x.u >= n || (int64_t)x.v >= n) return CHL_E_RANGE;
    if (x.w <= 0) return CHL_E_WEIGHT;
  }
  g.m_input = (int64_t)arcs.size() * (undirected ? 2 : 1);
  if (undirected) {
    size_t m = arcs.size();
    arcs.reserve(2 * m);
    for (size_t k = 0; k < m; ++k) arcs.push_back({arcs[k].v, arcs[k].u, arcs[k].w});
  }
  build_range(n, arcs, true, g.off[1], g.head[1], g.len[1]);
  build_range(n, arcs, false, g.off[0], g.head[0], g.len[0]);
  // symmetric iff the in-range equals the out-range row by row
  g.symmetric = g.off[0] == g.off[1] && g.head[0] == g.head[1] && g.len[0] == g.len[1];
  return CHL_OK;
}

// ---- readers ---------------------------------------------------------------

namespace {
inline bool blank_char(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\v' || c == '\f'; }
// A decimal token as sscanf's %lld takes it (leading blanks, optional sign);
// false for anything this fast path leaves to sscanf (more than 18 digits).
inline bool scan_ll(const char*& p, const char* e, long long& out) {
  while (p < e && blank_char(*p)) ++p;
  bool neg = false;
  if (p < e && (*p == '-' || *p == '+')) neg = *p++ == '-';
  if (p >= e || *p < '0' || *p > '9') return false;
  long long x = 0;
  int d = 0;
  while (p < e && *p >= '0' && *p <= '9') {
    x = x * 10 + (*p++ - '0');
    if (++d > 18) return false;
  }
  out = neg ? -x : x;
  return true;
}

// [q, e) split into T ranges of whole lines (e ends in a newline).
std::vector<const char*> line_cuts(const char* q, const char* e, int T) {
  std::vector<const char*> cut(T + 1, e);
  cut[0] = q;
  for (int t = 1; t < T; ++t) {
    const char* c = q + (size_t)(e - q) * t / T;
    if (c < cut[t - 1]) c = cut[t - 1];
    const char* nl = c < e ? (const char*)memchr(c, '\n', (size_t)(e - c)) : nullptr;
    cut[t] = nl ? nl + 1 : e;
  }
  return cut;
}

// Per-thread arc lists joined in range order.
void join_parts(std::vector<std::vector<Arc3>>& part, std::vector<Arc3>& arcs) {
  size_t total = 0;
  for (auto& v : part) total += v.size();
  arcs.clear();
  arcs.reserve(total);
  for (auto& v : part) {
    arcs.insert(arcs.end(), v.begin(), v.end());
    std::vector<Arc3>().swap(v);
  }
}
}  // namespace

// DIMACS on all host threads, for files in the plain layout: every line ends in
// a newline, no NUL byte, comment lines first and one 'p sp n m' line before any
// arc, then 'a u v w' and comment lines, non-comment lines under 512 bytes.
// Returns 1 with the same arcs, error and n as parse_dimacs, or 0 when the file
// is anything else: parse_dimacs then decides (accept or reject) as the
// reference reader does.
static int parse_dimacs_fast(const char* data, size_t size, int64_t& n, std::vector<Arc3>& arcs, int& err) {
  if (size == 0 || data[size - 1] != '\n' || memchr(data, 0, size)) return 0;
  const char* e = data + size;
  const char* q = data;
  // header: comment lines, then the one 'p' line
  for (;;) {
    if (q >= e) return 0;
    const char* nl = (const char*)memchr(q, '\n', (size_t)(e - q));
    if (*q == 'c') {
      q = nl + 1;
      continue;
    }
    if (*q != 'p' || nl - q + 1 > 511) return 0;
    char buf[512];
    memcpy(buf, q, (size_t)(nl - q + 1));
    buf[nl - q + 1] = 0;
    long long nn, mm;
    if (sscanf(buf, "p sp %lld %lld", &nn, &mm) != 2) return 0;
    n = nn;
    q = nl + 1;
    break;
  }
  const int T = host_threads();
  const std::vector<const char*> cut = line_cuts(q, e, T);
  std::vector<std::vector<Arc3>> part(T);
  std::vector<int> perr(T, CHL_OK);
  std::atomic<bool> plain(true);
  const long long nmax = n;
  parallel_for(T, T, 1, [&](int64_t t) {
    std::vector<Arc3>& out = part[t];
    for (const char* p = cut[t]; p < cut[t + 1] && plain.load(std::memory_order_relaxed);) {
      const char* nl = (const char*)memchr(p, '\n', (size_t)(cut[t + 1] - p));
      const char* le = nl;
      if (*p == 'c') {
        p = nl + 1;
        continue;
      }
      long long u, v, w;
      const char* r = p + 1;
      if (*p != 'a' || nl - p + 1 > 511 || !scan_ll(r, le, u) || !scan_ll(r, le, v) || !scan_ll(r, le, w) ||
          u < 1 || v < 1 || u > nmax || v > nmax) {
        plain = false;
        break;
      }
      if (w <= 0 || w > 2147483647LL) perr[t] = CHL_E_WEIGHT;
      out.push_back({(uint32_t)(u - 1), (uint32_t)(v - 1), (int32_t)w});
      p = nl + 1;
    }
  });
  if (!plain) return 0;
  join_parts(part, arcs);
  for (int t = 0; t < T; ++t)
    if (perr[t] != CHL_OK) err = perr[t];
  return 1;
}

// A character that ends a METIS token on a line (isspace without the newline).
inline bool metis_space(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\v' || c == '\f'; }

// The tokens of one METIS line [p, le): runs of digits separated by spaces,
// each passed to f(index, value) (f returns false to stop). Returns the token
// count, or -1 for any other character, a run over 18 digits or a stop.
template <class F>
int metis_tokens(const char* p, const char* le, F f) {
  int i = 0;
  while (p < le) {
    if (metis_space(*p)) {
      ++p;
      continue;
    }
    if (*p < '0' || *p > '9') return -1;
    long long x = 0;
    int d = 0;
    while (p < le && *p >= '0' && *p <= '9') {
      x = x * 10 + (*p++ - '0');
      if (++d > 18) return -1;
    }
    if (!f(i, x)) return -1;
    ++i;
  }
  return i;
}

// A line whose first non-space character is '%' (a METIS comment, graph.hpp:220).
inline bool metis_comment(const char* p, const char* le) {
  while (p < le && metis_space(*p)) ++p;
  return p < le && *p == '%';
}

// METIS on all host threads, for files in the plain layout: the first byte is
// not 'c', 'p' or 'a' (so the DIMACS reader rejects the file), the file ends in
// a newline, the first non-comment line is a valid header with n > 0, and every
// later line is a comment or a vertex line that parse_metis accepts. Vertex
// lines are numbered in a first pass, so ranges of lines parse independently.
// Returns 1 with the arcs, n and error that parse_metis gives, or 0 when the
// file is anything else: the sequential readers then decide.
static int parse_metis_fast(const char* data, size_t size, int64_t& n, std::vector<Arc3>& arcs, int& err) {
  if (size == 0 || data[size - 1] != '\n' || data[0] == 'c' || data[0] == 'p' || data[0] == 'a') return 0;
  const char* e = data + size;
  const char* q = data;
  while (q < e) {  // comment lines before the header
    const char* nl = (const char*)memchr(q, '\n', (size_t)(e - q));
    if (!metis_comment(q, nl)) break;
    q = nl + 1;
  }
  if (q >= e) return 0;
  const char* hl = (const char*)memchr(q, '\n', (size_t)(e - q));
  long long hv[4] = {0, 0, 0, 0};
  const int hc = metis_tokens(q, hl, [&](int i, long long x) {
    if (i >= 4) return false;
    hv[i] = x;
    return true;
  });
  if (hc < 1) return 0;
  const long long nn = hv[0];
  long long fmt = 0, skip = 0;
  if (hc >= 3) {
    fmt = hv[2];
    if (fmt % 10 > 1 || (fmt / 10) % 10 > 1 || fmt > 111) return 0;
    skip = (fmt >= 100) + (fmt % 100 >= 10);
  }
  if (hc == 4) {
    if (fmt % 100 < 10 || hv[3] > 1000) return 0;
    skip = (fmt >= 100) + hv[3];
  }
  if (nn <= 0 || nn > 0x7FFFFFFFll) return 0;
  const bool weighted = fmt % 10 == 1;
  q = hl + 1;
  const int T = host_threads();
  const std::vector<const char*> cut = line_cuts(q, e, T);
  // pass 1: vertex lines per range
  std::vector<int64_t> first(T + 1, 0);
  parallel_for(T, T, 1, [&](int64_t t) {
    int64_t c = 0;
    for (const char* p = cut[t]; p < cut[t + 1];) {
      const char* nl = (const char*)memchr(p, '\n', (size_t)(cut[t + 1] - p));
      if (!metis_comment(p, nl)) ++c;
      p = nl + 1;
    }
    first[t + 1] = c;
  });
  for (int t = 0; t < T; ++t) first[t + 1] += first[t];
  // pass 2: arcs; v is the 1-based vertex of a line, as in parse_metis
  std::vector<std::vector<Arc3>> part(T);
  std::vector<int> perr(T, CHL_OK);
  std::atomic<bool> plain(true);
  parallel_for(T, T, 1, [&](int64_t t) {
    std::vector<Arc3>& out = part[t];
    int64_t v = first[t] + 1;
    for (const char* p = cut[t]; p < cut[t + 1] && plain.load(std::memory_order_relaxed);) {
      const char* nl = (const char*)memchr(p, '\n', (size_t)(cut[t + 1] - p));
      if (metis_comment(p, nl)) {
        p = nl + 1;
        continue;
      }
      long long headv = -1;
      const int c = metis_tokens(p, nl, [&](int i, long long x) {
        if (i < skip) return true;
        if (x > 0x7FFFFFFFll) return false;  // the sequential reader decides
        if (weighted) {
          if ((i - skip) % 2 == 0) {
            headv = x - 1;
            return true;
          }
          if (headv < 0 || headv >= nn || v - 1 >= nn) return false;
          if (x <= 0 || x > 2147483647LL) perr[t] = CHL_E_WEIGHT;
          out.push_back({(uint32_t)(v - 1), (uint32_t)headv, (int32_t)x});
          return true;
        }
        if (x < 1 || x > nn || v - 1 >= nn) return false;
        out.push_back({(uint32_t)(v - 1), (uint32_t)(x - 1), 1});
        return true;
      });
      if (c < 0 || c < skip || (weighted && (c - skip) % 2 == 1)) {
        plain = false;
        break;
      }
      ++v;
      p = nl + 1;
    }
  });
  if (!plain) return 0;
  n = nn;
  join_parts(part, arcs);
  for (int t = 0; t < T; ++t)
    if (perr[t] != CHL_OK) err = perr[t];
  return 1;
}

// SNAP on all host threads, for files in the plain layout: the first byte is
// '#' (so the DIMACS and METIS readers reject the file), every line ends in a
// newline, no NUL byte, '#' lines only before the first data line, exactly one
// of them reading "# Nodes: n Edges: m", every line under 512 bytes, and each
// data line two ids in [0, n) (anything may follow them, as sscanf allows).
// Returns 1 with the arcs and n that parse_snap gives, or 0 when the file is
// anything else: the sequential readers then decide.
static int parse_snap_fast(const char* data, size_t size, int64_t& n, std::vector<Arc3>& arcs) {
  if (size == 0 || data[0] != '#' || data[size - 1] != '\n' || memchr(data, 0, size)) return 0;
  const char* e = data + size;
  const char* q = data;
  bool inited = false;
  while (q < e && *q == '#') {
    const char* nl = (const char*)memchr(q, '\n', (size_t)(e - q));
    if (nl - q + 1 > 511) return 0;
    char buf[512];
    memcpy(buf, q, (size_t)(nl - q + 1));
    buf[nl - q + 1] = 0;
    long long nn, mm;
    if (sscanf(buf, "# Nodes: %lld Edges: %lld", &nn, &mm) == 2) {
      if (inited) return 0;
      inited = true;
      n = nn;
    }
    q = nl + 1;
  }
  if (!inited) return 0;
  const int T = host_threads();
  const std::vector<const char*> cut = line_cuts(q, e, T);
  std::vector<std::vector<Arc3>> part(T);
  std::atomic<bool> plain(true);
  const long long nmax = n;
  parallel_for(T, T, 1, [&](int64_t t) {
    std::vector<Arc3>& out = part[t];
    for (const char* p = cut[t]; p < cut[t + 1] && plain.load(std::memory_order_relaxed);) {
      const char* nl = (const char*)memchr(p, '\n', (size_t)(cut[t + 1] - p));
      long long u, v;
      const char* r = p;
      if (*p == '#' || nl - p + 1 > 511 || !scan_ll(r, nl, u) || !scan_ll(r, nl, v) || u < 0 || v < 0 || u >= nmax ||
          v >= nmax) {
        plain = false;
        break;
      }
      out.push_back({(uint32_t)u, (uint32_t)v, 1});
      p = nl + 1;
    }
  });
  if (!plain) return 0;
  join_parts(part, arcs);
  return 1;
}

// Edge list on all host threads, for files in the plain layout: every line ends
// in a newline, no NUL byte, every line under 512 bytes holding three numbers
// u v w with 0 <= u, v < 2^31 (anything may follow them, as sscanf allows), the
// ids exactly 0 .. max, and a first line that the METIS header check rejects
// (a sign, or a third number that is no METIS fmt code, graph.hpp:224-236); the
// DIMACS and SNAP readers reject any file whose first line starts with a number.
// Returns 1 with the arcs, n and error that parse_edgelist gives, or 0 when the
// file is anything else: the sequential readers then decide.
static int parse_edgelist_fast(const char* data, size_t size, int64_t& n, std::vector<Arc3>& arcs, int& err) {
  if (size == 0 || data[size - 1] != '\n' || memchr(data, 0, size)) return 0;
  const char* e = data + size;
  {
    const char* nl = (const char*)memchr(data, '\n', size);
    const char* r = data;
    long long u, v, w;
    if (!scan_ll(r, nl, u) || !scan_ll(r, nl, v) || !scan_ll(r, nl, w)) return 0;
    const bool sign = memchr(data, '-', (size_t)(r - data)) || memchr(data, '+', (size_t)(r - data));
    const bool fmt_bad = w < 0 || w % 10 > 1 || (w / 10) % 10 > 1 || w > 111;
    if (!sign && !fmt_bad) return 0;
  }
  const int T = host_threads();
  const std::vector<const char*> cut = line_cuts(data, e, T);
  std::vector<std::vector<Arc3>> part(T);
  std::vector<int> perr(T, CHL_OK);
  std::vector<long long> pmax(T, -1);
  std::atomic<bool> plain(true);
  parallel_for(T, T, 1, [&](int64_t t) {
    std::vector<Arc3>& out = part[t];
    long long mx = -1;
    for (const char* p = cut[t]; p < cut[t + 1] && plain.load(std::memory_order_relaxed);) {
      const char* nl = (const char*)memchr(p, '\n', (size_t)(cut[t + 1] - p));
      long long u, v, w;
      const char* r = p;
      if (nl - p + 1 > 511 || !scan_ll(r, nl, u) || !scan_ll(r, nl, v) || !scan_ll(r, nl, w) || u < 0 || v < 0 ||
          u > 0x7FFFFFFFll || v > 0x7FFFFFFFll) {
        plain = false;
        break;
      }
      if (w <= 0 || w > 2147483647LL) perr[t] = CHL_E_WEIGHT;
      mx = std::max(mx, std::max(u, v));
      out.push_back({(uint32_t)u, (uint32_t)v, (int32_t)w});
      p = nl + 1;
    }
    pmax[t] = mx;
  });
  if (!plain) return 0;
  long long mx = -1;
  size_t total = 0;
  for (int t = 0; t < T; ++t) {
    mx = std::max(mx, pmax[t]);
    total += part[t].size();
  }
  // n counts the distinct ids (graph.hpp:297-304); the reader accepts only ids
  // below n, that is ids 0 .. max all present
  if (mx < 0 || (uint64_t)mx >= 2 * (uint64_t)total) return 0;
  std::vector<uint8_t> seen((size_t)mx + 1, 0);
  parallel_for(T, T, 1, [&](int64_t t) {
    for (const Arc3& a : part[t]) {
      __atomic_store_n(&seen[a.u], (uint8_t)1, __ATOMIC_RELAXED);
      __atomic_store_n(&seen[a.v], (uint8_t)1, __ATOMIC_RELAXED);
    }
  });
  std::atomic<bool> full(true);
  parallel_for((int64_t)seen.size(), T, 1 << 16, [&](int64_t i) {
    if (!seen[i]) full.store(false, std::memory_order_relaxed);
  });
  if (!full) return 0;
  n = mx + 1;
  join_parts(part, arcs);
  for (int t = 0; t < T; ++t)
    if (perr[t] != CHL_OK) err = perr[t];
  return 1;
}

static bool parse_dimacs(FILE* f, int64_t& n, std::vector<Arc3>& arcs, int& err) {
  char buf[512];
  bool inited = false;
  n = 0;
  while (fgets(buf, sizeof(buf), f)) {
    size_t L = strlen(buf);
    if (L == 0 || buf[L - 1] != '\n') {  // graph.hpp:175-177: only comments may be long or unterminated
      if (buf[0] != 'c') return false;
      int c;
      while ((c = fgetc(f)) != EOF && c != '\n') {}
      continue;
    }
    if (buf[0] == 'p') {
      long long nn, mm;
      if (inited || sscanf(buf, "p sp %lld %lld", &nn, &mm) != 2) return false;
      inited = true;
      n = nn;
      if (mm > 0) arcs.reserve((size_t)std::min<long long>(mm, 1 << 24));  // a hint, not trusted
    } else if (buf[0] == 'a') {
      long long u, v, w;
      if (sscanf(buf, "a %lld %lld %lld", &u, &v, &w) != 3) return false;
      if (u < 1 || v < 1 || u > n || v > n) return false;
      if (w <= 0 || w > 2147483647LL) err = CHL_E_WEIGHT;
      arcs.push_back({(uint32_t)(u - 1), (uint32_t)(v - 1), (int32_t)w});
    } else if (buf[0] != 'c') {
      return false;
    }
  }
  return true;
}

// graph.hpp:211-257. The reference's newline flag is never cleared, so a '%'
// anywhere starts a comment that runs to the end of the line, and that newline
// ends no line: the next line continues the current one (a token cut by the
// comment continues too). At the end of the file an unterminated line's tokens
// count, but its line-end check is not made. Ids are taken as the reference's
// 32-bit Vertex parameters take them (elem - 1, modulo 2^32).
static bool parse_metis(FILE* f, int64_t& n, std::vector<Arc3>& arcs, int& err) {
  int c, i = 0, fmt = 0, skip = 0;
  long long elem = 0;
  uint32_t v = 0, headv = 0xFFFFFFFFu;
  bool inelem = false;
  n = 0;
  do {
    c = fgetc(f);
    if (c == '%') {
      while (c != '\n' && c != EOF) c = fgetc(f);
    } else if (isdigit(c)) {
      inelem = true;
      elem = (long long)((unsigned long long)elem * 10 + (unsigned)(c - '0'));
    } else if (c == EOF || isspace(c)) {
      if (inelem) {
        if (v == 0) {
          if (i == 0) n = elem;
          else if (i == 2) {
            if (elem % 10 > 1 || (elem / 10) % 10 > 1 || elem > 111) return false;
            fmt = (int)elem;
            skip = (fmt >= 100) + (fmt % 100 >= 10);
          } else if (i == 3) {
            if (fmt % 100 < 10) return false;
            skip = (fmt >= 100) + (int)elem;
          } else if (i > 3) return false;
        } else if (i >= skip) {
          if (fmt % 10 == 1) {
            if ((i - skip) % 2 == 0) headv = (uint32_t)(elem - 1);
            else {
              if ((int64_t)(v - 1) >= n || (int64_t)headv >= n) return false;
              if (elem <= 0 || elem > 2147483647LL) err = CHL_E_WEIGHT;
              arcs.push_back({v - 1, headv, (int32_t)elem});
            }
          } else {
            const uint32_t h = (uint32_t)(elem - 1);
            if ((int64_t)(v - 1) >= n || (int64_t)h >= n) return false;
            arcs.push_back({v - 1, h, 1});
          }
        }
        ++i;
        elem = 0;
        inelem = false;
      }
      if (c == '\n') {
        if (v > 0 && (i < skip || (fmt % 10 == 1 && (i - skip) % 2 == 1))) return false;
        ++v;
        i = 0;
      }
    } else {
      return false;
    }
  } while (c != EOF);
  return n > 0;
}

static bool parse_snap(FILE* f, int64_t& n, std::vector<Arc3>& arcs) {
  char buf[512];
  bool inited = false;
  n = 0;
  while (fgets(buf, sizeof(buf), f)) {
    size_t L = strlen(buf);
    if (buf[0] == '#') {
      long long nn, mm;
      if (sscanf(buf, "# Nodes: %lld Edges: %lld", &nn, &mm) == 2) {
        if (inited) return false;
        inited = true;
        n = nn;
      }
      if (buf[L - 1] != '\n') {
        int c;
        while ((c = fgetc(f)) != EOF && c != '\n') {}
      }
    } else {
      if (buf[L - 1] != '\n' && !feof(f)) return false;
      long long u, v;
      if (sscanf(buf, "%lld %lld", &u, &v) != 2) return false;
      if (u < 0 || v < 0 || u >= n || v >= n) return false;
      arcs.push_back({(uint32_t)u, (uint32_t)v, 1});
    }
  }
  return true;
}

static bool parse_edgelist(FILE* f, int64_t& n, std::vector<Arc3>& arcs, int& err) {
  char buf[512];
  std::set<long long> ids;
  while (fgets(buf, sizeof(buf), f)) {
    size_t L = strlen(buf);
    if (buf[L - 1] != '\n' && !feof(f)) return false;
    long long u, v, w;
    if (sscanf(buf, "%lld %lld %lld", &u, &v, &w) != 3) return false;
    if (u < 0 || v < 0) return false;
    if (w <= 0 || w > 2147483647LL) err = CHL_E_WEIGHT;
    ids.insert(u);
    ids.insert(v);
    arcs.push_back({(uint32_t)u, (uint32_t)v, (int32_t)w});
  }
  n = (int64_t)ids.size();  // graph.hpp:297-304 counts distinct ids
  for (const Arc3& a : arcs)
    if ((int64_t)a.u >= n || (int64_t)a.v >= n) return false;  // ids must be contiguous
  return true;
}

int graph_read(Graph& g, const char* path, bool undirected) {
  {
    // DIMACS, METIS, SNAP or an edge list in the plain layout, on all host threads
    int fd = open(path, O_RDONLY);
    if (fd < 0) return CHL_E_IO;
    struct stat st;
    if (fstat(fd, &st) == 0 && st.st_size > 0) {
      void* m = mmap(nullptr, (size_t)st.st_size, PROT_READ, MAP_PRIVATE, fd, 0);
      if (m != MAP_FAILED) {
        int64_t n = 0;
        int err = CHL_OK;
        std::vector<Arc3> arcs;
        const char* data = (const char*)m;
        const size_t size = (size_t)st.st_size;
        int fmt = parse_dimacs_fast(data, size, n, arcs, err) ? 1 : 0;
        if (!fmt && parse_metis_fast(data, size, n, arcs, err)) fmt = 2;
        if (!fmt && parse_snap_fast(data, size, n, arcs)) fmt = 3;
        if (!fmt && parse_edgelist_fast(data, size, n, arcs, err)) fmt = 4;
        munmap(m, size);
        if (fmt) {
          close(fd);
          if (err != CHL_OK) return err;
          int r = graph_from_arcs(g, n, arcs, undirected);
          return r == CHL_OK ? fmt : r;
        }
      }
    }
    close(fd);
  }
  FILE* f = fopen(path, "r");
  if (!f) return CHL_E_IO;
  for (int fmt = 1; fmt <= 4; ++fmt) {
    rewind(f);
    int64_t n = 0;
    int err = CHL_OK;
    std::vector<Arc3> arcs;
    bool ok = false;
    switch (fmt) {
      case 1: ok = parse_dimacs(f, n, arcs, err); break;
      case 2: ok = parse_metis(f, n, arcs, err); break;
      case 3: ok = parse_snap(f, n, arcs); break;
      case 4: ok = parse_edgelist(f, n, arcs, err); break;
    }
    if (!ok) continue;
    fclose(f);
    if (err != CHL_OK) return err;
    int r = graph_from_arcs(g, n, arcs, undirected);
    return r == CHL_OK ? fmt : r;
  }
  fclose(f);
  return CHL_E_FORMAT;
}

int graph_write_dimacs(const Graph& g, const char* path) {
  FILE* f = fopen(path, "w");
  if (!f) return CHL_E_IO;
  std::vector<char> big(1 << 22);
  setvbuf(f, big.data(), _IOFBF, big.size());
  fprintf(f, "p sp %lld %lld\n", (long long)g.n, (long long)g.head[1].size());
  for (int64_t v = 0; v < g.n; ++v)
    for (uint64_t a = g.off[1][v]; a < g.off[1][v + 1]; ++a)
      fprintf(f, "a %lld %u %d\n", (long long)v + 1, g.head[1][a] + 1, g.len[1][a]);
  int r = fclose(f);
  return r == 0 ? CHL_OK : CHL_E_IO;
}

// ---- orders (ordering.hpp:37-67) -------------------------------------------

int order_read(const char* path, std::vector<uint32_t>& order) {
  FILE* f = fopen(path, "r");
  if (!f) return CHL_E_IO;
  long long s;
  if (fscanf(f, "%lld", &s) != 1 || s < 0) { fclose(f); return CHL_E_FORMAT; }
  {
    // every id takes at least two bytes: a larger count cannot be in the file
    struct stat st;
    if (fstat(fileno(f), &st) != 0 || s > (long long)st.st_size / 2 + 1) { fclose(f); return CHL_E_FORMAT; }
  }
  order.resize((size_t)s);
  for (long long i = 0; i < s; ++i) {
    long long x;
    if (fscanf(f, "%lld", &x) != 1 || x < 0) { fclose(f); return CHL_E_FORMAT; }
    order[i] = (uint32_t)x;
  }
  fclose(f);
  return CHL_OK;
}

int order_write(const char* path, const std::vector<uint32_t>& order) {
  FILE* f = fopen(path, "w");
  if (!f) return CHL_E_IO;
  fprintf(f, "%zu\n", order.size());
  for (uint32_t v : order) fprintf(f, "%u\n", v);
  return fclose(f) == 0 ? CHL_OK : CHL_E_IO;
}

// Label files (labeling.hpp:167-208): label_io.cpp.

}  // namespace chl
