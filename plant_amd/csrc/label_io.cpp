// Label files at scale (labeling.hpp:167-208 format): a parallel text writer
// and reader, and a binary sidecar of the device store.
//
// Text format: line 1 = n; then, for each vertex v and side s = 0, 1, a line
// "k h1 d1 ... hk dk". The writer formats chunks of vertices on all host threads
// and writes them in order; its bytes equal the reference writer's (ofstream
// with " " separators and std::endl). The reader follows the reference reader's
// token semantics (whitespace-separated tokens, any layout, only whitespace
// after the last label, labeling.hpp:185-204): a file laid out one row per line,
// as every writer lays it out, is parsed in parallel in two passes (row sizes,
// then entries straight into the final arrays); any other layout falls back to a
// sequential token parser with the same result.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <atomic>
#include <cstdio>
#include <cstring>

#include "../../include/chl.h"
#include "host.hpp"

namespace chl {
namespace {

inline char* put_u64(char* p, uint64_t x) {
  char t[24];
  int k = 0;
  do {
    t[k++] = (char)('0' + x % 10);
    x /= 10;
  } while (x);
  while (k) *p++ = t[--k];
  return p;
}

inline char* put_i32(char* p, int32_t x) {
  if (x < 0) {
    *p++ = '-';
    return put_u64(p, (uint64_t)(-(int64_t)x));
  }
  return put_u64(p, (uint64_t)x);
}

inline bool is_ws(char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\r' || c == '\v' || c == '\f'; }
inline bool is_blank(char c) { return c == ' ' || c == '\t' || c == '\r' || c == '\v' || c == '\f'; }

// A decimal integer token (optional sign) at p within [p, e); advances p.
inline bool get_int(const char*& p, const char* e, long long& out) {
  bool neg = false;
  if (p < e && (*p == '-' || *p == '+')) neg = *p++ == '-';
  if (p >= e || *p < '0' || *p > '9') return false;
  unsigned long long x = 0;
  int digits = 0;
  while (p < e && *p >= '0' && *p <= '9') {
    x = x * 10 + (unsigned)(*p++ - '0');
    if (++digits > 19) return false;
  }
  out = neg ? -(long long)x : (long long)x;
  return true;
}

struct Mapped {
  const char* p = nullptr;
  size_t size = 0;
  int fd = -1;
  ~Mapped() {
    if (p && size) munmap((void*)p, size);
    if (fd >= 0) close(fd);
  }
  bool open_file(const char* path) {
    fd = ::open(path, O_RDONLY);
    if (fd < 0) return false;
    struct stat st;
    if (fstat(fd, &st) != 0) return false;
    size = (size_t)st.st_size;
    if (size == 0) return true;
    void* m = mmap(nullptr, size, PROT_READ, MAP_PRIVATE, fd, 0);
    if (m == MAP_FAILED) {
      size = 0;
      return false;
    }
    madvise(m, size, MADV_SEQUENTIAL);
    p = (const char*)m;
    return true;
  }
};

// Reference token semantics, one thread: n, then per vertex and side k and k
// (hub, dist) pairs; only whitespace may follow.
int read_tokens(const char* p, const char* e, Labels& L) {
  long long n;
  while (p < e && is_ws(*p)) ++p;
  // every vertex has two rows of at least one token and a separator
  if (!get_int(p, e, n) || n < 0 || n > (e - p) / 4 + 1) return CHL_E_FORMAT;
  L.n = n;
  for (int s = 0; s < 2; ++s) {
    L.off[s].assign(n + 1, 0);
    L.ent[s].clear();
  }
  for (long long v = 0; v < n; ++v)
    for (int s = 0; s < 2; ++s) {
      long long k;
      while (p < e && is_ws(*p)) ++p;
      if (!get_int(p, e, k) || k < 0) return CHL_E_FORMAT;
      for (long long i = 0; i < k; ++i) {
        long long h, d;
        while (p < e && is_ws(*p)) ++p;
        if (!get_int(p, e, h)) return CHL_E_FORMAT;
        while (p < e && is_ws(*p)) ++p;
        if (!get_int(p, e, d)) return CHL_E_FORMAT;
        L.ent[s].push_back(((uint64_t)(uint32_t)h << 32) | (uint32_t)(int32_t)d);
      }
      L.off[s][v + 1] = L.ent[s].size();
    }
  while (p < e && is_ws(*p)) ++p;
  return p == e ? CHL_OK : CHL_E_FORMAT;
}

// One row per line: returns false (and leaves L undefined) when the layout is
// not that, so the caller falls back to read_tokens.
bool read_lines(const char* body, const char* e, long long n, Labels& L) {
  const int64_t rows = 2 * n;
  const int T = std::max(1, host_threads());
  // ranges of whole lines
  std::vector<const char*> cut(T + 1, e);
  cut[0] = body;
  for (int t = 1; t < T; ++t) {
    const char* q = body + (size_t)(e - body) * t / T;
    if (q < cut[t - 1]) q = cut[t - 1];
    const char* nl = (const char*)memchr(q, '\n', (size_t)(e - q));
    cut[t] = nl ? nl + 1 : e;
  }
  std::vector<int64_t> first(T + 1, 0);  // first line index of each range
  parallel_for(T, T, 1, [&](int64_t t) {
    int64_t c = 0;
    for (const char* q = cut[t]; q < cut[t + 1];) {
      const char* nl = (const char*)memchr(q, '\n', (size_t)(cut[t + 1] - q));
      ++c;
      q = nl ? nl + 1 : cut[t + 1];
    }
    first[t + 1] = c;
  });
  for (int t = 0; t < T; ++t) first[t + 1] += first[t];
  if (first[T] < rows) return false;
  // pass 1: row sizes, and the layout check
  std::vector<uint64_t> cnt(rows, 0);
  std::atomic<bool> ok(true);
  parallel_for(T, T, 1, [&](int64_t t) {
    int64_t line = first[t];
    for (const char* q = cut[t]; q < cut[t + 1] && ok.load(std::memory_order_relaxed); ++line) {
      const char* nl = (const char*)memchr(q, '\n', (size_t)(cut[t + 1] - q));
      const char* le = nl ? nl : cut[t + 1];
      const char* p = q;
      q = nl ? nl + 1 : cut[t + 1];
      if (line >= rows) {  // after the last row only whitespace
        while (p < le && is_blank(*p)) ++p;
        if (p != le) ok = false;
        continue;
      }
      while (p < le && is_blank(*p)) ++p;
      long long k;
      if (!get_int(p, le, k) || k < 0) {
        ok = false;
        break;
      }
      long long toks = 0, x;
      for (;;) {
        while (p < le && is_blank(*p)) ++p;
        if (p == le) break;
        if (!get_int(p, le, x) || (p < le && !is_blank(*p))) {
          ok = false;
          break;
        }
        ++toks;
      }
      if (toks != 2 * k) ok = false;
      cnt[line] = (uint64_t)k;
    }
  });
  if (!ok) return false;
  for (int s = 0; s < 2; ++s) {
    L.off[s].assign(n + 1, 0);
    for (long long v = 0; v < n; ++v) L.off[s][v + 1] = L.off[s][v] + cnt[2 * v + s];
    L.ent[s].resize(L.off[s][n]);
  }
  // pass 2: entries in place
  parallel_for(T, T, 1, [&](int64_t t) {
    int64_t line = first[t];
    for (const char* q = cut[t]; q < cut[t + 1] && line < rows; ++line) {
      const char* nl = (const char*)memchr(q, '\n', (size_t)(cut[t + 1] - q));
      const char* le = nl ? nl : cut[t + 1];
      const char* p = q;
      q = nl ? nl + 1 : cut[t + 1];
      const int s = (int)(line & 1);
      uint64_t* out = L.ent[s].data() + L.off[s][line >> 1];
      long long k = 0, h = 0, d = 0;  // the first pass checked every token
      while (p < le && is_blank(*p)) ++p;
      get_int(p, le, k);
      for (long long i = 0; i < k; ++i) {
        while (p < le && is_blank(*p)) ++p;
        get_int(p, le, h);
        while (p < le && is_blank(*p)) ++p;
        get_int(p, le, d);
        out[i] = ((uint64_t)(uint32_t)h << 32) | (uint32_t)(int32_t)d;
      }
    }
  });
  return true;
}

}  // namespace

int labels_write(const Labels& L, const char* path) {
  FILE* f = fopen(path, "wb");
  if (!f) return CHL_E_IO;
  const int64_t n = L.n;
  fprintf(f, "%lld\n", (long long)n);
  // chunks of consecutive vertices holding about 4M entries each
  std::vector<int64_t> cut{0};
  uint64_t acc = 0;
  for (int64_t v = 0; v < n; ++v) {
    acc += (L.off[0][v + 1] - L.off[0][v]) + (L.off[1][v + 1] - L.off[1][v]) + 1;
    if (acc >= ((uint64_t)1 << 22)) {
      cut.push_back(v + 1);
      acc = 0;
    }
  }
  if (cut.back() != n) cut.push_back(n);
  const int T = std::max(1, host_threads());
  std::vector<std::vector<char>> buf(T);
  std::vector<size_t> len(T);
  bool good = true;
  for (size_t c0 = 0; c0 + 1 < cut.size() && good; c0 += T) {
    const size_t c1 = std::min(cut.size() - 1, c0 + T);
    parallel_for((int64_t)(c1 - c0), T, 1, [&](int64_t i) {
      const int64_t a = cut[c0 + i], b = cut[c0 + i + 1];
      // at most 21 bytes per row count and newline, 23 per entry
      size_t bound = 0;
      for (int s = 0; s < 2; ++s) bound += 22 * (size_t)(b - a) + 23 * (size_t)(L.off[s][b] - L.off[s][a]);
      std::vector<char>& out = buf[i];
      if (out.size() < bound) out.resize(bound);
      char* p = out.data();
      for (int64_t v = a; v < b; ++v)
        for (int s = 0; s < 2; ++s) {
          p = put_u64(p, L.off[s][v + 1] - L.off[s][v]);
          for (uint64_t k = L.off[s][v]; k < L.off[s][v + 1]; ++k) {
            const uint64_t e = L.ent[s][k];
            *p++ = ' ';
            p = put_u64(p, (uint32_t)(e >> 32));
            *p++ = ' ';
            p = put_i32(p, (int32_t)(uint32_t)e);
          }
          *p++ = '\n';
        }
      len[i] = (size_t)(p - out.data());
    });
    for (size_t i = 0; i < c1 - c0 && good; ++i) good = fwrite(buf[i].data(), 1, len[i], f) == len[i];
  }
  if (fclose(f) != 0) good = false;
  return good ? CHL_OK : CHL_E_IO;
}

int labels_read(Labels& L, const char* path) {
  Mapped m;
  if (!m.open_file(path)) return CHL_E_IO;
  const char* p = m.p;
  const char* e = m.p + m.size;
  if (!p) return CHL_E_FORMAT;  // empty file
  // the header line holding n alone selects the line-parallel path
  const char* q = p;
  while (q < e && is_blank(*q)) ++q;
  long long n;
  if (get_int(q, e, n) && n >= 0 && n <= (long long)(m.size / 4) + 1) {
    while (q < e && is_blank(*q)) ++q;
    if (q < e && *q == '\n') {
      L.n = n;
      if (read_lines(q + 1, e, n, L)) return CHL_OK;
    }
  }
  return read_tokens(p, e, L);
}

}  // namespace chl
