// Streams a label file in the text format of shared_memory/labeling.hpp:167-183
// (line 1: n; then per vertex v a side-0 line and a side-1 line, each
// "k h1 d1 ... hk dk") and prints, per side, the entry count and the digest that
// chl_labels_checksum computes on the device: the sum mod 2^64 over every vertex
// v and entry (h, d) of v's row of
//   splitmix64_finalizer(((uint64_t)h << 32 | (uint32_t)d) ^ v * 0x9E3779B97F4A7C15).
// Test infrastructure: it turns the reference's own label file into a fixture
// small enough to commit. Usage: label_digest FILE  (or - for stdin).
// Output: one JSON object {"n":..,"count":[c0,c1],"digest":["0x..","0x.."]}.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace {

uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

struct Reader {
  FILE* f;
  char buf[1 << 20];
  size_t len = 0, pos = 0;
  int get() {
    if (pos == len) {
      len = fread(buf, 1, sizeof(buf), f);
      pos = 0;
      if (len == 0) return -1;
    }
    return (unsigned char)buf[pos++];
  }
  // next unsigned decimal token; false at end of input
  bool next(uint64_t& x) {
    int c = get();
    while (c == ' ' || c == '\n' || c == '\r' || c == '\t') c = get();
    if (c < 0) return false;
    if (c < '0' || c > '9') {
      fprintf(stderr, "label_digest: unexpected byte %d\n", c);
      exit(2);
    }
    x = 0;
    while (c >= '0' && c <= '9') {
      x = x * 10 + (uint64_t)(c - '0');
      c = get();
    }
    return true;
  }
};

}  // namespace

int main(int argc, char** argv) {
  if (argc != 2) {
    fprintf(stderr, "usage: label_digest FILE\n");
    return 1;
  }
  static Reader r;
  r.f = strcmp(argv[1], "-") == 0 ? stdin : fopen(argv[1], "rb");
  if (!r.f) {
    perror(argv[1]);
    return 1;
  }
  uint64_t n = 0;
  if (!r.next(n)) {
    fprintf(stderr, "label_digest: empty file\n");
    return 2;
  }
  uint64_t count[2] = {0, 0}, digest[2] = {0, 0};
  for (uint64_t v = 0; v < n; ++v) {
    const uint64_t key = v * 0x9E3779B97F4A7C15ull;
    for (int side = 0; side < 2; ++side) {
      uint64_t k = 0;
      if (!r.next(k)) {
        fprintf(stderr, "label_digest: file ends at vertex %llu\n", (unsigned long long)v);
        return 2;
      }
      for (uint64_t i = 0; i < k; ++i) {
        uint64_t h = 0, d = 0;
        if (!r.next(h) || !r.next(d)) {
          fprintf(stderr, "label_digest: row of vertex %llu cut short\n", (unsigned long long)v);
          return 2;
        }
        digest[side] += mix64(((h << 32) | (uint32_t)d) ^ key);
      }
      count[side] += k;
    }
  }
  uint64_t extra = 0;
  if (r.next(extra)) {
    fprintf(stderr, "label_digest: data after the last row\n");
    return 2;
  }
  printf("{\"n\": %llu, \"count\": [%llu, %llu], \"digest\": [\"0x%016llx\", \"0x%016llx\"]}\n",
         (unsigned long long)n, (unsigned long long)count[0], (unsigned long long)count[1],
         (unsigned long long)digest[0], (unsigned long long)digest[1]);
  return 0;
}
