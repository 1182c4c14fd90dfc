// Test harness only: the query kernel (plant_amd/csrc/query.hip) run on the
// host through hip/hip_runtime.h of this directory, every workgroup of the
// launch in turn. Input and output are little-endian binary files:
//   in:  u32 n, ks, q; then for side 0 and side 1, for each of the ks segments:
//        u32 fmt (5 or 8), u32 h0, u64 count, u64 off[n + 1] (rows by vertex
//        id), then the entries: fmt 5: u32 key[count], u8 lo[count];
//        fmt 8: u64 ent[count]; then u32 pairs[2 q]
//   out: i32 out[q]
#include "../../plant_amd/csrc/query.hip"

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

using namespace chl;

template <class T>
static bool get(FILE* f, T* p, size_t count) {
  return count == 0 || fread(p, sizeof(T), count, f) == count;
}

// 256-byte aligned storage, as hipMalloc gives (the SEG_F5 group loads read
// 16 bytes at a time)
static void* aligned(size_t bytes) {
  void* p = nullptr;
  if (posix_memalign(&p, 256, bytes ? bytes : 1)) return nullptr;
  memset(p, 0xA5, bytes);  // padding holds junk, as device memory may
  return p;
}

int main(int argc, char** argv) {
  if (argc != 3) return 2;
  FILE* f = fopen(argv[1], "rb");
  if (!f) return 2;
  uint32_t hdr[3];
  if (!get(f, hdr, 3)) return 2;
  const uint32_t n = hdr[0], ks = hdr[1];
  const uint64_t q = hdr[2];
  if (ks == 0 || ks > kQuerySeg) return 2;
  QueryView qv;
  memset(&qv, 0, sizeof(qv));
  qv.nseg = ks;
  bool all5 = true, all8 = true;
  std::vector<void*> held;
  for (int side = 0; side < 2; ++side) {
    QuerySide& S = side ? qv.s1 : qv.s0;
    std::vector<std::vector<uint64_t>> offs(ks);
    for (uint32_t s = 0; s < ks; ++s) {
      uint32_t fh[2];
      uint64_t count = 0;
      if (!get(f, fh, 2) || !get(f, &count, 1)) return 2;
      S.fmt[s] = fh[0];
      S.h0[s] = fh[1];
      offs[s].resize(n + 1);
      if (!get(f, offs[s].data(), n + 1)) return 2;
      if (fh[0] == SEG_F5) {
        uint32_t* key = (uint32_t*)aligned(4 * (count + kF5Pad));
        uint8_t* lo = (uint8_t*)aligned(count + kF5Pad);
        if (!get(f, key, count) || !get(f, lo, count)) return 2;
        S.key[s] = key;
        S.lo[s] = lo;
        held.push_back(key);
        held.push_back(lo);
        all8 = false;
      } else {
        uint64_t* ent = (uint64_t*)aligned(8 * count);
        if (!get(f, ent, count)) return 2;
        S.ent[s] = ent;
        held.push_back(ent);
        all5 = false;
      }
    }
    // row ranges, vertex-major: (begin, end) of every segment's row
    uint64_t* rng = (uint64_t*)aligned(16 * (size_t)n * ks);
    for (uint32_t v = 0; v < n; ++v)
      for (uint32_t s = 0; s < ks; ++s) {
        rng[2 * ((size_t)v * ks + s)] = offs[s][v];
        rng[2 * ((size_t)v * ks + s) + 1] = offs[s][v + 1];
      }
    S.rng = rng;
    held.push_back(rng);
  }
  qv.fmts = all5 ? SEG_F5 : all8 ? SEG_F8 : 0;
  std::vector<uint32_t> pairs(2 * q);
  if (!get(f, pairs.data(), 2 * q)) return 2;
  fclose(f);
  std::vector<int32_t> out(q, -7);
  if (const char* s = getenv("EMU_SEED")) emu_seed = strtoull(s, nullptr, 10);
  if (launch_query(qv, pairs.data(), n, q, out.data(), nullptr) != hipSuccess) return 3;
  FILE* g = fopen(argv[2], "wb");
  if (!g) return 2;
  fwrite(out.data(), sizeof(int32_t), q, g);
  fclose(g);
  for (void* p : held) free(p);
  return 0;
}
