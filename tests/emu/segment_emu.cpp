// Test harness only: the segment build of plant_amd/csrc/labels.hip (row count,
// scan, cursor copy, scatter, per-row sort) run on the host through
// hip/hip_runtime.h of this directory, every workgroup of every launch one after
// another. Input and output are little-endian binary files:
//   in:  u32 n (rows), u32 fmt (SEG_F5 or SEG_F8), u32 h0 (the batch's first
//        hub), u64 m; u32 row[m]; u64 hd[m] = (hub << 32) | dist, in any order
//   out: u64 off[n + 1]; per stored entry, in stored order, u32 hub, u32 dist
#include "../../plant_amd/csrc/tree.hip"  // (the fill kernels)
#include "../../plant_amd/csrc/labels.hip"

#include <cstdio>
#include <cstdlib>
#include <vector>

using namespace chl;

int main(int argc, char** argv) {
  if (argc != 3) return 2;
  FILE* f = fopen(argv[1], "rb");
  if (!f) return 2;
  uint32_t hdr[3];
  uint64_t m = 0;
  if (fread(hdr, 4, 3, f) != 3 || fread(&m, 8, 1, f) != 1) return 2;
  const uint32_t n = hdr[0], fmt = hdr[1], h0 = hdr[2];
  if (fmt != SEG_F5 && fmt != SEG_F8) return 2;
  std::vector<uint32_t> row(m);
  std::vector<uint64_t> hd(m);
  if (m && (fread(row.data(), 4, m, f) != m || fread(hd.data(), 8, m, f) != m)) return 2;
  fclose(f);
  for (uint64_t k = 0; k < m; ++k)
    if (row[k] >= n) return 2;
  if (const char* e = getenv("EMU_SEED")) emu_seed = strtoull(e, nullptr, 10);

  // the builder's order of launches (engine.cpp build_segment)
  std::vector<uint32_t> cnt(n, 0), longq(n + 1, 0);
  std::vector<uint64_t> off(n + 1, 0), cur(n, 0), tmp(scan_tmp_words(n) + 1, 0);
  std::vector<uint32_t> key(fmt == SEG_F5 ? m + 1 : 1);
  std::vector<uint8_t> lo(fmt == SEG_F5 ? m + 1 : 1);
  std::vector<uint64_t> ent(fmt == SEG_F8 ? m + 1 : 1);
  launch_count_rows(row.data(), m, cnt.data(), nullptr);
  launch_scan_u32(cnt.data(), n, off.data(), tmp.data(), nullptr);
  if (off[n] != m) return 3;
  if (m) {
    launch_copy_u64(off.data(), n, cur.data(), nullptr);
    launch_scatter_labels(row.data(), hd.data(), m, cur.data(), fmt, h0, key.data(), lo.data(), ent.data(), nullptr);
    launch_sort_rows(off.data(), n, fmt, key.data(), lo.data(), ent.data(), longq.data(), nullptr);
  }
  for (uint32_t r = 0; r < n; ++r)
    if (cur[r] != off[r + 1]) return 4;  // every cursor at its row's end

  FILE* o = fopen(argv[2], "wb");
  if (!o) return 2;
  fwrite(off.data(), 8, n + 1, o);
  for (uint64_t k = 0; k < m; ++k) {
    uint32_t h, d;
    if (fmt == SEG_F8) {
      h = (uint32_t)(ent[k] >> 32);
      d = (uint32_t)ent[k];
    } else {
      h = (key[k] >> 8) + h0;
      d = ((key[k] & 0xFFu) << 8) | lo[k];
    }
    fwrite(&h, 4, 1, o);
    fwrite(&d, 4, 1, o);
  }
  fclose(o);
  return 0;
}
