// Test harness only: the tree kernel (plant_amd/csrc/tree.hip) run on the host
// through hip/hip_runtime.h of this directory, one persistent workgroup taking
// every root of one launch. Input and output are little-endian binary files:
//   in:  u32 n, m, R, delta, threads, slots, flags; u32 off[n + 1]; u64 arc[m];
//        u32 rank[n]; u32 roots[R] (vertex ids)
//   out: u64 labels; (u64 key = row << 32 | hub, u64 ent = hub << 32 | dist)
//        per label; u32 1 when every search cell is consistent with its
//        slot's counts after the launch (rank kept, no state of a later tree
//        than the slot's last, no stamp at or above its next stamp base), else 0
// EMU_EPOCH0 / EMU_SBASE0 start every slot's tree count and stamp base at the
// given values, so a launch crosses their bounds and resets the slot's cells.
// EMU_SORT_MIN sets the smallest frontier sorted in LDS (default 2: every
// frontier within the kernel's bound; 0: none).
#include "../../plant_amd/csrc/tree.hip"

#include <cstdio>
#include <cstdlib>
#include <vector>

using namespace chl;

template <class T>
static bool get(FILE* f, std::vector<T>& v, size_t count) {
  v.resize(count);
  return count == 0 || fread(v.data(), sizeof(T), count, f) == count;
}

int main(int argc, char** argv) {
  if (argc != 3) return 2;
  FILE* f = fopen(argv[1], "rb");
  if (!f) return 2;
  std::vector<uint32_t> hdr, off, rank, roots;
  std::vector<uint64_t> arc;
  if (!get(f, hdr, 7)) return 2;
  const uint32_t n = hdr[0], m = hdr[1], R = hdr[2], delta = hdr[3], threads = hdr[4], S = hdr[5], flags = hdr[6];
  if (!get(f, off, n + 1) || !get(f, arc, m) || !get(f, rank, n) || !get(f, roots, R)) return 2;
  fclose(f);
  const size_t cells = (size_t)S * n;
  std::vector<uint64_t> cell(2 * cells), qd(cells);
  for (size_t i = 0; i < cells; ++i) {
    cell[2 * i] = kStateInf;
    cell[2 * i + 1] = (uint64_t)rank[i % n] << 32;
  }
  std::vector<uint32_t> qa(cells), qb(cells), vis(cells), spill(R), ts(R), lc(R), gk(1), gv(1), seq(2 * S, 0);
  for (uint32_t b = 0; b < S; ++b) {
    if (const char* e = getenv("EMU_EPOCH0")) seq[2 * b] = (uint32_t)strtoul(e, nullptr, 10);
    if (const char* e = getenv("EMU_SBASE0")) seq[2 * b + 1] = (uint32_t)strtoul(e, nullptr, 10);
  }
  const std::vector<uint32_t> seq0 = seq;
  const uint64_t cap = (uint64_t)n * R + 1;
  std::vector<uint32_t> orow(cap);
  std::vector<uint64_t> ohd(cap);
  std::vector<unsigned long long> ctr(CTR_COUNT, 0);
  ctr[CTR_SPILL_START] = ~0ull;
  TreeParams p{};
  p.off = off.data();
  p.arc = arc.data();
  p.rank = rank.data();
  p.n = n;
  p.roots = roots.data();
  p.R = R;
  p.delta = delta;
  p.cell = cell.data();
  p.seq = seq.data();
  p.qa = qa.data();
  p.qb = qb.data();
  p.vis = vis.data();
  p.qd = qd.data();
  p.gkey = gk.data();
  p.gval = gv.data();
  p.gcap = 1;
  p.out_row = orow.data();
  p.out_hd = ohd.data();
  p.out_cap = cap;
  p.spill = spill.data();
  p.tree_size = ts.data();
  p.label_count = lc.data();
  p.ctr = ctr.data();
  p.sort_min = 2;
  if (const char* e = getenv("EMU_SORT_MIN")) p.sort_min = (uint32_t)strtoul(e, nullptr, 10);
  emu_only_block = (int)S - 1;  // the last slot: the largest scratch offsets
  if (const char* e = getenv("EMU_SEED")) emu_seed = strtoull(e, nullptr, 10);
  launch_tree(p, flags, S, (int)threads, nullptr);
  uint32_t restored = 1;
  for (size_t i = 0; i < cells; ++i) {
    const uint32_t b = (uint32_t)(i / n);
    const uint64_t x = cell[2 * i], y = cell[2 * i + 1];
    if ((uint32_t)(y >> 32) != rank[i % n]) restored = 0;
    if (x != kStateInf && state_ekey(x) < kEpochMax - seq[2 * b]) restored = 0;
    if ((uint32_t)y != 0 && (uint32_t)y >= seq[2 * b + 1]) restored = 0;
    // a slot that ran trees has counted them
    if (x != kStateInf && seq[2 * b] == seq0[2 * b] && seq0[2 * b] < kEpochMax) restored = 0;
  }
  FILE* o = fopen(argv[2], "wb");
  if (!o) return 2;
  const uint64_t L = ctr[CTR_LABELS];
  fwrite(&L, 8, 1, o);
  for (uint64_t k = 0; k < L; ++k) {
    const uint64_t key = ((uint64_t)orow[k] << 32) | (ohd[k] >> 32);
    fwrite(&key, 8, 1, o);
    fwrite(&ohd[k], 8, 1, o);
  }
  fwrite(&restored, 4, 1, o);
  fclose(o);
  return ctr[CTR_OVERFLOW] ? 3 : 0;
}
