// Test harness only: the hashed-tree kernel (plant_amd/csrc/small.hip) run on
// the host through hip/hip_runtime.h of this directory, one wavefront per root
// over small tables, then optionally 256 threads per root over medium tables for
// the roots handed over, then the tree kernel (tree.hip) over the rest, as the
// label build does. Input and output are little-endian binary files:
//   in:  u32 n, m, R, delta, cap, slots, bigstep, cap2, slots2, bigstep2;
//        u32 off[n + 1]; u64 arc[m]; u32 rank[n]; u32 roots[R] (vertex ids).
//        Every bigstep-th root (0: none) is marked big for the small tables,
//        every bigstep2-th for the medium ones; cap and cap2 are the tables'
//        entries (powers of two; cap2 0: no medium stage).
//   out: u64 labels; (u64 key = row << 32 | hub, u64 ent = hub << 32 | dist)
//        per label; u32 roots the small stage completed; u32 roots the medium
//        stage completed; u32 1 when every table is empty again (key and stamp
//        0, state all ones), else 0
#include "../../plant_amd/csrc/tree.hip"
#include "../../plant_amd/csrc/small.hip"

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
  if (!get(f, hdr, 10)) return 2;
  const uint32_t n = hdr[0], m = hdr[1], R = hdr[2], delta = hdr[3], C = hdr[4], S = hdr[5], bigstep = hdr[6];
  const uint32_t C2 = hdr[7], S2 = hdr[8], bigstep2 = hdr[9];
  if (!get(f, off, n + 1) || !get(f, arc, m) || !get(f, rank, n) || !get(f, roots, R)) return 2;
  fclose(f);
  if (C < 64 || (C & (C - 1))) return 2;
  if (C2 && (C2 < 64 || (C2 & (C2 - 1)) || !S2)) return 2;
  if (const char* e = getenv("EMU_SEED")) emu_seed = strtoull(e, nullptr, 10);

  // small slots
  std::vector<uint64_t> sstate((size_t)S * C, kStateInf);
  std::vector<uint32_t> skey((size_t)S * C, 0), sstamp((size_t)S * C, 0), slist((size_t)S * 3 * C, 0);
  std::vector<uint8_t> big(R, 0);
  for (uint32_t k = 0; bigstep && k < R; ++k) big[k] = k % bigstep == 0;
  const uint64_t cap = (uint64_t)n * R + 1;
  std::vector<uint32_t> orow(cap), spill(R), ts(R), lc(R);
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
  p.out_row = orow.data();
  p.out_hd = ohd.data();
  p.out_cap = cap;
  p.spill = spill.data();
  p.tree_size = ts.data();
  p.label_count = lc.data();
  p.ctr = ctr.data();
  p.sstate = sstate.data();
  p.skey = skey.data();
  p.sstamp = sstamp.data();
  p.slist = slist.data();
  p.scap = C;
  p.big = big.data();
  emu_only_block = (int)S - 1;  // the last slot: the largest scratch offsets
  launch_small_trees(p, S, 64, nullptr);
  uint32_t empty = 1;
  for (size_t i = 0; i < sstate.size(); ++i)
    if (sstate[i] != kStateInf || skey[i] != 0 || sstamp[i] != 0) empty = 0;
  uint32_t handed = (uint32_t)ctr[CTR_OVERFLOW];
  const uint32_t done_small = R - handed;
  if (ctr[CTR_SKIPPED] != handed) return 4;  // the output never fills here
  // batch positions of the roots handed over
  std::vector<uint32_t> pos(handed);
  for (uint32_t k = 0; k < handed; ++k) pos[k] = spill[k];

  // the medium stage over the roots handed over (their own big flags)
  uint32_t done_medium = 0;
  if (C2) {
    std::vector<uint64_t> mstate((size_t)S2 * C2, kStateInf);
    std::vector<uint32_t> mkey((size_t)S2 * C2, 0), mstamp((size_t)S2 * C2, 0), mlist((size_t)S2 * 3 * C2, 0);
    std::vector<uint32_t> mroots(handed), mspill(handed + 1), mts(handed + 1), mlc(handed + 1);
    std::vector<uint8_t> mbig(handed + 1, 0);
    for (uint32_t k = 0; k < handed; ++k) {
      mroots[k] = roots[pos[k]];
      mbig[k] = bigstep2 && pos[k] % bigstep2 == 0;
    }
    ctr[CTR_OVERFLOW] = 0;
    ctr[CTR_NEXT] = 0;
    ctr[CTR_SKIPPED] = 0;
    p.roots = mroots.data();
    p.R = handed;
    p.spill = mspill.data();
    p.tree_size = mts.data();
    p.label_count = mlc.data();
    p.sstate = mstate.data();
    p.skey = mkey.data();
    p.sstamp = mstamp.data();
    p.slist = mlist.data();
    p.scap = C2;
    p.big = mbig.data();
    emu_only_block = (int)S2 - 1;
    if (handed) launch_small_trees(p, S2, 256, nullptr);
    for (size_t i = 0; i < mstate.size(); ++i)
      if (mstate[i] != kStateInf || mkey[i] != 0 || mstamp[i] != 0) empty = 0;
    const uint32_t h2 = (uint32_t)ctr[CTR_OVERFLOW];
    if (ctr[CTR_SKIPPED] != h2) return 4;
    done_medium = handed - h2;
    std::vector<uint32_t> pos2(h2);
    for (uint32_t k = 0; k < h2; ++k) pos2[k] = pos[mspill[k]];
    pos.swap(pos2);
    handed = h2;
  }

  // the tree kernel over the roots handed over
  std::vector<uint32_t> roots2(handed);
  for (uint32_t k = 0; k < handed; ++k) roots2[k] = roots[pos[k]];
  const size_t cells = (size_t)n;
  std::vector<uint64_t> cell(2 * cells), qd(cells);
  for (size_t i = 0; i < cells; ++i) {
    cell[2 * i] = kStateInf;
    cell[2 * i + 1] = (uint64_t)rank[i] << 32;
  }
  std::vector<uint32_t> qa(cells), qb(cells), vis(cells), gk(1), gv(1), seq(2, 0), spill2(R + 1);
  ctr[CTR_OVERFLOW] = 0;
  ctr[CTR_NEXT] = 0;
  ctr[CTR_SKIPPED] = 0;
  p.roots = roots2.data();
  p.R = handed;
  p.cell = cell.data();
  p.seq = seq.data();
  p.qa = qa.data();
  p.qb = qb.data();
  p.vis = vis.data();
  p.qd = qd.data();
  p.gkey = gk.data();
  p.gval = gv.data();
  p.gcap = 1;
  p.spill = spill2.data();
  p.big = nullptr;
  emu_only_block = 0;
  if (handed) launch_tree(p, TREE_FLAGS, 1, 64, nullptr);
  if (ctr[CTR_OVERFLOW]) return 3;

  FILE* o = fopen(argv[2], "wb");
  if (!o) return 2;
  const uint64_t L = ctr[CTR_LABELS];
  fwrite(&L, 8, 1, o);
  for (uint64_t k = 0; k < L; ++k) {
    const uint64_t key = ((uint64_t)orow[k] << 32) | (ohd[k] >> 32);
    fwrite(&key, 8, 1, o);
    fwrite(&ohd[k], 8, 1, o);
  }
  fwrite(&done_small, 4, 1, o);
  fwrite(&done_medium, 4, 1, o);
  fwrite(&empty, 4, 1, o);
  fclose(o);
  return 0;
}
