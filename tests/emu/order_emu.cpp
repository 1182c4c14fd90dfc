// Test harness only: the device sampled-betweenness pipeline (the tree kernel
// of plant_amd/csrc/tree.hip in its TREE_KEEP mode, then the four passes of
// plant_amd/csrc/order.hip) run on the host through hip/hip_runtime.h of this
// directory, every workgroup of every launch one after another, over arrays
// sized exactly as order_device.cpp sizes them for `slots` slots. Built with
// AddressSanitizer by its test, so an index past any array ends the run.
//   in:  u32 n, m_out, m_in, R, delta, slots, threads; u32 off[n + 1];
//        u64 arc[m_out]; u32 ioff[n + 1]; u64 iarc[m_in]; u32 inv[n];
//        u32 roots[R] (device ids)
//   out: u64 score[n] (per device id); u32 1 when every search cell came back
//        to its unreached state, else 0
#include "../../plant_amd/csrc/tree.hip"
#include "../../plant_amd/csrc/order.hip"

#include <algorithm>
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
  std::vector<uint32_t> hdr, off, ioff, inv, roots;
  std::vector<uint64_t> arc, iarc;
  if (!get(f, hdr, 7)) return 2;
  const uint32_t n = hdr[0], mo = hdr[1], mi = hdr[2], R = hdr[3], delta = hdr[4], S = hdr[5], T = hdr[6];
  if (!get(f, off, n + 1) || !get(f, arc, mo) || !get(f, ioff, n + 1) || !get(f, iarc, mi) || !get(f, inv, n) ||
      !get(f, roots, R))
    return 2;
  fclose(f);
  const size_t cells = (size_t)S * n;
  std::vector<uint64_t> cell(2 * cells), qd(cells), acc(cells), score(n, 0);
  std::vector<uint32_t> qa(cells), qb(cells), vis(cells), parent(cells), leaf(cells), zero(n, 0);
  std::vector<uint32_t> touched(S), nleaf(S), tsize(S), lcount(S), spill(S), seq(2 * S, 0);
  std::vector<unsigned long long> ctr(CTR_COUNT, 0);
  launch_init_cells(cell.data(), cells, zero.data(), n, nullptr);
  TreeParams p{};
  p.off = off.data();
  p.arc = arc.data();
  p.rank = zero.data();
  p.n = n;
  p.delta = delta;
  p.cell = cell.data();
  p.seq = seq.data();
  p.qa = qa.data();
  p.qb = qb.data();
  p.vis = vis.data();
  p.qd = qd.data();
  p.spill = spill.data();
  p.tree_size = tsize.data();
  p.label_count = lcount.data();
  p.touched = touched.data();
  p.ctr = ctr.data();
  BcTrees t{};
  t.cell = cell.data();
  t.vis = vis.data();
  t.touched = touched.data();
  t.parent = parent.data();
  t.acc = reinterpret_cast<unsigned long long*>(acc.data());
  t.leaf = leaf.data();
  t.nleaf = nleaf.data();
  t.ioff = ioff.data();
  t.iarc = iarc.data();
  t.inv = inv.data();
  t.score = reinterpret_cast<unsigned long long*>(score.data());
  t.n = n;
  if (const char* e = getenv("EMU_SEED")) emu_seed = strtoull(e, nullptr, 10);
  for (uint32_t r0 = 0; r0 < R; r0 += S) {
    const uint32_t s = std::min<uint32_t>(S, R - r0);
    p.roots = roots.data() + r0;
    p.R = s;
    t.roots = roots.data() + r0;
    t.S = s;
    std::fill(ctr.begin(), ctr.end(), 0);
    std::fill(nleaf.begin(), nleaf.end(), 0);
    launch_tree(p, TREE_KEEP, s, (int)T, nullptr);
    launch_bc_trees(t, n, nullptr);
    launch_init_cells(cell.data(), (size_t)s * n, zero.data(), n, nullptr);
  }
  uint32_t restored = 1;
  for (size_t i = 0; i < cells; ++i)
    if (cell[2 * i] != kStateInf || cell[2 * i + 1] != 0) restored = 0;
  FILE* o = fopen(argv[2], "wb");
  if (!o) return 2;
  fwrite(score.data(), 8, n, o);
  fwrite(&restored, 4, 1, o);
  fclose(o);
  return 0;
}
