// Test harness only: the sampled-betweenness kernel (plant_amd/csrc/order.hip)
// run on the host through hip/hip_runtime.h of this directory, with every
// scratch array sized exactly as order_device.cpp sizes it for `slots` slots.
// One chosen workgroup (the last slot: the largest scratch offsets) takes every
// root of the persistent launch. Built with AddressSanitizer by its test, so an
// index past any array ends the run.
//   in:  u32 n, m_out, m_in, R, delta, slots; u32 off[n + 1]; u64 arc[m_out];
//        u32 ioff[n + 1]; u64 iarc[m_in]; u32 inv[n]; u32 roots[R] (device ids)
//   out: u64 score[n] (per device id); u32 1 when every scratch word came back
//        to its rest state, else 0
#include "../../plant_amd/csrc/order.hip"

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
  if (!get(f, hdr, 6)) return 2;
  const uint32_t n = hdr[0], mo = hdr[1], mi = hdr[2], R = hdr[3], delta = hdr[4], S = hdr[5];
  if (!get(f, off, n + 1) || !get(f, arc, mo) || !get(f, ioff, n + 1) || !get(f, iarc, mi) || !get(f, inv, n) ||
      !get(f, roots, R))
    return 2;
  fclose(f);
  // the sizes of order_device.cpp: cell u64 per vertex per slot (two u32
  // words), acc u64, three u32 lists
  const size_t cells = (size_t)S * n;
  std::vector<uint64_t> cell(cells, 0x00000000FFFFFFFFull), acc(cells, 1), score(n, 0);
  std::vector<uint32_t> qa(cells), qb(cells), vis(cells);
  std::vector<unsigned long long> ctr(BC_COUNT, 0);
  BcParams p{};
  p.off = off.data();
  p.arc = arc.data();
  p.ioff = ioff.data();
  p.iarc = iarc.data();
  p.inv = inv.data();
  p.roots = roots.data();
  p.R = R;
  p.n = n;
  p.delta = delta;
  p.cell = reinterpret_cast<uint32_t*>(cell.data());
  p.acc = reinterpret_cast<unsigned long long*>(acc.data());
  p.qa = qa.data();
  p.qb = qb.data();
  p.vis = vis.data();
  p.score = reinterpret_cast<unsigned long long*>(score.data());
  p.ctr = ctr.data();
  emu_only_block = (int)S - 1;
  if (const char* e = getenv("EMU_SEED")) emu_seed = strtoull(e, nullptr, 10);
  launch_betweenness(p, S, nullptr);
  uint32_t restored = 1;
  for (size_t i = 0; i < cells; ++i)
    if (cell[i] != 0x00000000FFFFFFFFull || acc[i] != 1) restored = 0;
  FILE* o = fopen(argv[2], "wb");
  if (!o) return 2;
  fwrite(score.data(), 8, n, o);
  fwrite(&restored, 4, 1, o);
  fclose(o);
  return ctr[BC_NEXT] >= R ? 0 : 3;
}
