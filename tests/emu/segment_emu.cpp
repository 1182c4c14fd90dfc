// Test harness only: the label-store kernels of plant_amd/csrc/labels.hip run
// on the host through hip/hip_runtime.h of this directory, every workgroup of
// every launch one after another:
//  * the segment build (row count, scan, cursor copy, scatter, per-row sort) of
//    each input batch, launched as engine.cpp's build_segment launches them;
//  * then, by mode: 0 nothing more; 1 the merge of all the batches' segments
//    (row totals, scan, concatenation) into one segment of format out_fmt, as
//    store.cpp's merge_range does; 2 the compaction of the first segment under
//    a keep flag per entry (count kept, scan, copy kept), as compact_side does,
//    with keep = ((hub * 2654435761) ^ row) % 3 != 0; 3 the cleaning of the
//    first segment against the second as the other side's batch rows
//    (clean_side), the keep flags written after the entries.
// Input and output are little-endian binary files:
//   in:  u32 n (rows), u32 batches, u32 mode, u32 out_fmt; per batch: u32 fmt
//        (SEG_F5 or SEG_F8), u32 h0 (the batch's first hub), u64 m, u32
//        row[m], u64 hd[m] = (hub << 32) | dist in any order
//   out: u64 off[n + 1] of the result; per entry in stored order u32 hub, u32 dist;
//        mode 3: then u8 keep per entry of the first segment
#include "../../plant_amd/csrc/tree.hip"  // (the fill kernels)
#include "../../plant_amd/csrc/labels.hip"

#include <cstdio>
#include <cstdlib>
#include <vector>

using namespace chl;

struct HostSeg {
  uint32_t fmt = SEG_F8, h0 = 0;
  uint64_t m = 0;
  std::vector<uint64_t> off;
  std::vector<uint32_t> key;
  std::vector<uint8_t> lo;
  std::vector<uint64_t> ent;
  void alloc(uint32_t n, uint64_t count) {
    off.assign(n + 1, 0);
    m = count;
    key.assign(fmt == SEG_F5 ? count + kF5Pad : 1, 0);
    lo.assign(fmt == SEG_F5 ? count + kF5Pad : 1, 0);
    ent.assign(fmt == SEG_F8 ? count + 1 : 1, 0);
  }
  SegView view() const { return SegView{off.data(), key.data(), lo.data(), ent.data(), h0, fmt}; }
  void entry(uint64_t k, uint32_t& h, uint32_t& d) const {
    if (fmt == SEG_F8) {
      h = (uint32_t)(ent[k] >> 32);
      d = (uint32_t)ent[k];
    } else {
      h = (key[k] >> 8) + h0;
      d = ((key[k] & 0xFFu) << 8) | lo[k];
    }
  }
};

int main(int argc, char** argv) {
  if (argc != 3) return 2;
  FILE* f = fopen(argv[1], "rb");
  if (!f) return 2;
  uint32_t hdr[4];
  if (fread(hdr, 4, 4, f) != 4) return 2;
  const uint32_t n = hdr[0], nb = hdr[1], mode = hdr[2], out_fmt = hdr[3];
  if (!nb || mode > 3 || (mode == 3 && nb != 2) || (out_fmt != SEG_F5 && out_fmt != SEG_F8)) return 2;
  if (const char* e = getenv("EMU_SEED")) emu_seed = strtoull(e, nullptr, 10);
  std::vector<HostSeg> segs(nb);
  std::vector<uint32_t> cnt(n), longq(n + 1);
  std::vector<uint64_t> cur(n), tmp(scan_tmp_words(n) + 1);
  for (uint32_t b = 0; b < nb; ++b) {
    uint32_t bh[2];
    uint64_t m = 0;
    if (fread(bh, 4, 2, f) != 2 || fread(&m, 8, 1, f) != 1) return 2;
    std::vector<uint32_t> row(m);
    std::vector<uint64_t> hd(m);
    if (m && (fread(row.data(), 4, m, f) != m || fread(hd.data(), 8, m, f) != m)) return 2;
    for (uint64_t k = 0; k < m; ++k)
      if (row[k] >= n) return 2;
    HostSeg& g = segs[b];
    g.fmt = bh[0];
    g.h0 = bh[1];
    if (g.fmt != SEG_F5 && g.fmt != SEG_F8) return 2;
    g.alloc(n, m);
    std::fill(cnt.begin(), cnt.end(), 0);
    std::fill(longq.begin(), longq.end(), 0);
    launch_count_rows(row.data(), m, cnt.data(), nullptr);
    launch_scan_u32(cnt.data(), n, g.off.data(), tmp.data(), nullptr);
    if (g.off[n] != m) return 3;
    if (m) {
      launch_copy_u64(g.off.data(), n, cur.data(), nullptr);
      launch_scatter_labels(row.data(), hd.data(), m, cur.data(), g.fmt, g.h0, g.key.data(), g.lo.data(),
                            g.ent.data(), nullptr);
      launch_sort_rows(g.off.data(), n, g.fmt, g.key.data(), g.lo.data(), g.ent.data(), longq.data(), nullptr);
      for (uint32_t r = 0; r < n; ++r)
        if (cur[r] != g.off[r + 1]) return 4;  // every cursor at its row's end
    }
  }
  fclose(f);

  HostSeg res;
  std::vector<uint8_t> clean_keep;
  if (mode == 0) {
    res = segs[0];
  } else if (mode == 3) {
    res = segs[0];
    clean_keep.assign(segs[0].m + 1, 0xFF);
    std::vector<unsigned long long> ctr(CTR_COUNT, 0);
    launch_clean(segs[0].view(), segs[1].view(), n, clean_keep.data(), ctr.data(), nullptr);
  } else if (mode == 1) {
    std::vector<SegView> views(nb);
    for (uint32_t b = 0; b < nb; ++b) views[b] = segs[b].view();
    launch_row_total(views.data(), nb, n, cnt.data(), nullptr);
    res.fmt = out_fmt;
    res.h0 = segs[0].h0;
    std::vector<uint64_t> noff(n + 1);
    launch_scan_u32(cnt.data(), n, noff.data(), tmp.data(), nullptr);
    res.alloc(n, noff[n]);
    res.off = noff;
    launch_concat_rows(views.data(), nb, n, res.off.data(), res.fmt, res.h0, res.key.data(), res.lo.data(),
                       res.ent.data(), nullptr);
  } else {
    const HostSeg& g = segs[0];
    std::vector<uint8_t> keep(g.m + 1, 0);
    for (uint32_t r = 0; r < n; ++r)
      for (uint64_t k = g.off[r]; k < g.off[r + 1]; ++k) {
        uint32_t h, d;
        g.entry(k, h, d);
        keep[k] = ((h * 2654435761u) ^ r) % 3 != 0;
      }
    launch_count_kept(g.off.data(), keep.data(), n, cnt.data(), nullptr);
    std::vector<uint64_t> noff(n + 1);
    launch_scan_u32(cnt.data(), n, noff.data(), tmp.data(), nullptr);
    res.fmt = g.fmt;
    res.h0 = g.h0;
    res.alloc(n, noff[n]);
    res.off = noff;
    launch_copy_kept(g.view(), keep.data(), n, res.off.data(), res.key.data(), res.lo.data(), res.ent.data(), nullptr);
  }

  FILE* o = fopen(argv[2], "wb");
  if (!o) return 2;
  fwrite(res.off.data(), 8, n + 1, o);
  for (uint64_t k = 0; k < res.m; ++k) {
    uint32_t h, d;
    res.entry(k, h, d);
    fwrite(&h, 4, 1, o);
    fwrite(&d, 4, 1, o);
  }
  if (mode == 3) fwrite(clean_keep.data(), 1, res.m, o);
  fclose(o);
  return 0;
}
