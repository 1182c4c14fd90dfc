// The segmented label store (host side): one segment per batch, appended in
// O(batch), merged as sizes and memory allow, consolidated after the build.
//
// The reference keeps every row as a nested host vector and appends each
// superstep's cleaned labels to it (shared_memory/labeling.hpp:43-54,110-120,
// driven by akiba.hpp:197-208). Here a superstep (batch) becomes a segment: a
// CSR over row ids holding the labels of the batch's hub range, in the compact
// SEG_F5 form when its distances and hub span allow and SEG_F8 otherwise. A
// vertex's row is the concatenation of its rows in the segments, in segment
// order, which is hub order. Both sides of a store always have the same
// segments (the same hub ranges), so a query merges segment by segment.
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <vector>

#include "dev.hpp"
#include "engine.hpp"
#include "internal.hpp"

namespace chl {

SegView Segment::view() const {
  SegView v;
  memset(&v, 0, sizeof(v));
  v.off = off.p;
  v.key = key.p;
  v.lo = lo.p;
  v.ent = ent.p;
  v.h0 = h0;
  v.fmt = fmt;
  return v;
}

uint64_t Segment::entry_bytes() const { return total * (fmt == SEG_F8 ? 8ull : 5ull); }

int Segment::alloc(uint64_t rows, uint64_t count, uint32_t f) {
  fmt = f;
  total = count;
  CHL_HIP(off.alloc(rows + 1));
  if (fmt == SEG_F8) {
    CHL_HIP(ent.alloc(count));
  } else {
    CHL_HIP(key.alloc(count + kF5Pad));
    CHL_HIP(lo.alloc(count + kF5Pad));
  }
  return CHL_OK;
}

int LabelStore::init(int64_t nn, bool shared, int dev) {
  n = nn;
  shared_sides = shared;
  device = dev;
  for (int s = 0; s < 2; ++s) {
    seg[s].clear();
    total[s] = 0;
    CHL_HIP(dview[s].reserve(kSegMax));
  }
  return CHL_OK;
}

uint64_t LabelStore::bytes() const {
  uint64_t b = 0;
  for (int s = 0; s < nsides(); ++s)
    for (auto& g : seg[s]) b += g->entry_bytes() + 8ull * (uint64_t)(n + 1);
  return b;
}

int LabelStore::sync_views() {
  for (int s = 0; s < nsides(); ++s) {
    std::vector<SegView> v;
    for (auto& g : seg[s]) v.push_back(g->view());
    if (!v.empty()) CHL_HIP(hipMemcpy(dview[s].p, v.data(), sizeof(SegView) * v.size(), hipMemcpyHostToDevice));
  }
  return CHL_OK;
}

int LabelStore::append(std::unique_ptr<Segment> s0, std::unique_ptr<Segment> s1) {
  if (seg[0].size() >= kSegMax) return set_error(CHL_E_NOMEM, "label store: more than %u segments", kSegMax);
  total[0] += s0->total;
  seg[0].push_back(std::move(s0));
  if (!shared_sides) {
    total[1] += s1->total;
    seg[1].push_back(std::move(s1));
  }
  return sync_views();
}

// Merges segments [a, b) of every stored side into one (rows concatenated in
// segment order). Needs the merged size in free device memory.
int LabelStore::merge_range(size_t a, size_t b, hipStream_t st) {
  if (b <= a + 1) return CHL_OK;
  const uint64_t rows = (uint64_t)n;
  DBuf<uint32_t> cnt;
  DBuf<uint64_t> tmp;
  CHL_HIP(cnt.reserve(std::max<uint64_t>(rows, 1)));
  CHL_HIP(tmp.reserve(scan_tmp_words(rows)));
  const auto wall = [] { return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count(); };
  double t_alloc = 0, t_copy = 0, t_free = 0;
  uint64_t moved = 0;
  for (int s = 0; s < nsides(); ++s) {
    auto& L = seg[s];
    uint64_t count = 0;
    bool f5 = L[b - 1]->h1 - L[a]->h0 <= kF5MaxWidth;
    for (size_t i = a; i < b; ++i) {
      count += L[i]->total;
      f5 = f5 && L[i]->fmt == SEG_F5;
    }
    std::unique_ptr<Segment> m(new Segment());
    m->h0 = L[a]->h0;
    m->h1 = L[b - 1]->h1;
    double t = wall();
    int rc = m->alloc(rows, count, f5 ? SEG_F5 : SEG_F8);
    if (rc) return rc;
    t_alloc += wall() - t;
    t = wall();
    const SegView* views = dview[s].p + a;  // the device copies of segments a..b
    CHL_HIP(launch_row_total(views, (uint32_t)(b - a), (uint32_t)rows, cnt.p, st));
    CHL_HIP(launch_scan_u32(cnt.p, rows, m->off.p, tmp.p, st));
    CHL_HIP(launch_concat_rows(views, (uint32_t)(b - a), (uint32_t)rows, m->off.p, m->fmt, m->h0, m->key.p, m->lo.p,
                               m->ent.p, st));
    CHL_HIP(hipStreamSynchronize(st));
    t_copy += wall() - t;
    moved += m->entry_bytes();
    t = wall();
    L.erase(L.begin() + a, L.begin() + b);
    L.insert(L.begin() + a, std::move(m));
    t_free += wall() - t;
  }
  if (verbose)
    fprintf(stderr, "merge segments [%zu, %zu): %.2f GB, alloc %.3f s, copy %.3f s, free %.3f s\n", a, b, moved / 1e9,
            t_alloc, t_copy, t_free);
  return sync_views();
}

namespace {
uint64_t seg_bytes(const LabelStore& S, size_t i) {
  uint64_t b = 0;
  for (int s = 0; s < S.nsides(); ++s) b += S.seg[s][i]->entry_bytes();
  return b;
}
size_t free_bytes() {
  size_t f = 0, t = 0;
  if (hipMemGetInfo(&f, &t) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return f;
}
}  // namespace

// After a batch: merge the newest segment into its predecessor while it is at
// least half as large (sizes then fall geometrically, so an entry is merged a
// logarithmic number of times), and merge the smallest neighbouring pair while
// there are more than kSegSoft segments; each merge only where the merged copy
// fits in `budget` bytes of free memory (0: a quarter of what is free).
int LabelStore::tidy(size_t budget, hipStream_t st) {
  constexpr size_t kSegSoft = 24;
  for (;;) {
    const size_t k = seg[0].size();
    if (k < 2) return CHL_OK;
    const size_t room = budget ? budget : free_bytes() / 4;
    const uint64_t last = seg_bytes(*this, k - 1), prev = seg_bytes(*this, k - 2);
    if (2 * last >= prev && last + prev + 16ull * (uint64_t)n <= room) {
      int rc = merge_range(k - 2, k, st);
      if (rc) return rc;
      continue;
    }
    if (k > kSegSoft) {
      size_t best = 0;
      uint64_t bb = ~0ull;
      for (size_t i = 0; i + 1 < k; ++i) {
        const uint64_t c = seg_bytes(*this, i) + seg_bytes(*this, i + 1);
        if (c < bb) {
          bb = c;
          best = i;
        }
      }
      if (bb + 16ull * (uint64_t)n <= free_bytes() / 2) {
        int rc = merge_range(best, best + 2, st);
        if (rc) return rc;
        continue;
      }
    }
    return CHL_OK;
  }
}

// After the build: one segment when its copy fits in free memory; otherwise the
// segments are cut into kSegFinal contiguous groups of about equal bytes and
// each group is merged in one pass (one allocation per group, every entry
// copied once; merging neighbouring pairs instead allocated and copied each
// entry two or three times, and on the 16M-vertex store the allocations alone
// took 11 s). A group whose copy does not fit in free memory is left as it is.
int LabelStore::consolidate(hipStream_t st) {
  constexpr size_t kSegFinal = 4;
  const size_t k = seg[0].size();
  if (k < 2) return CHL_OK;
  uint64_t all = 0;
  for (size_t i = 0; i < k; ++i) all += seg_bytes(*this, i);
  if (all + 16ull * (uint64_t)n <= free_bytes() * 9 / 10) return merge_range(0, k, st);
  if (k <= kSegFinal) return CHL_OK;
  // group ends: close a group once it holds its share of what is left
  std::vector<size_t> ends;
  {
    size_t i = 0;
    uint64_t left = all;
    for (size_t g = 0; g < kSegFinal && i < k; ++g) {
      const uint64_t share = left / (kSegFinal - g);
      uint64_t acc = 0;
      size_t j = i;
      // at least one segment per group, and leave one for each later group
      while (j < k && (j == i || acc + seg_bytes(*this, j) / 2 <= share) && k - j > kSegFinal - g - 1) {
        acc += seg_bytes(*this, j);
        ++j;
      }
      if (g == kSegFinal - 1) j = k;
      ends.push_back(j);
      for (size_t t = i; t < j; ++t) left -= seg_bytes(*this, t);
      i = j;
    }
  }
  // merge from the last group down, so the earlier groups' indices stay valid
  for (size_t g = ends.size(); g-- > 0;) {
    const size_t lo = g ? ends[g - 1] : 0, hi = ends[g];
    if (hi - lo < 2) continue;
    uint64_t bytes = 0;
    for (size_t t = lo; t < hi; ++t) bytes += seg_bytes(*this, t);
    if (bytes + 16ull * (uint64_t)n > free_bytes() * 9 / 10) continue;
    int rc = merge_range(lo, hi, st);
    if (rc) return rc;
  }
  return CHL_OK;
}

int LabelStore::build_ranges(const uint32_t* row_of, hipStream_t st) {
  while (nseg() > kQuerySeg) {
    // a kept store of many segments: merge the smallest neighbouring pair
    size_t best = 0;
    uint64_t bb = ~0ull;
    for (size_t i = 0; i + 1 < nseg(); ++i) {
      const uint64_t c = seg_bytes(*this, i) + seg_bytes(*this, i + 1);
      if (c < bb) {
        bb = c;
        best = i;
      }
    }
    int rc = merge_range(best, best + 2, st);
    if (rc) return rc;
  }
  const uint32_t nn = (uint32_t)n, k = nseg();
  for (int s = 0; s < nsides(); ++s) {
    CHL_HIP(rng[s].alloc(2 * (size_t)std::max<uint32_t>(nn, 1) * std::max<uint32_t>(k, 1)));
    if (k) CHL_HIP(launch_row_ranges(dview[s].p, k, row_of, nn, rng[s].p, st));
  }
  CHL_HIP(hipStreamSynchronize(st));
  return CHL_OK;
}

QueryView LabelStore::query_view() const {
  QueryView q;
  memset(&q, 0, sizeof(q));
  q.nseg = nseg();
  bool all5 = true, all8 = true;
  for (int side = 0; side < 2; ++side) {
    QuerySide& Q = side ? q.s1 : q.s0;
    const int s = shared_sides ? 0 : side;
    Q.rng = rng[s].p;
    for (uint32_t i = 0; i < q.nseg && i < kQuerySeg; ++i) {
      const Segment& g = *seg[s][i];
      Q.key[i] = g.key.p;
      Q.lo[i] = g.lo.p;
      Q.ent[i] = g.ent.p;
      Q.h0[i] = g.h0;
      Q.fmt[i] = g.fmt;
      all5 = all5 && g.fmt == SEG_F5;
      all8 = all8 && g.fmt == SEG_F8;
    }
  }
  q.fmts = all5 ? SEG_F5 : all8 ? SEG_F8 : 0;
  return q;
}

// Host rows in row-id order, entries (hub << 32) | dist.
int LabelStore::download(int side, std::vector<uint64_t>& hoff, std::vector<uint64_t>& hent) const {
  const int s = shared_sides ? 0 : side;
  const auto& L = seg[s];
  const size_t k = L.size();
  std::vector<std::vector<uint64_t>> so(k), se(k);
  DBuf<uint64_t> dec;
  for (size_t i = 0; i < k; ++i) {
    so[i].resize(n + 1);
    CHL_HIP(hipMemcpy(so[i].data(), L[i]->off.p, sizeof(uint64_t) * (n + 1), hipMemcpyDeviceToHost));
    se[i].resize(L[i]->total);
    if (!L[i]->total) continue;
    if (L[i]->fmt == SEG_F8) {
      CHL_HIP(hipMemcpy(se[i].data(), L[i]->ent.p, sizeof(uint64_t) * L[i]->total, hipMemcpyDeviceToHost));
    } else {
      // decode on the device in chunks
      const uint64_t chunk = (uint64_t)1 << 26;
      CHL_HIP(dec.reserve(std::min<uint64_t>(chunk, L[i]->total)));
      SegView v = L[i]->view();
      for (uint64_t c0 = 0; c0 < L[i]->total; c0 += chunk) {
        const uint64_t c = std::min(chunk, L[i]->total - c0);
        SegView w = v;
        w.key = v.key + c0;
        w.lo = v.lo + c0;
        CHL_HIP(launch_decode(w, c, dec.p, nullptr));
        CHL_HIP(hipMemcpy(se[i].data() + c0, dec.p, sizeof(uint64_t) * c, hipMemcpyDeviceToHost));
      }
    }
  }
  hoff.assign(n + 1, 0);
  for (int64_t r = 0; r < n; ++r) {
    uint64_t len = 0;
    for (size_t i = 0; i < k; ++i) len += so[i][r + 1] - so[i][r];
    hoff[r + 1] = hoff[r] + len;
  }
  hent.resize(hoff[n]);
  parallel_for(n, host_threads(), 4096, [&](int64_t r) {
    uint64_t o = hoff[r];
    for (size_t i = 0; i < k; ++i)
      for (uint64_t j = so[i][r]; j < so[i][r + 1]; ++j) hent[o++] = se[i][j];
  });
  return CHL_OK;
}

// One segment from host rows (hub-ascending): SEG_F5 when every hub is below
// 2^24 and every distance at most 65535, SEG_F8 otherwise.
int LabelStore::upload(int side, const std::vector<uint64_t>& hoff, const std::vector<uint64_t>& hent) {
  bool f5 = true;
  uint32_t hmax = 0;
  for (uint64_t e : hent) {
    const uint32_t h = (uint32_t)(e >> 32), d = (uint32_t)e;
    hmax = std::max(hmax, h);
    if (h >= kF5MaxWidth || d > kF5MaxDist) f5 = false;
  }
  std::unique_ptr<Segment> g(new Segment());
  g->h0 = 0;
  g->h1 = hent.empty() ? 0 : hmax + 1;
  int rc = g->alloc((uint64_t)n, hent.size(), f5 ? SEG_F5 : SEG_F8);
  if (rc) return rc;
  CHL_HIP(hipMemcpy(g->off.p, hoff.data(), sizeof(uint64_t) * (n + 1), hipMemcpyHostToDevice));
  if (!hent.empty()) {
    if (f5) {
      std::vector<uint32_t> key(hent.size());
      std::vector<uint8_t> lo(hent.size());
      parallel_for((int64_t)hent.size(), host_threads(), (int64_t)1 << 20, [&](int64_t k) {
        const uint32_t h = (uint32_t)(hent[k] >> 32), d = (uint32_t)hent[k];
        key[k] = (h << 8) | (d >> 8);
        lo[k] = (uint8_t)(d & 0xFFu);
      });
      CHL_HIP(hipMemcpy(g->key.p, key.data(), sizeof(uint32_t) * key.size(), hipMemcpyHostToDevice));
      CHL_HIP(hipMemcpy(g->lo.p, lo.data(), lo.size(), hipMemcpyHostToDevice));
    } else {
      CHL_HIP(hipMemcpy(g->ent.p, hent.data(), sizeof(uint64_t) * hent.size(), hipMemcpyHostToDevice));
    }
  }
  total[side] = hent.size();
  seg[side].clear();
  seg[side].push_back(std::move(g));
  return CHL_OK;
}

}  // namespace chl
