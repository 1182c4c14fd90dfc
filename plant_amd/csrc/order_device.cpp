// Host side of the device vertex orders. Sampled betweenness: roots, scratch
// slots, the persistent tree launch, then a stable radix sort of (~score, id).
// Degree: per-vertex keys, then one or two stable radix sorts.
#include <algorithm>
#include <cstring>

#include "engine.hpp"
#include "internal.hpp"

namespace chl {

namespace {

// Waits for a stream's queued work when it leaves scope.
struct StreamDrain {
  hipStream_t s;
  ~StreamDrain() {
    if (s) (void)hipStreamSynchronize(s);
  }
};

int order_impl(const Graph& g, DeviceGraph& dg, int device, int64_t roots, uint64_t seed,
               std::vector<uint32_t>& order, std::vector<uint64_t>* score, double* device_ms, size_t* scratch) {
  const int64_t n = g.n;
  order.resize(n);
  if (score) score->assign(n, 0);
  if (device_ms) *device_ms = 0;
  if (n == 0) return CHL_OK;
  if (n >= 0xFFFFFFFFll) return set_error(CHL_E_RANGE, "more than 2^32 - 1 vertices");
  int rc = ensure_device(device);
  if (rc) return rc;
  if (dg.device != device || dg.n != n) {
    rc = dg.build(g, device);
    if (rc) return rc;
  }
  std::vector<uint32_t> rs;
  sample_roots(n, roots, seed, rs);
  const uint32_t R = (uint32_t)rs.size();
  for (auto& r : rs) r = dg.pos[r];

  // bucket width: the builds' default
  const uint32_t delta = (uint32_t)default_delta(g);

  hipStream_t stream = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  CHL_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  struct Guard {
    hipStream_t& s;
    hipEvent_t &a, &b;
    ~Guard() {
      if (a) (void)hipEventDestroy(a);
      if (b) (void)hipEventDestroy(b);
      if (s) (void)hipStreamDestroy(s);
    }
  } guard{stream, e0, e1};
  CHL_HIP(hipEventCreate(&e0));
  CHL_HIP(hipEventCreate(&e1));

  // slots: 48 bytes of scratch per vertex each (the tree kernel's 16-byte cell,
  // its three lists and pile; parent, accumulator and leaf list), within a
  // third of the free memory
  size_t freeb = 0, totalb = 0;
  CHL_HIP(hipMemGetInfo(&freeb, &totalb));
  const size_t per = (size_t)n * 48;
  const size_t fit = std::max<size_t>(1, (freeb / 3) / per);
  const uint32_t slots = (uint32_t)std::max<size_t>(1, std::min<size_t>({fit, (size_t)R, (size_t)256}));
  const size_t cells = (size_t)slots * n;
  *scratch = cells * 48;
  DBuf<uint64_t> cell, qd, acc, score_d, key, val, skey, sval;
  DBuf<uint32_t> qa, qb, vis, parent, leaf, zero, inv, droots, touched, nleaf, tsize, lcount, spill, seq;
  DBuf<unsigned long long> ctr;
  DBuf<uint8_t> tmp;
  StreamDrain drain{stream};
  CHL_HIP(cell.reserve(2 * cells));
  CHL_HIP(qd.reserve(cells));
  CHL_HIP(acc.reserve(cells));
  CHL_HIP(qa.reserve(cells));
  CHL_HIP(qb.reserve(cells));
  CHL_HIP(vis.reserve(cells));
  CHL_HIP(parent.reserve(cells));
  CHL_HIP(leaf.reserve(cells));
  CHL_HIP(zero.reserve(n));
  CHL_HIP(score_d.reserve(n));
  CHL_HIP(inv.reserve(n));
  CHL_HIP(droots.reserve(std::max<uint32_t>(R, 1)));
  for (auto* b : {&touched, &nleaf, &tsize, &lcount, &spill}) CHL_HIP(b->reserve(slots));
  CHL_HIP(seq.reserve(2 * (size_t)slots));
  CHL_HIP(ctr.reserve(CTR_COUNT));
  CHL_HIP(hipMemcpyAsync(inv.p, dg.inv.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice, stream));
  if (R) CHL_HIP(hipMemcpyAsync(droots.p, rs.data(), sizeof(uint32_t) * R, hipMemcpyHostToDevice, stream));
  // every vertex has rank 0 here: the searches expand every vertex they reach
  CHL_HIP(launch_fill_u32(zero.p, 0, n, stream));
  CHL_HIP(launch_init_cells(cell.p, cells, zero.p, (uint32_t)n, stream));
  CHL_HIP(launch_fill_u32(seq.p, 0, 2 * (size_t)slots, stream));
  CHL_HIP(launch_fill_u64(score_d.p, 0, n, stream));

  const int ir = dg.sym ? 1 : 0;
  TreeParams p;
  memset(&p, 0, sizeof(p));
  p.off = dg.off[1].p;
  p.arc = dg.arc[1].p;
  p.rank = zero.p;
  p.n = (uint32_t)n;
  p.delta = delta;
  p.cell = cell.p;
  p.seq = seq.p;
  p.qa = qa.p;
  p.qb = qb.p;
  p.vis = vis.p;
  p.qd = qd.p;
  p.spill = spill.p;
  p.tree_size = tsize.p;
  p.label_count = lcount.p;
  p.touched = touched.p;
  p.ctr = ctr.p;
  BcTrees t;
  memset(&t, 0, sizeof(t));
  t.cell = cell.p;
  t.vis = vis.p;
  t.touched = touched.p;
  t.parent = parent.p;
  t.acc = reinterpret_cast<unsigned long long*>(acc.p);
  t.leaf = leaf.p;
  t.nleaf = nleaf.p;
  t.ioff = dg.off[ir].p;
  t.iarc = dg.arc[ir].p;
  t.inv = inv.p;
  t.score = reinterpret_cast<unsigned long long*>(score_d.p);
  t.n = (uint32_t)n;
  // large graphs: a workgroup of 1024 threads per search
  const int tpr = n >= (int64_t)1 << 20 ? 1024 : 256;
  CHL_HIP(hipEventRecord(e0, stream));
  for (uint32_t r0 = 0; r0 < R; r0 += slots) {
    const uint32_t S = std::min<uint32_t>(slots, R - r0);
    p.roots = droots.p + r0;
    p.R = S;
    t.roots = droots.p + r0;
    t.S = S;
    CHL_HIP(launch_fill_u64((uint64_t*)ctr.p, 0, CTR_COUNT, stream));
    CHL_HIP(launch_fill_u32(nleaf.p, 0, S, stream));
    CHL_HIP(launch_tree(p, TREE_KEEP, S, tpr, stream));
    CHL_HIP(launch_bc_trees(t, (unsigned)n, stream));
    // the slots back to their unreached state for the next group
    CHL_HIP(launch_init_cells(cell.p, (size_t)S * n, zero.p, (uint32_t)n, stream));
  }

  // order: stable ascending sort of ~score over ids in ascending order
  CHL_HIP(key.reserve(n));
  CHL_HIP(val.reserve(n));
  CHL_HIP(skey.reserve(n));
  CHL_HIP(sval.reserve(n));
  CHL_HIP(launch_score_keys(t.score, inv.p, (uint32_t)n, key.p, val.p, stream));
  size_t tb = 0;
  CHL_HIP(sort_pairs_u64(nullptr, tb, key.p, skey.p, val.p, sval.p, (uint64_t)n, 64, stream));
  CHL_HIP(tmp.reserve(tb));
  CHL_HIP(sort_pairs_u64(tmp.p, tb, key.p, skey.p, val.p, sval.p, (uint64_t)n, 64, stream));
  CHL_HIP(hipEventRecord(e1, stream));

  std::vector<uint64_t> hv(n);
  CHL_HIP(hipMemcpyAsync(hv.data(), sval.p, sizeof(uint64_t) * n, hipMemcpyDeviceToHost, stream));
  std::vector<uint64_t> hs;
  if (score) {
    hs.resize(n);
    CHL_HIP(hipMemcpyAsync(hs.data(), score_d.p, sizeof(uint64_t) * n, hipMemcpyDeviceToHost, stream));
  }
  CHL_HIP(hipStreamSynchronize(stream));
  for (int64_t i = 0; i < n; ++i) order[i] = (uint32_t)hv[i];
  if (score)
    for (int64_t v = 0; v < n; ++v) (*score)[v] = hs[dg.pos[v]];
  float ms = 0;
  CHL_HIP(hipEventElapsedTime(&ms, e0, e1));
  if (device_ms) *device_ms = ms;
  return CHL_OK;
}

}  // namespace

// Degree order (distributed_memory/degree.cpp:62-67 with reference ties): the
// degrees come from the two arc-offset arrays, one stable radix sort orders
// (degree descending, id descending); with seeded ties a first sort orders the
// ids by their seeded key (equal keys keep ids ascending) and a second, stable
// sort by degree keeps that order among equal degrees.
int order_degree_device(const Graph& g, int device, uint64_t seed, bool reference_ties, std::vector<uint32_t>& order,
                        double* device_ms) {
  const int64_t n = g.n;
  order.resize(n);
  if (device_ms) *device_ms = 0;
  if (n == 0) return CHL_OK;
  if (n >= 0x7FFFFFFFll) return set_error(CHL_E_RANGE, "more than 2^31 - 1 vertices");
  if (g.off[0][n] + g.off[1][n] >= 0xFFFFFFFFull) return set_error(CHL_E_RANGE, "2^32 or more arcs");
  int rc = ensure_device(device);
  if (rc) return rc;
  hipStream_t stream = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  CHL_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  struct Guard {
    hipStream_t& s;
    hipEvent_t &a, &b;
    ~Guard() {
      if (a) (void)hipEventDestroy(a);
      if (b) (void)hipEventDestroy(b);
      if (s) (void)hipStreamDestroy(s);
    }
  } guard{stream, e0, e1};
  CHL_HIP(hipEventCreate(&e0));
  CHL_HIP(hipEventCreate(&e1));
  DBuf<uint64_t> off1, off0, key, val, skey, sval;
  DBuf<uint8_t> tmp;
  // declared after the buffers, so it runs first on every return: queued work
  // ends before the buffers go back to the allocation cache
  StreamDrain drain{stream};
  CHL_HIP(off1.reserve(n + 1));
  CHL_HIP(off0.reserve(n + 1));
  CHL_HIP(key.reserve(n));
  CHL_HIP(val.reserve(n));
  CHL_HIP(skey.reserve(n));
  CHL_HIP(sval.reserve(n));
  CHL_HIP(hipMemcpyAsync(off1.p, g.off[1].data(), sizeof(uint64_t) * (n + 1), hipMemcpyHostToDevice, stream));
  CHL_HIP(hipMemcpyAsync(off0.p, g.off[0].data(), sizeof(uint64_t) * (n + 1), hipMemcpyHostToDevice, stream));
  size_t tb = 0, tb2 = 0;
  CHL_HIP(sort_pairs_u64(nullptr, tb, key.p, skey.p, val.p, sval.p, (uint64_t)n, 64, stream));
  CHL_HIP(sort_pairs_u64(nullptr, tb2, key.p, skey.p, sval.p, val.p, (uint64_t)n, 32, stream));
  CHL_HIP(tmp.reserve(std::max<size_t>({tb, tb2, 1})));
  CHL_HIP(hipEventRecord(e0, stream));
  CHL_HIP(launch_degree_keys(off1.p, off0.p, (uint32_t)n, seed, reference_ties ? 1 : 0, key.p, val.p, stream));
  CHL_HIP(sort_pairs_u64(tmp.p, tb, key.p, skey.p, val.p, sval.p, (uint64_t)n, 64, stream));
  const uint64_t* result = sval.p;
  if (!reference_ties) {
    // second pass: degree keys of the tie-ordered ids, sorted stably (32 bits)
    CHL_HIP(launch_degree_rekey(off1.p, off0.p, sval.p, (uint32_t)n, key.p, stream));
    CHL_HIP(sort_pairs_u64(tmp.p, tb2, key.p, skey.p, sval.p, val.p, (uint64_t)n, 32, stream));
    result = val.p;
  }
  CHL_HIP(hipEventRecord(e1, stream));
  std::vector<uint64_t> hv(n);
  CHL_HIP(hipMemcpyAsync(hv.data(), result, sizeof(uint64_t) * n, hipMemcpyDeviceToHost, stream));
  CHL_HIP(hipStreamSynchronize(stream));
  for (int64_t i = 0; i < n; ++i) order[i] = (uint32_t)hv[i];
  float ms = 0;
  CHL_HIP(hipEventElapsedTime(&ms, e0, e1));
  if (device_ms) *device_ms = ms;
  return CHL_OK;
}

int order_betweenness_device(const Graph& g, DeviceGraph& dg, int device, int64_t roots, uint64_t seed,
                             std::vector<uint32_t>& order, std::vector<uint64_t>* score, double* device_ms) {
  size_t scratch = 0;
  const int rc = order_impl(g, dg, device, roots, seed, order, score, device_ms, &scratch);
  // large scratch goes back to the device rather than to the allocation cache,
  // so that a following build on a large graph sees it as free
  size_t f = 0, t = 0;
  if (hipMemGetInfo(&f, &t) == hipSuccess && scratch > t / 16) dev_cache_flush();
  return rc;
}

}  // namespace chl
