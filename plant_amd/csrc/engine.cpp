// Label construction on one GPU: the drop-in replacement of run_paraPLL
// (shared_memory/akiba.hpp:132-217) and of the hybrid run
// (distributed_memory/akiba.hpp:190-510) at one rank, plus the label handle and
// the batched distance query (shared_memory/labeling.hpp:56-65,
// distributed_memory/query_funcs.hpp:23-30).
//
// Device layout (all in HBM for the whole build):
//   * graph in device ids, a breadth-first locality order kept with the graph
//     handle across builds; per arc range an offset array u32[n+1] and packed
//     arcs u64[A] = (head << 32) | length, plus the rank of every device id
//     u32[n] for the current label build (copied into the search cells). A
//     symmetric graph keeps one range, since both trees of a root are then the
//     same.
//   * label store per side, in rank ids: a list of segments, one per batch
//     (merged as they accumulate), each a CSR over rank rows with 5-byte
//     entries (8 bytes where a distance or the hub span does not fit); a
//     symmetric graph stores side 0 only (side 1 is identical).
//   * per-slot search scratch (tree.hip), S slots x 36 n bytes.
//   * batch output: row u32 and (hub << 32) | dist u64 per label (12 bytes);
//     counted per row, scattered into the batch's new segment, rows sorted,
//     cleaned (GLL) and appended to the store as one segment: no existing row
//     moves.
//
// A batch is a contiguous rank range [r0, r1). Its trees read only labels of
// hubs < r0, which are final, so the labels do not depend on how roots are
// spread over workgroups, on the batch width or on the bucket width.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <random>
#include <string>
#include <vector>

#include "dev.hpp"
#include "engine.hpp"
#include "internal.hpp"

using namespace chl;

namespace chl {

static double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int ensure_device(int device) {
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count == 0) return set_error(CHL_E_DEVICE, "no HIP device visible");
  if (device < 0 || device >= count) return set_error(CHL_E_DEVICE, "device %d out of range (%d visible)", device, count);
  CHL_HIP(hipSetDevice(device));
  return CHL_OK;
}

// Locality order of the vertices, so that a search front touches few cache
// lines of the per-vertex state. First a breadth-first order over the arcs of
// both ranges from a vertex of maximum degree (then from the next unvisited
// vertex by id); then clusters of up to kLocalityCluster vertices, each grown
// breadth-first from the first unassigned vertex of that order, take
// consecutive ids: a tree that enters a cluster finds its neighbours' states on
// the same lines. inv[id] = vertex.
constexpr size_t kLocalityCluster = 64;

static void locality_order(const Graph& g, std::vector<uint32_t>& inv) {
  const int64_t n = g.n;
  std::vector<uint32_t> bfs;
  bfs.reserve(n);
  std::vector<uint8_t> seen(n, 0);
  int64_t best = 0;
  uint64_t bdeg = 0;
  for (int64_t v = 0; v < n; ++v) {
    const uint64_t d = (g.off[1][v + 1] - g.off[1][v]) + (g.off[0][v + 1] - g.off[0][v]);
    if (d > bdeg) {
      bdeg = d;
      best = v;
    }
  }
  for (int64_t k = -1; k < n; ++k) {
    const uint32_t s = (uint32_t)(k < 0 ? best : k);
    if (seen[s]) continue;
    seen[s] = 1;
    size_t head = bfs.size();
    bfs.push_back(s);
    while (head < bfs.size()) {
      const uint32_t u = bfs[head++];
      for (int rg = 1; rg >= 0; --rg)
        for (uint64_t a = g.off[rg][u]; a < g.off[rg][u + 1]; ++a) {
          const uint32_t v = g.head[rg][a];
          if (!seen[v]) {
            seen[v] = 1;
            bfs.push_back(v);
          }
        }
    }
  }
  inv.clear();
  inv.reserve(n);
  std::fill(seen.begin(), seen.end(), 0);
  std::vector<uint32_t> q;
  for (const uint32_t s : bfs) {
    if (seen[s]) continue;
    q.clear();
    q.push_back(s);
    seen[s] = 1;
    for (size_t head = 0; head < q.size() && q.size() < kLocalityCluster; ++head) {
      const uint32_t u = q[head];
      for (int rg = 1; rg >= 0 && q.size() < kLocalityCluster; --rg)
        for (uint64_t a = g.off[rg][u]; a < g.off[rg][u + 1] && q.size() < kLocalityCluster; ++a) {
          const uint32_t v = g.head[rg][a];
          if (!seen[v]) {
            seen[v] = 1;
            q.push_back(v);
          }
        }
    }
    inv.insert(inv.end(), q.begin(), q.end());
  }
}

// CSR of one arc range in device ids.
static void relabel_range(const Graph& g, int range, const std::vector<uint32_t>& inv, const std::vector<uint32_t>& pos,
                          std::vector<uint32_t>& off, std::vector<uint64_t>& arc) {
  const int64_t n = g.n;
  off.assign(n + 1, 0);
  for (int64_t i = 0; i < n; ++i) {
    uint32_t v = inv[i];
    off[i + 1] = off[i] + (uint32_t)(g.off[range][v + 1] - g.off[range][v]);
  }
  arc.resize(off[n]);
  parallel_for(n, host_threads(), 4096, [&](int64_t i) {
    uint32_t v = inv[i];
    uint64_t o = off[i];
    for (uint64_t a = g.off[range][v]; a < g.off[range][v + 1]; ++a)
      arc[o++] = ((uint64_t)pos[g.head[range][a]] << 32) | (uint32_t)g.len[range][a];
  });
}

int DeviceGraph::build(const Graph& g, int dev) {
  device = dev;
  n = g.n;
  sym = g.symmetric;
  locality_order(g, inv);
  pos.assign(n, 0);
  for (int64_t i = 0; i < n; ++i) pos[inv[i]] = (uint32_t)i;
  for (int r = 0; r < 2; ++r) {
    if (sym && r == 0) continue;  // in-arcs equal out-arcs
    if (g.off[r][n] >= 0xFFFFFFFFull) {
      device = -1;
      return set_error(CHL_E_RANGE, "more than 2^32 arcs in a range");
    }
    std::vector<uint32_t> o;
    std::vector<uint64_t> a;
    relabel_range(g, r, inv, pos, o, a);
    CHL_HIP(off[r].reserve(n + 1));
    CHL_HIP(arc[r].reserve(std::max<size_t>(a.size(), 1)));
    CHL_HIP(hipMemcpy(off[r].p, o.data(), sizeof(uint32_t) * (n + 1), hipMemcpyHostToDevice));
    if (!a.empty()) CHL_HIP(hipMemcpy(arc[r].p, a.data(), sizeof(uint64_t) * a.size(), hipMemcpyHostToDevice));
  }
  return CHL_OK;
}

// ---------------------------------------------------------------------------
// Builder
// ---------------------------------------------------------------------------
struct SideBatch {
  DBuf<uint32_t> row;  // tree output: row id per label
  DBuf<uint64_t> hd;   // (hub << 32) | dist per label
  DBuf<uint8_t> keep;  // GLL cleaning flags
  std::unique_ptr<Segment> seg;  // the batch's new segment
  uint64_t m = 0;       // labels emitted (before cleaning)
  uint64_t kept = 0;    // labels appended
  uint32_t maxd = 0;    // largest label distance
};

// Default search slots: 2048 workgroups of 256 threads, eight wavefronts per
// SIMD on a 256-CU device (the tree kernel uses 32 VGPRs). The searches are
// bound by memory latency, not bandwidth: on the 1M-vertex road-like graph
// 512, 1024 and 2048 slots took 1122, 934 and 875 ms of tree time per build
// while their fetched bytes rose from 1.80e12 to 2.54e12.
constexpr uint32_t kDefaultSlots = 2048;
// Below this many slots (memory-limited: a large graph) the default
// granularity gives large trees a workgroup of 1024 threads.
constexpr uint32_t kWideBelow = 1024;

// Small-tree slots (small.hip): 4096 wavefronts, each with a hash table of
// 32768 entries (28 bytes each: 3.8 GB in all). A root goes there when its
// predicted tree processes at most kSmallRoute vertices, an eighth of the
// entries a tree may insert before it is handed to the tree kernel (the
// predictions' mean error is a factor of about 2.4).
#ifndef CHL_SMALL_CAP
#define CHL_SMALL_CAP 32768
#endif
#ifndef CHL_SMALL_ROUTE_DIV
#define CHL_SMALL_ROUTE_DIV 16
#endif
constexpr uint32_t kSmallSlots = 4096;
constexpr uint32_t kSmallCap = CHL_SMALL_CAP;
constexpr uint32_t kSmallRoute = kSmallCap / CHL_SMALL_ROUTE_DIV;

// Medium-tree slots (small.hip, 256 threads per root; build option
// medium_trees): kMediumSlots workgroups, each with a hash table of kMediumCap
// entries (28 bytes each: 7.5 GB in all), for the roots the small slots leave
// whose predicted tree processes at most kMediumRoute vertices; a medium pass
// runs when at least kMediumMin such roots are in the batch. Off by default: on
// the 16M-vertex road-like graph a medium pass costs 0.4-0.8 ns per processed
// vertex against 0.25-0.38 ns in the tree kernel, and the tree kernel's pass
// after it, bound by the batch's largest trees, hardly shortens (learned build
// 105.5 s against 98.4 s without).
#ifndef CHL_MEDIUM_CAP
#define CHL_MEDIUM_CAP 262144
#endif
#ifndef CHL_MEDIUM_SLOTS
#define CHL_MEDIUM_SLOTS 1024
#endif
#ifndef CHL_MEDIUM_ROUTE_DIV
#define CHL_MEDIUM_ROUTE_DIV 8
#endif
constexpr uint32_t kMediumSlots = CHL_MEDIUM_SLOTS;
constexpr uint32_t kMediumCap = CHL_MEDIUM_CAP;
constexpr uint32_t kMediumRoute = kMediumCap / CHL_MEDIUM_ROUTE_DIV;
constexpr uint32_t kMediumMin = 256;

// Frontier sorting in the tree kernel (tree.hip): frontiers of at least
// kSortMin entries, in launches of at least kSortWaves wavefronts (half of what
// 256 compute units hold). 1M road-like graph (2048 slots of 256 threads): tree
// time 0.915 -> 0.864 s per learned build; 16M (a few hundred slots, fewer as
// the store grows): sorting in every launch cost 2-16% of the tree time.
#ifndef CHL_SORT_MIN
#define CHL_SORT_MIN 512
#endif
constexpr uint32_t kSortMin = CHL_SORT_MIN;
constexpr uint64_t kSortWaves = 4096;

// Device memory a memory-limited build keeps free beside the search slots (for
// the batch buffers, segment building and the store's appends and merges): a
// kReserveDiv-th of the device.
#ifndef CHL_RESERVE_DIV
#define CHL_RESERVE_DIV 4
#endif
constexpr size_t kReserveDiv = CHL_RESERVE_DIV;

// The default bucket width: kDeltaMeanArcs times the mean arc length.
constexpr double kDeltaMeanArcs = 8.0;
int32_t default_delta(const Graph& g) {
  double sum = 0;
  for (int32_t w : g.len[1]) sum += w;
  const double mean = g.len[1].empty() ? 1.0 : sum / (double)g.len[1].size();
  return (int32_t)std::max(1.0, std::min(1e9, std::floor(kDeltaMeanArcs * mean)));
}

// Search slots for a granularity: base_slots workgroups of 256 threads, as
// many threads in flight at 64 or 1024 threads per root.
static uint32_t slots_for(uint32_t base_slots, uint32_t tpr) {
  return std::max<uint32_t>(1, (uint32_t)((uint64_t)base_slots * 256 / tpr));
}
static int64_t granularity_of(uint32_t tpr) { return tpr == 64 ? 1 : tpr == 1024 ? 3 : 2; }

struct Builder {
  const Graph& g;
  DeviceGraph& dg;
  const std::vector<uint32_t>& order;
  chl_build_opts o;
  chl_build_stats* stats;
  BuildHooks* hooks;
  int64_t n;
  bool sym;
  int nsides;
  hipStream_t stream = nullptr;

  SlotBuf<uint64_t> cell;
  SlotBuf<uint32_t> qa, qb, vis;
  DBuf<uint32_t> gkey, gval, grank, seq;
  SlotBuf<uint64_t> qd;  // deferred piles
  // small-tree slots (small.hip): tables of kSmallCap entries, lists of 3 kSmallCap
  DBuf<uint64_t> sstate;
  DBuf<uint32_t> skey, sstamp, slist;
  DBuf<uint8_t> bigd;
  std::vector<uint8_t> bigh;
  uint32_t sslots = 0;
  // medium-tree slots: tables of kMediumCap entries; bigmh[k] = 1 for batch
  // position k predicted too large for them, bigmd its copy for the current roots
  DBuf<uint64_t> mstate;
  DBuf<uint32_t> mkey, mstamp, mlist;
  DBuf<uint8_t> bigmd;
  std::vector<uint8_t> bigmh, bigm_cur;
  uint32_t mslots = 0;
  std::vector<uint32_t> id_of_rank;  // rank -> device id
  std::vector<uint32_t> hids;        // staging of root ids
  uint32_t slots = 0, gcap = 0;
  uint32_t active = 0;  // slots used by launches (at most slots)
  uint32_t tpr = 256;   // threads per root: 64 (wavefront), 256 or 1024 (workgroups)
  DBuf<uint32_t> roots, roots2, spill, tree_size, label_count, touched, origd;
  uint32_t W = 1;            // widest batch (scratch offset of the statistics arrays)
  int64_t stats_batches = 0;
  DBuf<unsigned long long> ctr;
  DBuf<unsigned long long> misc;  // [0] switch flag, [1] max row
  // segment building: per-row counts, scan scratch, scatter cursors, the list
  // of long rows of the row sort
  DBuf<uint32_t> cnt, longq;
  DBuf<uint64_t> scan_tmp, cur;
  SideBatch sb[2];
  static constexpr double kBatchBytes = 32.0;  // device bytes per batch label in flight
  // mean processed vertices per root above which memory-limited slots run
  // 1024-thread workgroups
#ifndef CHL_WIDE_TREE
#define CHL_WIDE_TREE 65536.0
#endif
  static constexpr double kWideTree = CHL_WIDE_TREE;
  uint64_t out_cap = 0;
  uint64_t next_need = 0;  // output estimate of the next batch (labels)
  uint64_t out_limit = ~0ull;  // output capacity the memory cap allows for the current batch
  double tail_labels = -1;  // labels per root at the top ranks of the last batch
  hipEvent_t ev[8];

  Builder(const Graph& gg, DeviceGraph& d, const std::vector<uint32_t>& ord, const chl_build_opts& opts,
          chl_build_stats* s, BuildHooks* h)
      : g(gg), dg(d), order(ord), o(opts), stats(s), hooks(h), n(gg.n), sym(gg.symmetric) {
    nsides = sym ? 1 : 2;
    for (auto& e : ev) e = nullptr;
  }
  ~Builder() {
    // an error return can leave work queued: let it finish before the member
    // buffers go back to the allocation cache, where another call may take them
    if (stream) (void)hipStreamSynchronize(stream);
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
  }

  // The device graph (built once per graph handle and device) and the current
  // rank of every device id.
  int upload_graph(const std::vector<uint32_t>& rank) {
    if (dg.device != o.device || dg.n != n) {
      int rc = dg.build(g, o.device);
      if (rc) return rc;
    }
    std::vector<uint32_t> rank_of_id(n);
    id_of_rank.resize(n);
    for (int64_t i = 0; i < n; ++i) {
      rank_of_id[i] = rank[dg.inv[i]];
      id_of_rank[rank[dg.inv[i]]] = (uint32_t)i;
    }
    CHL_HIP(grank.reserve(n));
    CHL_HIP(hipMemcpyAsync(grank.p, rank_of_id.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice, stream));
    CHL_HIP(hipStreamSynchronize(stream));
    return CHL_OK;
  }

  // Search slots. Each holds 36 bytes per vertex (16-byte cell, three u32
  // lists and a u64 pile). On a graph where the wanted count does not fit, the
  // slots take the free memory beyond a reserve of a quarter of the device (the
  // label store is still small when the largest trees run, at the top ranks),
  // in arrays whose trailing slots can go back to the device (ChunkBuf), and
  // fit_slots gives slots back as the store grows.
  static constexpr size_t kSlotBytesPerVertex = 36;
  uint32_t slots_wanted = 0;

  int alloc_slots(uint32_t want) {
    size_t freeb = 0, totalb = 0;
    CHL_HIP(hipMemGetInfo(&freeb, &totalb));
    freeb += dev_cached_bytes();  // blocks the allocator keeps are free to reuse
    const size_t per = (size_t)n * kSlotBytesPerVertex;
    const size_t reserve = totalb / kReserveDiv;
    size_t budget = freeb > reserve + totalb / 8 ? freeb - reserve : freeb / 3;
    if (o.slot_memory_mb > 0) budget = std::min(budget, (size_t)o.slot_memory_mb << 20);
    uint32_t fit = (uint32_t)std::max<size_t>(1, budget / std::max<size_t>(per, 1));
    slots_wanted = want;
    slots = std::max<uint32_t>(1, std::min(want, fit));
    const bool limited = slots < want;
    if (limited) dev_cache_flush();  // the chunks come from the device, not the cache
    const size_t cells = (size_t)slots * n;
    // chunks of whole slots, about 512 MB each
    auto chunk = [&](size_t slot_bytes) {
      return slot_bytes * std::max<size_t>(1, ((size_t)512 << 20) / std::max<size_t>(slot_bytes, 1));
    };
    CHL_HIP(cell.alloc(2 * cells, limited, chunk(16 * (size_t)n)));
    CHL_HIP(qa.alloc(cells, limited, chunk(4 * (size_t)n)));
    CHL_HIP(qb.alloc(cells, limited, chunk(4 * (size_t)n)));
    CHL_HIP(vis.alloc(cells, limited, chunk(4 * (size_t)n)));
    CHL_HIP(qd.alloc(cells, limited, chunk(8 * (size_t)n)));
    CHL_HIP(seq.reserve(2 * (size_t)slots));
    CHL_HIP(launch_init_cells(cell.p, cells, grank.p, (uint32_t)n, stream));
    CHL_HIP(launch_fill_u32(seq.p, 0, 2 * (size_t)slots, stream));
    return CHL_OK;
  }

  // Before a batch of a memory-limited build: when the free memory has fallen
  // below four fifths of the reserve (the store grows behind the batches), slots
  // go back until the reserve is free again, keeping at least kMinSlots. The
  // slots that stay keep their memory and contents (cells, counts).
  static constexpr uint32_t kMinSlots = 32;
  int fit_slots() {
    if (slots >= slots_wanted || slots <= kMinSlots || !cell.chunked) return CHL_OK;
    size_t freeb = 0, totalb = 0;
    CHL_HIP(hipMemGetInfo(&freeb, &totalb));
    freeb += dev_cached_bytes();
    if (freeb >= totalb * 4 / (5 * kReserveDiv)) return CHL_OK;
    const size_t per = (size_t)n * kSlotBytesPerVertex;
    const size_t give = (totalb / kReserveDiv - freeb + per - 1) / per;
    const uint32_t keep = (uint32_t)std::max<int64_t>(kMinSlots, (int64_t)slots - (int64_t)give);
    if (keep >= slots) return CHL_OK;
    const double t0 = now_s();
    CHL_HIP(hipStreamSynchronize(stream));
    const size_t cells = (size_t)keep * n;
    CHL_HIP(cell.shrink(2 * cells));
    CHL_HIP(qa.shrink(cells));
    CHL_HIP(qb.shrink(cells));
    CHL_HIP(vis.shrink(cells));
    CHL_HIP(qd.shrink(cells));
    if (o.verbose) fprintf(stderr, "slots: %u -> %u (%.1f GB free before)\n", slots, keep, freeb / 1e9);
    slots = keep;
    active = std::min(active, slots);
    if (hooks) hooks->slots = slots;
    if (stats) {
      stats->slot_shrinks++;
      stats->slot_s += now_s() - t0;
    }
    if (o.verbose) fprintf(stderr, "slots: resized in %.3f s\n", now_s() - t0);
    return CHL_OK;
  }

  int ensure_gtab(uint64_t max_row) {
    uint32_t need = 1;
    while (need < 2 * max_row + 2) need <<= 1;
    if (need <= gcap) return CHL_OK;
    gcap = need;
    CHL_HIP(gkey.reserve((size_t)slots * gcap));
    CHL_HIP(gval.reserve((size_t)slots * gcap));
    return CHL_OK;
  }

  int ensure_small() {
    if (sslots) return CHL_OK;
    const size_t cells = (size_t)kSmallSlots * kSmallCap;
    CHL_HIP(sstate.reserve(cells));
    CHL_HIP(skey.reserve(cells));
    CHL_HIP(sstamp.reserve(cells));
    CHL_HIP(slist.reserve(3 * cells));
    TreeParams p;
    memset(&p, 0, sizeof(p));
    p.sstate = sstate.p;
    p.skey = skey.p;
    p.sstamp = sstamp.p;
    p.scap = kSmallCap;
    CHL_HIP(launch_init_small(p, kSmallSlots, stream));
    sslots = kSmallSlots;
    return CHL_OK;
  }

  int ensure_medium() {
    if (mslots) return CHL_OK;
    const size_t cells = (size_t)kMediumSlots * kMediumCap;
    CHL_HIP(mstate.reserve(cells));
    CHL_HIP(mkey.reserve(cells));
    CHL_HIP(mstamp.reserve(cells));
    CHL_HIP(mlist.reserve(3 * cells));
    TreeParams p;
    memset(&p, 0, sizeof(p));
    p.sstate = mstate.p;
    p.skey = mkey.p;
    p.sstamp = mstamp.p;
    p.scap = kMediumCap;
    CHL_HIP(launch_init_small(p, kMediumSlots, stream));
    mslots = kMediumSlots;
    return CHL_OK;
  }

  int ensure_out(SideBatch& b, uint64_t cap) {
    CHL_HIP(b.row.reserve(cap));
    CHL_HIP(b.hd.reserve(cap));
    return CHL_OK;
  }

  // Grows a batch output buffer to cap entries keeping its first `used` entries.
  template <class T>
  int grow_keep(DBuf<T>& buf, uint64_t cap, uint64_t used) {
    if (buf.n >= cap) return CHL_OK;
    DBuf<T> nb;
    CHL_HIP(nb.reserve(cap));
    if (used) CHL_HIP(hipMemcpyAsync(nb.p, buf.p, sizeof(T) * used, hipMemcpyDeviceToDevice, stream));
    buf.swap(nb);
    CHL_HIP(hipStreamSynchronize(stream));
    return CHL_OK;
  }

  // When the device runs short (the store grows in place behind the batches),
  // batch buffers far larger than the next batch needs go back to the device
  // and so does the allocation cache: the first batches of a large graph grow
  // them to many times what later batches use.
  int trim_batch_buffers(uint64_t need) {
    if (!need) return CHL_OK;
    size_t freeb = 0, totalb = 0;
    CHL_HIP(hipMemGetInfo(&freeb, &totalb));
    if (freeb > totalb / kReserveDiv) return CHL_OK;
    const uint64_t keep = 2 * need + ((uint64_t)1 << 20);
    size_t freed = 0;
    auto trim = [&](auto& buf) {
      if (buf.n > keep) {
        freed += buf.bytes;
        buf.release();
      }
    };
    CHL_HIP(hipStreamSynchronize(stream));
    for (auto& b : sb) {
      trim(b.row);
      trim(b.hd);
      trim(b.keep);
    }
    if (freed) {
      out_cap = std::min<uint64_t>(out_cap, keep);
      dev_cache_flush();
      if (o.verbose) fprintf(stderr, "trim: %.2f GB of batch buffers back to the device\n", freed / 1e9);
    }
    return CHL_OK;
  }

  // Runs the trees of roots[0..R) for one side into the side's output buffer;
  // b.m = labels written. Roots whose labels did not fit, and the roots not
  // started once the buffer was full, run in a further launch after the buffer
  // has grown by what the former reported and an estimate for the latter.
  // small_pass: the first launch runs the small-tree kernel over the batch, with
  // bigd[k] = 1 for roots it leaves to later launches (they, and trees that
  // outgrow a small slot, go to the spill list like roots not started).
  // medium_pass: the next launch runs the medium-tree kernel over the roots left
  // (bigmh by batch position: those it leaves to the tree kernel).
  int run_trees(int side, uint32_t R, const uint32_t* hroots, uint32_t prefix, uint32_t flags, LabelStore& L,
                double& ms, bool small_pass = false, bool medium_pass = false) {
    SideBatch& b = sb[side];
    const int range = side == 0 ? 1 : 0;  // forward tree walks out-arcs and fills side 0
    int rc = ensure_out(b, out_cap);
    if (rc) return rc;
    TreeParams p;
    memset(&p, 0, sizeof(p));
    p.off = dg.off[sym ? 1 : range].p;
    p.arc = dg.arc[sym ? 1 : range].p;
    p.n = (uint32_t)n;
    p.scan_view = L.d_views(side);
    p.root_view = L.d_views(1 - side);
    p.scan_nseg = p.root_nseg = L.nseg();
    p.prefix = prefix;
    p.roots = roots.p;
    p.R = R;
    p.delta = (uint32_t)o.delta;
    p.rank = grank.p;
    p.cell = cell.p;
    p.seq = seq.p;
    p.qa = qa.p;
    p.qb = qb.p;
    p.vis = vis.p;
    p.qd = qd.p;
    p.gkey = gkey.p;
    p.gval = gval.p;
    p.gcap = gcap;
    p.out_row = b.row.p;
    p.out_hd = b.hd.p;
    p.out_cap = std::min<uint64_t>(b.row.n, b.hd.n);
    p.spill = spill.p;
    p.tree_size = tree_size.p;
    p.touched = o.verbose ? touched.p : nullptr;
    p.label_count = label_count.p;
    p.ctr = ctr.p;
    if (small_pass) {
      p.sstate = sstate.p;
      p.skey = skey.p;
      p.sstamp = sstamp.p;
      p.slist = slist.p;
      p.scap = kSmallCap;
      p.big = bigd.p;
    }
    CHL_HIP(launch_fill_u64((uint64_t*)ctr.p, 0, CTR_COUNT, stream));
    CHL_HIP(launch_fill_u64((uint64_t*)ctr.p + CTR_SPILL_START, ~0ull, 1, stream));
    std::vector<uint32_t> cur(hroots, hroots + R), idx, next, orig(R), next_orig;
    for (uint32_t j = 0; j < R; ++j) orig[j] = j;  // position of each current root in the batch
    unsigned long long hc[CTR_COUNT];
    uint64_t lab0 = 0;  // label counter at the start of the pass
    const int medium_at = medium_pass ? (small_pass ? 1 : 0) : -1;  // the medium launch's pass
    for (int pass = 0;; ++pass) {
      if (pass == medium_at) {
        // the medium tables, and their big flags for the roots of this pass
        bigm_cur.resize(p.R);
        for (uint32_t j = 0; j < p.R; ++j) bigm_cur[j] = bigmh[orig[j]];
        CHL_HIP(bigmd.reserve(std::max<size_t>(p.R, 1)));
        CHL_HIP(hipMemcpyAsync(bigmd.p, bigm_cur.data(), p.R, hipMemcpyHostToDevice, stream));
        p.sstate = mstate.p;
        p.skey = mkey.p;
        p.sstamp = mstamp.p;
        p.slist = mlist.p;
        p.scap = kMediumCap;
        p.big = bigmd.p;
      }
      CHL_HIP(hipEventRecord(ev[0], stream));
      if (pass == 0 && small_pass) CHL_HIP(launch_small_trees(p, std::min<uint32_t>(sslots, p.R), 64, stream));
      else if (pass == medium_at) CHL_HIP(launch_small_trees(p, std::min<uint32_t>(mslots, p.R), 256, stream));
      else {
        // sorting the frontiers pays when the launch keeps the memory system busy
        // (many trees in flight), not when a few slots wait on latency
        const uint32_t wgs = std::min<uint32_t>(active, p.R);
        p.sort_min = (uint64_t)wgs * tpr / 64 >= kSortWaves ? kSortMin : 0;
        CHL_HIP(launch_tree(p, flags, wgs, (int)tpr, stream));
      }
      CHL_HIP(hipEventRecord(ev[1], stream));
      if (pass > 0) {
        // this pass's statistics to the roots' positions in the batch
        CHL_HIP(launch_scatter_u32(p.tree_size, origd.p, p.R, tree_size.p, stream));
        CHL_HIP(launch_scatter_u32(p.label_count, origd.p, p.R, label_count.p, stream));
        if (p.touched) CHL_HIP(launch_scatter_u32(p.touched, origd.p, p.R, touched.p, stream));
      }
      CHL_HIP(hipMemcpyAsync(hc, ctr.p, sizeof(hc), hipMemcpyDeviceToHost, stream));
      CHL_HIP(hipStreamSynchronize(stream));
      float t = 0;
      CHL_HIP(hipEventElapsedTime(&t, ev[0], ev[1]));
      ms += t;
      if (stats) {
        stats->relax_ms += t;
        stats->tree_launches++;
      }
      if (o.verbose)
        fprintf(stderr, "  side %d pass %d: %u roots, %.3f ms, written %llu, spilled %llu (%llu labels, %llu not started), processed %llu arcs %llu rounds %llu\n", side, pass, p.R, t, hc[CTR_LABELS], hc[CTR_OVERFLOW], hc[CTR_SPILL], hc[CTR_SKIPPED], hc[CTR_PROCESSED], hc[CTR_ARCS], hc[CTR_ROUNDS]);
      if (pass == 0 && small_pass && stats) stats->small_roots += (int64_t)(p.R - hc[CTR_OVERFLOW]);
      if (pass == medium_at && stats) stats->medium_roots += (int64_t)(p.R - hc[CTR_OVERFLOW]);
      if (!hc[CTR_OVERFLOW]) break;
      // next launch for the roots that did not fit or were not started
      const uint32_t ns = (uint32_t)hc[CTR_OVERFLOW];
      const uint64_t skipped = hc[CTR_SKIPPED];
      idx.resize(ns);
      CHL_HIP(hipMemcpy(idx.data(), spill.p, sizeof(uint32_t) * ns, hipMemcpyDeviceToHost));
      next.resize(ns);
      next_orig.resize(ns);
      for (uint32_t j = 0; j < ns; ++j) {
        next[j] = cur[idx[j]];
        next_orig[j] = orig[idx[j]];
      }
      cur.swap(next);
      orig.swap(next_orig);
      // contiguous prefix of fitting roots (all labels written when only roots
      // not started, such as the small-tree kernel's hand-overs, are left)
      const uint64_t written = std::min<uint64_t>(hc[CTR_SPILL_START], hc[CTR_LABELS]);
      // labels still to write: exact for the roots that ran, and for the roots
      // not started 1.5 times the mean of the roots that ran in this pass
      // (bounded by a 48th of the free memory: key and entry now, their sorted
      // copies and the store append later; a short estimate costs one more pass)
      const uint64_t ran = p.R - skipped;
      const double mean = ran ? (double)(hc[CTR_LABELS] - lab0) / (double)ran : 0.0;
      size_t freeb = 0, totalb = 0;
      CHL_HIP(hipMemGetInfo(&freeb, &totalb));
      const double est = std::min((double)freeb / 48.0, 1.5 * mean * (double)skipped);
      const uint64_t need = written + hc[CTR_SPILL] + (uint64_t)est + 4096;
      out_cap = std::max<uint64_t>(out_cap, need + need / 8);
      rc = grow_keep(b.row, out_cap, written);
      if (rc) return rc;
      rc = grow_keep(b.hd, out_cap, written);
      if (rc) return rc;
      hids.resize(ns);
      for (uint32_t j = 0; j < ns; ++j) hids[j] = id_of_rank[cur[j]];
      CHL_HIP(hipMemcpy(roots2.p, hids.data(), sizeof(uint32_t) * ns, hipMemcpyHostToDevice));
      CHL_HIP(hipMemcpy(origd.p, orig.data(), sizeof(uint32_t) * ns, hipMemcpyHostToDevice));
      p.roots = roots2.p;
      p.R = ns;
      p.out_row = b.row.p;
      p.out_hd = b.hd.p;
      p.out_cap = out_cap;
      p.tree_size = tree_size.p + W;  // scratch, scattered back after the launch
      p.label_count = label_count.p + W;
      if (p.touched) p.touched = touched.p + W;
      lab0 = written;
      CHL_HIP(launch_fill_u64((uint64_t*)ctr.p + CTR_LABELS, written, 1, stream));
      CHL_HIP(launch_fill_u64((uint64_t*)ctr.p + CTR_OVERFLOW, 0, 3, stream));  // OVERFLOW, NEXT, SPILL
      CHL_HIP(launch_fill_u64((uint64_t*)ctr.p + CTR_SPILL_START, ~0ull, 1, stream));
      CHL_HIP(launch_fill_u64((uint64_t*)ctr.p + CTR_SKIPPED, 0, 1, stream));
    }
    if (o.verbose >= 2 && touched.p) {
      // per-root distributions of processed and reached vertices
      std::vector<uint32_t> a(R), c(R);
      CHL_HIP(hipMemcpy(a.data(), tree_size.p, sizeof(uint32_t) * R, hipMemcpyDeviceToHost));
      CHL_HIP(hipMemcpy(c.data(), touched.p, sizeof(uint32_t) * R, hipMemcpyDeviceToHost));
      double wsum = 0, wfit[4] = {0, 0, 0, 0};
      size_t fit[4] = {0, 0, 0, 0};
      for (uint32_t k = 0; k < R; ++k) {
        wsum += a[k];
        for (int q = 0; q < 4; ++q)
          if (c[k] <= (2048u << q)) {
            ++fit[q];
            wfit[q] += a[k];
          }
      }
      std::sort(a.begin(), a.end());
      std::sort(c.begin(), c.end());
      auto pct = [&](const std::vector<uint32_t>& x, double f) { return x[(size_t)(f * (x.size() - 1))]; };
      fprintf(stderr,
              "    processed p50 %u p90 %u p99 %u max %u | reached p50 %u p90 %u p99 %u max %u | roots reaching "
              "<= 2k/4k/8k/16k: %.3f %.3f %.3f %.3f (work %.3f %.3f %.3f %.3f)\n",
              pct(a, .5), pct(a, .9), pct(a, .99), a.back(), pct(c, .5), pct(c, .9), pct(c, .99), c.back(),
              (double)fit[0] / R, (double)fit[1] / R, (double)fit[2] / R, (double)fit[3] / R, wfit[0] / wsum,
              wfit[1] / wsum, wfit[2] / wsum, wfit[3] / wsum);
    }
    b.m = hc[CTR_LABELS];
    b.maxd = (uint32_t)hc[CTR_MAXD];
    if (stats) {
      stats->settled += (int64_t)hc[CTR_PROCESSED];
      stats->arcs_scanned += (int64_t)hc[CTR_ARCS];
      stats->settled_distinct += (int64_t)hc[CTR_SETTLED];
      stats->settled_arcs += (int64_t)hc[CTR_SETTLED_ARCS];
      stats->prune_entries += (int64_t)hc[CTR_PRUNE_ENT];
      stats->candidates += (int64_t)hc[CTR_LABELS];
      // the minimal work of the trees: per distinct settled vertex its row
      // bounds and state, per arc it scans (once) the arc and the head's
      // state, per prune probe one label entry, per label its entry
      stats->relax_bytes += 16.0 * hc[CTR_SETTLED] + 16.0 * hc[CTR_SETTLED_ARCS] + 8.0 * hc[CTR_PRUNE_ENT] +
                            8.0 * hc[CTR_LABELS];
      // the same counted over every processing, re-processing included
      stats->relax_bytes_processed +=
          16.0 * hc[CTR_PROCESSED] + 16.0 * hc[CTR_ARCS] + 8.0 * hc[CTR_PRUNE_ENT] + 8.0 * hc[CTR_LABELS];
    }
    if (hooks && hooks->on_tree_launch) hooks->on_tree_launch(hooks->ctx, side, R, (float)ms, hc);
    return CHL_OK;
  }

  // Autotuner: times the PLaNT trees of a fixed sample of roots (ranks
  // log-uniform over [max(1024, n/64), n); labels discarded) under each candidate
  // per-root granularity (workgroup, wavefront) at the default bucket width,
  // then the better one at half and twice it, and keeps the fastest. Labels depend
  // on neither. The wide workgroup (1024 threads) is not a candidate: on a
  // sample it wins by finishing the largest sampled tree sooner, while over the
  // batches of a whole build (many roots per slot) it was slower in every batch
  // but the first on the 1M-vertex road-like graph.
  int autotune(LabelStore& L, double& lt, uint32_t base_slots) {
    const double t0 = now_s();
    std::mt19937_64 rng(0x5EEDu);
    const int64_t lo = std::min<int64_t>(std::max<int64_t>(1024, n / 64), n / 2);
    const uint32_t R = (uint32_t)std::min<int64_t>(std::min<uint32_t>(W, 2 * base_slots), n - lo);
    std::vector<uint32_t> sample(R);
    std::uniform_real_distribution<double> u(std::log((double)lo + 1.0), std::log((double)n));
    for (auto& r : sample) r = (uint32_t)std::min<int64_t>(n - 1, std::max<int64_t>(lo, (int64_t)std::exp(u(rng)) - 1));
    hids.resize(R);
    for (uint32_t k = 0; k < R; ++k) hids[k] = id_of_rank[sample[k]];
    CHL_HIP(hipMemcpyAsync(roots.p, hids.data(), sizeof(uint32_t) * R, hipMemcpyHostToDevice, stream));
    const int32_t base_delta = o.delta;
    double best = 1e300;
    uint32_t best_tpr = tpr;
    int32_t best_delta = o.delta;
    uint32_t best_active = active;
    double warm = 0;
    int rc = run_trees(0, R, sample.data(), 0, TREE_FLAGS, L, warm);  // first launch of the kernel
    if (rc) return rc;
    lt += warm;
    std::vector<std::pair<uint32_t, int32_t>> cand = {{256u, 4}, {64u, 4}};  // (threads, bucket width in quarters)
    for (size_t c = 0; c < cand.size(); ++c) {
      tpr = cand[c].first;
      active = std::min<uint32_t>(slots, slots_for(base_slots, tpr));
      o.delta = std::max<int32_t>(1, (int32_t)((int64_t)base_delta * cand[c].second / 4));
      double ms = 0;
      rc = run_trees(0, R, sample.data(), 0, TREE_FLAGS, L, ms);
      if (rc) return rc;
      lt += ms;
      if (o.verbose) fprintf(stderr, "autotune: %u threads per root, delta %d slots %u: %.3f ms\n", tpr, o.delta,
                             active, ms);
      if (ms < best) {
        best = ms;
        best_tpr = tpr;
        best_delta = o.delta;
        best_active = active;
      }
      if (c == 1) {
        // second stage: the bucket width at the best granularity
        cand.push_back({best_tpr, 2});
        cand.push_back({best_tpr, 8});
      }
    }
    tpr = best_tpr;
    o.delta = best_delta;
    active = best_active;
    if (stats) {
      stats->autotune_ms = (now_s() - t0) * 1e3;
      stats->delta_used = o.delta;
      stats->slots_used = active;
      stats->granularity_used = granularity_of(tpr);
    }
    return CHL_OK;
  }

  // ms += the device time between two events recorded on the stream.
  int elapsed(hipEvent_t a, hipEvent_t b, double& ms) {
    CHL_HIP(hipEventSynchronize(b));
    float t = 0;
    CHL_HIP(hipEventElapsedTime(&t, a, b));
    ms += t;
    if (stats) stats->segment_ms += t;
    return CHL_OK;
  }

  // The batch's labels of one side into a new segment for the hub range
  // [r0, r0 + R): counted per row, offsets by a scan, scattered into the rows,
  // rows sorted by hub. O(batch): no stored row moves. SEG_F5 unless a distance
  // exceeds 65535 or the range is wider than 2^24 ranks.
  int build_segment(int side, uint32_t r0, uint32_t R, double& ms) {
    SideBatch& b = sb[side];
    const uint64_t m = b.m;
    CHL_HIP(hipEventRecord(ev[2], stream));
    std::unique_ptr<Segment> g(new Segment());
    g->h0 = r0;
    g->h1 = r0 + R;
    const bool f5 = b.maxd <= kF5MaxDist && R <= kF5MaxWidth;
    int rc = g->alloc((uint64_t)n, m, f5 ? SEG_F5 : SEG_F8);
    if (rc) return rc;
    CHL_HIP(launch_fill_u32(cnt.p, 0, (size_t)n, stream));
    CHL_HIP(launch_count_rows(b.row.p, m, cnt.p, stream));
    CHL_HIP(launch_scan_u32(cnt.p, (uint64_t)n, g->off.p, scan_tmp.p, stream));
    if (m) {
      CHL_HIP(launch_copy_u64(g->off.p, (uint64_t)n, cur.p, stream));
      CHL_HIP(launch_scatter_labels(b.row.p, b.hd.p, m, cur.p, g->fmt, g->h0, g->key.p, g->lo.p, g->ent.p, stream));
      CHL_HIP(launch_sort_rows(g->off.p, (uint32_t)n, g->fmt, g->key.p, g->lo.p, g->ent.p, longq.p, stream));
    }
    CHL_HIP(hipEventRecord(ev[3], stream));
    b.seg = std::move(g);
    return elapsed(ev[2], ev[3], ms);
  }

  // GLL cleaning of one side's new segment against the batch labels of both
  // sides (clean_cover over the superstep's labels, shared_memory/akiba.hpp:
  // 197-208 with labeling.hpp:78-88).
  int clean_side(int side, double& ms) {
    SideBatch& b = sb[side];
    SideBatch& ob = sb[sym ? side : 1 - side];
    const uint64_t m = b.m;
    if (!m) return CHL_OK;
    CHL_HIP(b.keep.reserve(m));
    CHL_HIP(launch_fill_u64((uint64_t*)ctr.p + CTR_PRUNE_ENT, 0, 1, stream));
    CHL_HIP(hipEventRecord(ev[4], stream));
    CHL_HIP(launch_clean(b.seg->view(), ob.seg->view(), (uint32_t)n, b.keep.p, ctr.p, stream));
    CHL_HIP(hipEventRecord(ev[5], stream));
    double t = 0;
    int rc = elapsed(ev[4], ev[5], t);
    if (rc) return rc;
    ms += t;
    if (stats) {
      // per candidate: its entry, its flag and the two offsets of the other
      // row; per entry of the other rows read: the entry
      unsigned long long reads = 0;
      CHL_HIP(hipMemcpy(&reads, ctr.p + CTR_PRUNE_ENT, sizeof(reads), hipMemcpyDeviceToHost));
      const double eb = b.seg->fmt == SEG_F8 ? 8.0 : 5.0, oeb = ob.seg->fmt == SEG_F8 ? 8.0 : 5.0;
      stats->clean_entries += (int64_t)reads;
      stats->clean_bytes += (eb + 1.0 + 16.0) * (double)m + oeb * (double)reads;
      stats->clean_kernel_ms += t;
    }
    return CHL_OK;
  }

  // Drops the covered candidates: a new segment of the kept entries.
  int compact_side(int side, double& ms) {
    SideBatch& b = sb[side];
    if (!b.m) return CHL_OK;
    CHL_HIP(hipEventRecord(ev[4], stream));
    const Segment& o = *b.seg;
    CHL_HIP(launch_count_kept(o.off.p, b.keep.p, (uint32_t)n, cnt.p, stream));
    CHL_HIP(launch_scan_u32(cnt.p, (uint64_t)n, cur.p, scan_tmp.p, stream));
    uint64_t kept = 0;
    CHL_HIP(hipMemcpyAsync(&kept, cur.p + n, sizeof(kept), hipMemcpyDeviceToHost, stream));
    CHL_HIP(hipStreamSynchronize(stream));
    std::unique_ptr<Segment> g(new Segment());
    g->h0 = o.h0;
    g->h1 = o.h1;
    int rc = g->alloc((uint64_t)n, kept, o.fmt);
    if (rc) return rc;
    CHL_HIP(launch_copy_u64(cur.p, (uint64_t)n + 1, g->off.p, stream));
    CHL_HIP(launch_copy_kept(o.view(), b.keep.p, (uint32_t)n, g->off.p, g->key.p, g->lo.p, g->ent.p, stream));
    CHL_HIP(hipEventRecord(ev[5], stream));
    rc = elapsed(ev[4], ev[5], ms);
    if (rc) return rc;
    b.seg = std::move(g);
    b.m = kept;
    return CHL_OK;
  }

  // Appends the batch's segments to the store (no stored row moves), then lets
  // the store merge small segments.
  int append_batch(LabelStore& L, double& ms) {
    for (int s = 0; s < nsides; ++s) sb[s].kept = sb[s].m;
    CHL_HIP(hipEventRecord(ev[6], stream));
    int rc = L.append(std::move(sb[0].seg), nsides == 2 ? std::move(sb[1].seg) : nullptr);
    if (rc) return rc;
    rc = L.tidy(0, stream);
    if (rc) return rc;
    CHL_HIP(hipEventRecord(ev[7], stream));
    return elapsed(ev[6], ev[7], ms);
  }

  int run(LabelStore& L) {
    const double t0 = now_s();
    CHL_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    for (auto& e : ev) CHL_HIP(hipEventCreate(&e));
    std::vector<uint32_t> rank(n);
    for (int64_t i = 0; i < n; ++i) rank[order[i]] = (uint32_t)i;
    int rc = upload_graph(rank);
    if (rc) return rc;
    if (o.delta <= 0) {
      // default bucket width: kDeltaMeanArcs times the mean arc length. Four
      // times against twice: 6% less tree time on the 1M road-like graph and on
      // R-MAT scale 20 with lengths in [1, 16], the same on unit lengths. Eight
      // against four, once frontiers are sorted (wider buckets, fewer rounds):
      // 1M road-like learned build 1.024-1.036 -> 0.997-1.011 s, 16M 97-99 ->
      // 95.3 s; R-MAT scale 20 the same with unit and [1, 16] lengths.
      o.delta = default_delta(g);
    }
    // the same threads in flight at every granularity: a wavefront per root runs
    // four roots where a workgroup runs one, a wide workgroup a quarter
    tpr = o.granularity == 1 ? 64u : o.granularity == 3 ? 1024u : 256u;
    const uint32_t base_slots = o.slots > 0 ? (uint32_t)o.slots : kDefaultSlots;
    rc = alloc_slots(o.autotune && o.slots <= 0 ? 4 * base_slots : slots_for(base_slots, tpr));
    active = std::min<uint32_t>(slots, slots_for(base_slots, tpr));
    if (rc) return rc;
    // When device memory allows fewer than kWideBelow search slots (a large
    // graph: every slot holds 36 bytes per vertex), the default granularity
    // gives each slot a workgroup of 1024 threads while the trees are large, so
    // that more threads stay in flight; once the mean tree of a batch is small
    // the next batch uses 256 threads again.
    const bool auto_wide = o.granularity == 0 && !o.autotune && slots < kWideBelow;
    double last_tree_mean = -1.0;  // processed vertices per root in the last batch (-1: none yet)
    // Static schedule: a fixed width when batch_roots is given; otherwise widths
    // doubling from the slot count (GLL: from 1024 roots) up to 65536 roots
    // (later roots have small trees, and fewer batches mean fewer appends to the
    // store). A GLL batch prunes only by the labels of earlier batches, so wider
    // first batches leave more candidates to clean: at 200K vertices a first
    // batch of 2048 roots doubled the build time against 1024.
    const bool fixed = o.batch_roots > 0;
    // (a batching policy may ask for batches of any width)
    const uint32_t static_cap = (uint32_t)std::max<int64_t>(1, std::min<int64_t>(n, fixed ? o.batch_roots : 65536));
    W = (uint32_t)(hooks && hooks->next_width && !fixed ? n : static_cap);
    uint32_t width = fixed ? static_cap
                           : std::min<uint32_t>(static_cap, o.mode == CHL_MODE_GLL ? 1024u : std::max<uint32_t>(slots, 1024));
    CHL_HIP(roots.reserve(W));
    CHL_HIP(roots2.reserve(W));
    CHL_HIP(spill.reserve(W));
    CHL_HIP(origd.reserve(W));
    CHL_HIP(tree_size.reserve(2 * (size_t)W));
    if (o.verbose) CHL_HIP(touched.reserve(2 * (size_t)W));
    CHL_HIP(label_count.reserve(2 * (size_t)W));
    CHL_HIP(ctr.reserve(CTR_COUNT));
    CHL_HIP(misc.reserve(4));
    CHL_HIP(launch_fill_u64((uint64_t*)misc.p, 0, 4, stream));
    CHL_HIP(cnt.reserve(std::max<int64_t>(n, 1)));
    CHL_HIP(cur.reserve(n + 1));
    CHL_HIP(longq.reserve(n + 1));
    CHL_HIP(scan_tmp.reserve(scan_tmp_words((uint64_t)n)));
    // first batch: the top roots label a large part of the graph
    if (o.out_capacity > 0) {
      out_cap = (uint64_t)o.out_capacity;
    } else {
      // the top roots label a large part of the graph: width * n / 4, bounded
      // by an eighth of the free memory (12 bytes per label)
      size_t freeb = 0, totalb = 0;
      CHL_HIP(hipMemGetInfo(&freeb, &totalb));
      const uint64_t want = (uint64_t)width * (uint64_t)n / 4;
      out_cap = std::max<uint64_t>((uint64_t)1 << 20, std::min<uint64_t>(want, freeb / 8 / 12));
    }
    if (hooks) hooks->slots = slots;
    rc = L.init(n, sym, o.device);
    if (rc) return rc;
    L.verbose = o.verbose;
    if (stats) {
      stats->switch_rank = -1;
      stats->delta_used = o.delta;
      stats->slots_used = active;
      stats->granularity_used = granularity_of(tpr);
      stats->batch_used = W;
    }

    bool plant_phase = o.mode != CHL_MODE_GLL;
    const bool policy_switch = o.mode == CHL_MODE_HYBRID && hooks && hooks->plant_next;
    std::vector<uint32_t> hroots(W), hts(W), hlc(W);
    double lt = 0, stt = 0, ct = 0;
    // PLaNT mode only: in hybrid mode the PLaNT trees of the sampled ranks are
    // the slow trees that the switch exists to avoid (on R-MAT scale 20 the
    // trials alone took several times the whole build)
    if (o.autotune && o.mode == CHL_MODE_PLANT && n >= 4096) {
      rc = autotune(L, lt, base_slots);
      if (rc) return rc;
    }
    if (plant_phase && hooks && hooks->pilot && hooks->on_pilot) {
      // measured PLaNT trees of sampled roots (labels discarded; timed as tree time)
      const uint32_t R = hooks->pilot(hooks->ctx, hroots.data(), std::min<uint32_t>(W, 4096));
      if (R) {
        hids.resize(R);
        for (uint32_t k = 0; k < R; ++k) hids[k] = id_of_rank[hroots[k]];
        CHL_HIP(hipMemcpyAsync(roots.p, hids.data(), sizeof(uint32_t) * R, hipMemcpyHostToDevice, stream));
        rc = run_trees(0, R, hroots.data(), 0, TREE_FLAGS, L, lt);
        if (rc) return rc;
        CHL_HIP(hipMemcpy(hts.data(), tree_size.p, sizeof(uint32_t) * R, hipMemcpyDeviceToHost));
        CHL_HIP(hipMemcpy(hlc.data(), label_count.p, sizeof(uint32_t) * R, hipMemcpyDeviceToHost));
        hooks->on_pilot(hooks->ctx, hroots.data(), hts.data(), hlc.data(), R);
      }
    }
    for (int64_t r0 = 0; r0 < n;) {
      // memory for the next batch: oversized batch buffers first, then (a
      // memory-limited build) search slots go back to the device
      rc = trim_batch_buffers(next_need);
      if (rc) return rc;
      rc = fit_slots();
      if (rc) return rc;
      uint32_t bw = width;
      if (hooks && hooks->next_width) {
        const uint32_t lw = hooks->next_width(hooks->ctx, r0, width, W);
        if (lw) bw = std::min<uint32_t>(lw, W);
      }
      {
        // memory cap under either policy: a batch's labels (output 12 bytes,
        // new segment up to 8, cleaning flag and compacted copy 9: kBatchBytes
        // each) within the free memory, at
        // 1.5 times the labels per root of the last batch's highest eighth of
        // ranks (labels per root fall with rank); before any batch, a root
        // may label every vertex
        size_t freeb = 0, totalb = 0;
        CHL_HIP(hipMemGetInfo(&freeb, &totalb));
        const double per_root = tail_labels >= 0 ? std::max(1.0, 1.5 * tail_labels) : (double)n;
        const double cap = (double)freeb / kBatchBytes / per_root / nsides;
        bw = (uint32_t)std::max<double>(1.0, std::min<double>(bw, cap));
        out_limit = std::max<uint64_t>((uint64_t)1 << 20, (uint64_t)(freeb / kBatchBytes / nsides));
      }
      const uint32_t R = (uint32_t)std::min<int64_t>(bw, n - r0);
      const uint32_t prefix = (uint32_t)r0;
      if (auto_wide) {
        tpr = (last_tree_mean < 0 || last_tree_mean > kWideTree) ? 1024u : 256u;
        active = slots;
        if (stats) stats->granularity_used = granularity_of(tpr);
      }
      const double t_batch = lt + stt + ct;
      // hybrid under a policy: the policy picks each batch's tree kind until it
      // makes the pruned kind final (a pruned batch is exact as any other)
      bool plant_batch = plant_phase;
      if (policy_switch && plant_phase && prefix > 0) {
        int fin = 0;
        plant_batch = hooks->plant_next(hooks->ctx, r0, R, &fin) != 0;
        if (stats && !plant_batch && stats->switch_rank < 0) stats->switch_rank = r0;
        if (stats && plant_batch) stats->switch_rank = -1;
        if (!plant_batch && fin) plant_phase = false;
      }
      for (uint32_t k = 0; k < R; ++k) hroots[k] = (uint32_t)(r0 + k);
      if (hooks && hooks->order_roots) hooks->order_roots(hooks->ctx, hroots.data(), R);
      if (hooks && hooks->out_estimate) {
        const uint64_t e = hooks->out_estimate(hooks->ctx, hroots.data(), R);
        if (e) out_cap = std::max<uint64_t>(e, (uint64_t)1 << 16);
      }
      // the output buffer within the memory cap as well: an estimate made for a
      // wider batch than the cap allowed would not fit (a batch that needs more
      // spills and grows the buffer by what it measured)
      out_cap = std::min(out_cap, out_limit);
      hids.resize(R);
      for (uint32_t k = 0; k < R; ++k) hids[k] = id_of_rank[hroots[k]];
      CHL_HIP(hipMemcpyAsync(roots.p, hids.data(), sizeof(uint32_t) * R, hipMemcpyHostToDevice, stream));
      uint32_t flags = 0;
      if (plant_batch) flags = TREE_FLAGS | ((o.plant_prune && prefix > 0) ? TREE_PRUNE : 0u);
      else flags = prefix > 0 ? TREE_PRUNE : 0u;
      if (flags == 0 && o.out_capacity <= 0) {
        // unpruned trees without flags label every vertex they reach
        size_t freeb = 0, totalb = 0;
        CHL_HIP(hipMemGetInfo(&freeb, &totalb));
        out_cap = std::max<uint64_t>(out_cap, std::min<uint64_t>((uint64_t)R * (uint64_t)n + 4096, freeb / 4 / 12));
      }
      if (flags & TREE_PRUNE) {
        uint64_t mr = 0;
        for (int s = 0; s < nsides; ++s) {
          CHL_HIP(launch_fill_u64((uint64_t*)misc.p + 1, 0, 1, stream));
          CHL_HIP(launch_max_row(L.d_views(s), L.nseg(), (uint32_t)n, misc.p + 1, stream));
          unsigned long long h = 0;
          CHL_HIP(hipMemcpyAsync(&h, misc.p + 1, sizeof(h), hipMemcpyDeviceToHost, stream));
          CHL_HIP(hipStreamSynchronize(stream));
          mr = std::max<uint64_t>(mr, h);
        }
        rc = ensure_gtab(mr);
        if (rc) return rc;
      }
      // learned batching, PLaNT batches of a memory-limited build (too few search
      // slots for the many small trees of the lower ranks): roots predicted to
      // grow small trees run on the small-tree kernel first. With the full slot
      // count the tree kernel is the faster of the two for small trees too (1M
      // road-like graph: 22.6 against 18.2 ms for the batch of 250K roots
      // after rank 400K).
      // Roots predicted between the small tables and kMediumRoute run on the
      // medium-tree kernel next (256 threads per root, a table of kMediumCap
      // entries: about 80 times as many trees in flight per byte as a dense slot).
      bool small_pass = false, medium_pass = false;
      if (flags == TREE_FLAGS && hooks && hooks->route_small && slots < slots_wanted) {
        bigh.resize(R);
        bigmh.resize(R);
        const uint32_t ns = hooks->route_small(hooks->ctx, hroots.data(), R, kSmallRoute, bigh.data());
        if (ns >= std::min<uint32_t>(R, 1024)) {
          rc = ensure_small();
          if (rc) return rc;
          CHL_HIP(bigd.reserve(std::max<size_t>(W, 1)));
          CHL_HIP(hipMemcpyAsync(bigd.p, bigh.data(), R, hipMemcpyHostToDevice, stream));
          small_pass = true;
        }
        const uint32_t nm =
            o.medium_trees ? hooks->route_small(hooks->ctx, hroots.data(), R, kMediumRoute, bigmh.data()) : 0;
        if (nm >= (small_pass ? ns : 0) + kMediumMin) {
          rc = ensure_medium();
          if (rc) return rc;
          medium_pass = true;
        }
      }
      for (int s = 0; s < nsides; ++s) {
        rc = run_trees(s, R, hroots.data(), prefix, flags, L, lt, small_pass, medium_pass);
        if (rc) return rc;
        if (s == 0 && nsides == 2) {
          // tree statistics are per side; keep the forward side for the hooks
          CHL_HIP(hipMemcpy(hts.data(), tree_size.p, sizeof(uint32_t) * R, hipMemcpyDeviceToHost));
          CHL_HIP(hipMemcpy(hlc.data(), label_count.p, sizeof(uint32_t) * R, hipMemcpyDeviceToHost));
        }
      }
      if (nsides == 1) {
        CHL_HIP(hipMemcpy(hts.data(), tree_size.p, sizeof(uint32_t) * R, hipMemcpyDeviceToHost));
        CHL_HIP(hipMemcpy(hlc.data(), label_count.p, sizeof(uint32_t) * R, hipMemcpyDeviceToHost));
      }
      const bool device_switch = plant_phase && o.mode == CHL_MODE_HYBRID && !policy_switch;
      if (device_switch)
        CHL_HIP(launch_switch_test(tree_size.p, label_count.p, R, o.switch_ratio, misc.p, stream));
      for (int s = 0; s < nsides; ++s) {
        rc = build_segment(s, (uint32_t)r0, R, stt);
        if (rc) return rc;
      }
      if (!(flags & TREE_FLAGS)) {
        for (int s = 0; s < nsides; ++s) {
          rc = clean_side(s, ct);
          if (rc) return rc;
        }
        for (int s = 0; s < nsides; ++s) {
          rc = compact_side(s, ct);
          if (rc) return rc;
        }
      }
      rc = append_batch(L, ct);
      if (rc) return rc;
      if (stats) {
        stats->batches++;
        if (flags & TREE_FLAGS) stats->trees_plant += (int64_t)R * nsides;
        else stats->trees_pruned += (int64_t)R * nsides;
      }
      if (hooks && hooks->on_batch)
        hooks->on_batch(hooks->ctx, hroots.data(), hts.data(), hlc.data(), R, flags, lt + stt + ct - t_batch,
                        sb[0].kept + (nsides == 2 ? sb[1].kept : 0));
      if (device_switch) {
        unsigned long long f = 0;
        CHL_HIP(hipMemcpyAsync(&f, misc.p, sizeof(f), hipMemcpyDeviceToHost, stream));
        CHL_HIP(hipStreamSynchronize(stream));
        if (f) {
          plant_phase = false;
          if (stats) stats->switch_rank = r0 + R;
        }
      }
      if (o.verbose) {
        uint64_t emitted = sb[0].kept + (nsides == 2 ? sb[1].kept : 0);
        fprintf(stderr,
                "batch %lld: ranks [%lld, %lld) flags %u labels %llu store %llu (%u segments, %.2f GB) lt %.2f st "
                "%.2f ct %.2f s wall %.2f s\n",
                (long long)stats_batches, (long long)r0, (long long)(r0 + R), flags, (unsigned long long)emitted,
                (unsigned long long)L.total[0], L.nseg(), L.bytes() / 1e9, lt / 1e3, stt / 1e3, ct / 1e3,
                now_s() - t0);
      }
      ++stats_batches;
      {
        double sum = 0;
        for (uint32_t k = 0; k < R; ++k) sum += hts[k];
        last_tree_mean = R ? sum / R : 0.0;
      }
      {
        // labels per root (candidates for pruned trees) of the batch's highest
        // eighth of ranks, forward side
        const uint32_t lo = (uint32_t)(r0 + R - std::max<uint32_t>(1, R / 8));
        double sum = 0, cnt = 0;
        for (uint32_t k = 0; k < R; ++k)
          if (hroots[k] >= lo) {
            sum += hlc[k];
            cnt += 1;
          }
        if (cnt > 0) tail_labels = sum / cnt;
      }
      // output estimate for the next batch from this one's labels per root
      {
        const uint32_t nextR = fixed ? static_cap : std::min<uint32_t>(static_cap, width * 2);
        uint64_t emitted = std::max(sb[0].kept, nsides == 2 ? sb[1].kept : 0);
        uint64_t est = emitted / std::max<uint32_t>(R, 1) * nextR;
        next_need = est + est / 4;
        out_cap = std::max<uint64_t>(out_cap, next_need);
        width = nextR;
      }
      r0 += R;
    }
    CHL_HIP(hipStreamSynchronize(stream));
    {
      // the store's segments are merged into one where memory allows (queries
      // then merge-join one row pair per side); when the merged copy does not
      // fit beside the search slots and batch buffers, those go back to the
      // device first (otherwise they stay in the allocation cache for the next
      // build)
      size_t freeb = 0, totalb = 0;
      CHL_HIP(hipMemGetInfo(&freeb, &totalb));
      if (!o.keep_segments && L.nseg() > 1 && (double)L.bytes() * 1.2 > (double)freeb) {
        cell.release();
        qd.release();
        for (auto* b : {&qa, &qb, &vis}) b->release();
        for (auto* b : {&gkey, &gval, &seq}) b->release();
        for (auto& b : sb) {
          b.row.release();
          b.hd.release();
          b.keep.release();
          b.seg.reset();
        }
        dev_cache_flush();
      }
      CHL_HIP(hipEventRecord(ev[6], stream));
      if (!o.keep_segments) {
        rc = L.consolidate(stream);
        if (rc) return rc;
      }
      CHL_HIP(hipEventRecord(ev[7], stream));
      rc = elapsed(ev[6], ev[7], ct);
      if (rc) return rc;
    }
    const double t1 = now_s();
    if (stats) {
      stats->segments = L.nseg();
      stats->slots_final = active;
      stats->store_bytes = (double)L.bytes();
      stats->lt = lt / 1e3;
      stats->st = stt / 1e3;
      stats->ct = ct / 1e3;
      stats->total_s = t1 - t0;
      stats->labels = (int64_t)(L.total[0] + (sym ? L.total[0] : L.total[1]));
      stats->avg_label = (double)stats->labels / (double)n / 2.0;
      stats->clean_ms += ct;
    }
    return CHL_OK;
  }
};

int build_labels(const Graph& g, DeviceGraph& dg, const std::vector<uint32_t>& order, const chl_build_opts& opts,
                 chl_build_stats* stats, BuildHooks* hooks, LabelStore& L) {
  Builder b(g, dg, order, opts, stats, hooks);
  return b.run(L);
}

}  // namespace chl
