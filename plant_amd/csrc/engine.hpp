// Host-side interface of the device engine (label store, build, hooks).
#pragma once
#include <cstdint>
#include <memory>
#include <vector>

#include "dev.hpp"

namespace chl {

// One segment of a label side: the labels of hub ranks [h0, h1) as a CSR over
// row ids (off, rows + 1 offsets), entries in SEG_F5 (key, lo) or SEG_F8 (ent)
// form.
struct Segment {
  uint32_t h0 = 0, h1 = 0;
  uint32_t fmt = SEG_F8;
  uint64_t total = 0;
  BigBuf<uint64_t> off;
  BigBuf<uint32_t> key;
  BigBuf<uint8_t> lo;
  BigBuf<uint64_t> ent;
  SegView view() const;
  uint64_t entry_bytes() const;
  int alloc(uint64_t rows, uint64_t count, uint32_t fmt);
};

// Labels resident on the device. Rows are indexed by row id (the rank for built
// labels, the vertex id for labels read from a file); each side is a list of
// segments with ascending, disjoint hub ranges, the same for both sides.
// shared_sides: side 1 is side 0 (symmetric graph).
struct LabelStore {
  int64_t n = 0;
  bool shared_sides = false;
  int device = 0;
  std::vector<std::unique_ptr<Segment>> seg[2];
  DBuf<SegView> dview[2];  // device copies of the segment views
  BigBuf<uint64_t> rng[2];  // per side, built for queries: vertex-major row bounds in every segment
  uint64_t total[2] = {0, 0};
  int verbose = 0;  // 1: a line per merge on stderr
  int nsides() const { return shared_sides ? 1 : 2; }
  const std::vector<std::unique_ptr<Segment>>& side(int s) const { return seg[shared_sides ? 0 : s]; }
  const SegView* d_views(int s) const { return dview[shared_sides ? 0 : s].p; }
  uint32_t nseg() const { return (uint32_t)seg[0].size(); }
  uint64_t side_total(int s) const { return total[shared_sides ? 0 : s]; }
  uint64_t bytes() const;  // device bytes of entries and offsets
  int init(int64_t n, bool shared, int dev);
  int sync_views();
  // appends one batch's segments (s1 ignored with shared sides)
  int append(std::unique_ptr<Segment> s0, std::unique_ptr<Segment> s1);
  int merge_range(size_t a, size_t b, hipStream_t st);
  int tidy(size_t budget, hipStream_t st);
  int consolidate(hipStream_t st);
  // merges down to kQuerySeg segments, then builds the range tables of queries
  int build_ranges(const uint32_t* row_of, hipStream_t st);
  QueryView query_view() const;
  int download(int side, std::vector<uint64_t>& off, std::vector<uint64_t>& ent) const;
  int upload(int side, const std::vector<uint64_t>& off, const std::vector<uint64_t>& ent);
};

// The graph resident on the device, kept in the graph handle across builds:
// vertices renumbered into a breadth-first locality order (device ids), one CSR
// per arc range with arcs packed as (head id << 32) | length. A symmetric graph
// keeps only the out-arc range.
struct DeviceGraph {
  int device = -1;
  int64_t n = 0;
  bool sym = false;
  std::vector<uint32_t> inv;  // device id -> vertex
  std::vector<uint32_t> pos;  // vertex -> device id
  DBuf<uint32_t> off[2];
  DBuf<uint64_t> arc[2];
  int build(const Graph& g, int device);
};

// Callbacks into the batching policy (static order, learned packing).
struct BuildHooks {
  void* ctx = nullptr;
  // May permute roots[0..R) of the next batch (the set is fixed).
  void (*order_roots)(void* ctx, uint32_t* roots, uint32_t R) = nullptr;
  // Per-root measured tree sizes and label counts of a finished batch (flags:
  // TREE_*), its device time (trees, sort, clean and merge) and the labels it
  // added to the store.
  void (*on_batch)(void* ctx, const uint32_t* roots, const uint32_t* tree_size, const uint32_t* label_count,
                   uint32_t R, uint32_t flags, double ms, uint64_t labels) = nullptr;
  // Hybrid mode: 1 if the batch of R roots at rank r0 runs PLaNT trees, 0 if it
  // runs pruned trees; *final = 1 makes every later batch pruned. nullptr: the
  // device-side tree/label ratio test decides.
  int (*plant_next)(void* ctx, int64_t r0, uint32_t R, int* final) = nullptr;
  // One tree launch finished: device time and the launch counters (CTR_*).
  void (*on_tree_launch)(void* ctx, int side, uint32_t R, float ms, const unsigned long long* ctr) = nullptr;
  // Width of the batch starting at rank r0 (at most max_width); 0 keeps the static width.
  uint32_t (*next_width)(void* ctx, int64_t r0, uint32_t static_width, uint32_t max_width) = nullptr;
  // Output buffer estimate (labels) for roots[0..R); 0 keeps the static estimate.
  uint64_t (*out_estimate)(void* ctx, const uint32_t* roots, uint32_t R) = nullptr;
  // Before the first batch (PLaNT trees only): ranks whose trees are measured and
  // whose labels are discarded (returns the count, at most cap), and their sizes.
  uint32_t (*pilot)(void* ctx, uint32_t* roots, uint32_t cap) = nullptr;
  void (*on_pilot)(void* ctx, const uint32_t* roots, const uint32_t* tree_size, const uint32_t* label_count,
                   uint32_t R) = nullptr;
  // PLaNT batch of roots[0..R): big[k] = 1 for a root predicted to grow a tree of
  // more than `limit` processed vertices, else 0; returns the roots predicted
  // small (0: no prediction, every root to the tree kernel).
  uint32_t (*route_small)(void* ctx, const uint32_t* roots, uint32_t R, uint32_t limit, uint8_t* big) = nullptr;
  uint32_t slots = 0;  // set by the builder before the first batch
};

int ensure_device(int device);
// The default bucket width of the tree searches: a multiple of the mean arc length.
int32_t default_delta(const Graph& g);
// dg: the graph on the device (built into *dg first when dg->device differs).
int build_labels(const Graph& g, DeviceGraph& dg, const std::vector<uint32_t>& order, const chl_build_opts& opts,
                 chl_build_stats* stats, BuildHooks* hooks, LabelStore& L);
// Sampled-betweenness order computed on the device; the same order (and
// scores, per vertex id) as the host order_betweenness. device_ms: kernel and
// sort time.
int order_betweenness_device(const Graph& g, DeviceGraph& dg, int device, int64_t roots, uint64_t seed,
                             std::vector<uint32_t>& order, std::vector<uint64_t>* score, double* device_ms);
// Degree order computed on the device; the same order as the host
// order_degree. device_ms: key and sort time.
int order_degree_device(const Graph& g, int device, uint64_t seed, bool reference_ties, std::vector<uint32_t>& order,
                        double* device_ms);
// Build under the batching policy of opts.batching (static rank order or learned packing).
int run_build(const Graph& g, DeviceGraph& dg, const std::vector<uint32_t>& order, const chl_build_opts& opts,
              chl_build_stats* stats, LabelStore& L);

}  // namespace chl
