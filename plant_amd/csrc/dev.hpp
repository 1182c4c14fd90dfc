// Device-side declarations shared by the HIP kernels and the host engine.
#pragma once
#include <hip/hip_runtime.h>

#include <utility>
#include <vector>

#include <cstdint>

#include "../../include/chl.h"
#include "host.hpp"

namespace chl {

// Returns CHL_E_DEVICE (and records the message) from the enclosing int-returning
// function when a HIP call fails.
#define CHL_HIP(x)                                                                                  \
  do {                                                                                              \
    hipError_t e_ = (x);                                                                            \
    if (e_ != hipSuccess)                                                                           \
      return ::chl::set_error(CHL_E_DEVICE, "%s:%d %s: %s", __FILE__, __LINE__, #x,                 \
                              hipGetErrorString(e_));                                               \
  } while (0)

// Process-wide cache of device allocations: a released block is kept (up to a
// cap) and handed to a later request of at least its size and at most twice
// it, so repeated builds and query batches do not pay hipMalloc/hipFree. On an
// allocation failure the cache is emptied and the allocation retried.
hipError_t dev_alloc(void** p, size_t bytes, size_t* got);
void dev_free(void* p, size_t bytes);
void dev_cache_flush();
size_t dev_cached_bytes();  // bytes of freed blocks the allocator keeps for the current device

// Device array whose trailing part can go back to the device without moving
// what is kept: a virtual range is reserved and backed by physical chunks of a
// fixed size (HIP virtual memory management); shrink unmaps and releases the
// chunks wholly past the kept length. The search slots of a memory-limited
// build live in these, so that slots go back as the label store grows without
// freeing, allocating and clearing the slots that stay. Where the runtime
// refuses virtual memory the array is one plain allocation and shrink keeps it.
struct ChunkArray {
  char* base = nullptr;
  size_t reserved = 0;  // bytes of the virtual range
  size_t mapped = 0;    // usable bytes
  size_t chunk = 0;     // bytes per physical chunk
  int device = 0;
  bool vmm = false;
  std::vector<hipMemGenericAllocationHandle_t> maps;  // chunk handles in address order
  ChunkArray() = default;
  ChunkArray(const ChunkArray&) = delete;
  ChunkArray& operator=(const ChunkArray&) = delete;
  ~ChunkArray() { release(); }
  // A new array of at least `bytes` (contents undefined) in chunks of about
  // chunk_bytes.
  hipError_t alloc(size_t bytes, size_t chunk_bytes);
  // Keeps the first `bytes` (rounded up to whole chunks); returns the rest.
  hipError_t shrink(size_t bytes);
  void release();
};

template <class T>
struct ChunkBuf {
  ChunkArray a;
  T* p = nullptr;
  size_t n = 0;  // elements usable
  hipError_t alloc(size_t count, size_t chunk_bytes) {
    release();
    hipError_t e = a.alloc((count ? count : 1) * sizeof(T), chunk_bytes);
    p = reinterpret_cast<T*>(a.base);
    n = a.mapped / sizeof(T);
    return e;
  }
  hipError_t shrink(size_t count) {
    hipError_t e = a.shrink(count * sizeof(T));
    n = a.mapped / sizeof(T);
    return e;
  }
  void release() {
    a.release();
    p = nullptr;
    n = 0;
  }
};

// Device buffer that returns its memory to the cache.
template <class T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  size_t bytes = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() { release(); }
  void release() {
    if (p) dev_free(p, bytes);
    p = nullptr;
    n = 0;
    bytes = 0;
  }
  // At least `count` elements (contents undefined); keeps a large enough buffer.
  hipError_t reserve(size_t count) {
    if (count <= n && p) return hipSuccess;
    release();
    size_t got = 0;
    hipError_t e = dev_alloc((void**)&p, (count ? count : 1) * sizeof(T), &got);
    if (e == hipSuccess) {
      bytes = got;
      n = got / sizeof(T);
    } else {
      p = nullptr;
    }
    return e;
  }
  void swap(DBuf& o) {
    T* tp = p; p = o.p; o.p = tp;
    size_t tn = n; n = o.n; o.n = tn;
    size_t tb = bytes; bytes = o.bytes; o.bytes = tb;
  }
};

// Search-slot array: from the allocation cache (repeated builds reuse it), or,
// in a memory-limited build, a ChunkBuf whose trailing slots can go back.
template <class T>
struct SlotBuf {
  DBuf<T> d;
  ChunkBuf<T> c;
  T* p = nullptr;
  bool chunked = false;
  hipError_t alloc(size_t count, bool in_chunks, size_t chunk_bytes) {
    release();
    chunked = in_chunks;
    hipError_t e = chunked ? c.alloc(count, chunk_bytes) : d.reserve(count);
    p = chunked ? c.p : d.p;
    return e;
  }
  // Keeps the first `count` elements; returns whole chunks past them.
  hipError_t shrink(size_t count) { return chunked ? c.shrink(count) : hipSuccess; }
  void release() {
    d.release();
    c.release();
    p = nullptr;
  }
};

// ---- segmented label store ------------------------------------------------
// A label side is a list of segments. A segment holds the labels whose hub rank
// lies in [h0, h1) (one batch, or several merged) as a CSR over row ids (the
// rank for built labels, the vertex id for labels read from a file), rows
// hub-ascending. Entry formats:
//   SEG_F5: key u32 = ((hub - h0) << 8) | (dist >> 8) and lo u8 = dist & 0xFF,
//           5 bytes; needs hub - h0 < 2^24 and dist <= 65535. key is ascending
//           along a row (the hub is in its top bits).
//   SEG_F8: ent u64 = (hub << 32) | dist, 8 bytes, any value (the escape).
// Hub ranges of the segments of a side are disjoint, so a merge-join over a
// vertex's whole row is the minimum of the merge-joins over its segments.
enum : uint32_t { SEG_F5 = 5, SEG_F8 = 8 };
constexpr uint32_t kSegMax = 48;  // segments per side at most (more are merged first)
constexpr uint32_t kF5MaxDist = 65535;
constexpr uint32_t kF5MaxWidth = 1u << 24;
constexpr uint64_t kF5Pad = 4;

struct SegView {
  const uint64_t* off;  // row offsets, n + 1, segment-relative
  const uint32_t* key;  // SEG_F5 (allocated with kF5Pad entries to spare, for 4-entry group loads)
  const uint8_t* lo;    // SEG_F5
  const uint64_t* ent;  // SEG_F8
  uint32_t h0;
  uint32_t fmt;
};

// Entry k of a segment as (hub, dist).
__device__ __forceinline__ void seg_entry(const SegView& s, uint64_t k, uint32_t& h, uint32_t& d) {
  if (s.fmt == SEG_F8) {
    const uint64_t e = s.ent[k];
    h = (uint32_t)(e >> 32);
    d = (uint32_t)e;
  } else {
    const uint32_t x = s.key[k];
    h = s.h0 + (x >> 8);
    d = ((x & 0xFFu) << 8) | s.lo[k];
  }
}

// Tree-kernel mode bits.
enum : uint32_t {
  TREE_FLAGS = 1u,  // propagate the most important rank on the path (PLaNT, distributed_memory/akiba.hpp:71,89)
  TREE_PRUNE = 2u,  // prune by the completed rank-prefix labels (shared_memory/akiba.hpp:94-110)
  TREE_KEEP = 4u,   // alone: a plain search per workgroup (root = workgroup index, R <= slots) whose
                    // final distances stay in the slot's cells and whose reached vertices stay in
                    // its vis list (touched[b] of them); no labels, no reset
};

// Packed per-vertex search state: (epoch key << 34) | (dist << 2) | tag.
//  * tag: kTagClean when every vertex on the best path found so far, the vertex
//    itself included, is ranked at or below the root (TREE_FLAGS; without it
//    every reached vertex is clean), kTagCovered when some vertex on it is more
//    important than the root, kTagPruned once the vertex is pruned at that
//    distance. Among equal distances the 64-bit minimum keeps the smaller tag,
//    so a path through a more important vertex wins a tie, and an
//    equal-distance update cannot revive a pruned vertex.
//  * epoch key: kEpochMax - (the slot's tree count). Every tree a slot starts
//    gets a smaller key, so a cell left by an earlier tree of the slot compares
//    above every state of the current one and reads as unreached: the cells
//    need no reset between trees. Unreached = all ones (key kEpochMax).
static constexpr uint64_t kStateInf = ~0ull;
static constexpr uint32_t kTagPruned = 0, kTagCovered = 1, kTagClean = 2;
static constexpr uint32_t kEpochMax = (1u << 30) - 1;
// A slot's stamps grow by two per round over all its trees; past this bound
// (and past kEpochMax trees) the slot's cells are reset and both counts restart.
static constexpr uint32_t kStampMax = 0x7FFFFFFFu;
__host__ __device__ inline uint64_t state_word(uint32_t ekey, uint32_t d, uint32_t tag) {
  return ((uint64_t)ekey << 34) | ((uint64_t)d << 2) | tag;
}
__host__ __device__ inline uint32_t state_dist(uint64_t x) { return (uint32_t)(x >> 2); }
__host__ __device__ inline uint32_t state_tag(uint64_t x) { return (uint32_t)x & 3u; }
__host__ __device__ inline uint32_t state_ekey(uint64_t x) { return (uint32_t)(x >> 34); }

// Counters of algorithmic work, summed over a launch.
// CTR_PROCESSED / CTR_ARCS: vertices processed and arcs scanned, every
// processing counted (a label-correcting search processes a vertex again when
// its state improves). CTR_SETTLED / CTR_SETTLED_ARCS: distinct vertices
// settled (final distance below the last bucket bound) and the arcs a search
// scans when it processes each of them once: the minimal work of the trees.
// CTR_LABELS: labels written; CTR_OVERFLOW: roots whose labels did not fit (their
// indices go to spill[]); CTR_SPILL: labels those roots still have to write.
// CTR_SPILL_START: smallest output offset reserved by a root that did not fit
// (all ones when none), so [0, CTR_SPILL_START) is the written prefix.
// CTR_SKIPPED: roots not started because the output was already full (they are
// in spill[] too, with unknown label counts). CTR_MAXD: largest label distance
// written.
enum { CTR_PROCESSED = 0, CTR_ARCS = 1, CTR_PRUNE_ENT = 2, CTR_LABELS = 3, CTR_OVERFLOW = 4, CTR_NEXT = 5,
       CTR_SPILL = 6, CTR_ROUNDS = 7, CTR_SPILL_START = 8, CTR_SKIPPED = 9, CTR_MAXD = 10, CTR_SETTLED = 11,
       CTR_SETTLED_ARCS = 12, CTR_COUNT = 13 };

struct TreeParams {
  // graph in device ids (a locality order): arcs of the range walked by the tree
  // (out-arcs for a forward tree), arc = (head << 32) | length; rank[id] = the
  // vertex's position in the vertex order
  const uint32_t* off;
  const uint64_t* arc;
  const uint32_t* rank;
  uint32_t n;
  // label store (rows in rank ids, hub-sorted, all hubs < prefix): device
  // arrays of segment views of the side read at a visited vertex and of the
  // side read at the root
  const SegView* scan_view;
  const SegView* root_view;
  uint32_t scan_nseg, root_nseg;
  uint32_t prefix;
  // batch (roots as device ids)
  const uint32_t* roots;
  uint32_t R;
  uint32_t delta;
  // per-slot scratch, slot = workgroup index: cell holds 2 n words per slot
  // (state word, stamp and rank), the lists and the pile n entries per slot;
  // seq holds two words per slot (trees started, next stamp base), zero when
  // the slot's cells are initialised and carried from launch to launch
  uint64_t* cell;
  uint32_t* seq;
  uint32_t* qa;
  uint32_t* qb;
  uint32_t* vis;
  uint64_t* qd;     // deferred pile, n entries per slot
  uint32_t* gkey;   // global root-label hash per slot (gcap entries) when the row is too long for LDS
  uint32_t* gval;
  uint32_t gcap;
  // output: emitted labels, row id (the vertex's rank) and (hub << 32) | dist
  uint32_t* out_row;
  uint64_t* out_hd;
  uint64_t out_cap;
  uint32_t* spill;  // [R] indices into roots whose labels did not fit
  // per root statistics and launch counters
  uint32_t* tree_size;
  uint32_t* label_count;
  uint32_t* touched;  // optional: vertices reached per root
  unsigned long long* ctr;  // CTR_* (CTR_NEXT is the work-queue cursor)
  // small-tree searches (small.hip), per small slot: a hash table of scap
  // entries (state word, vertex + 1 with 0 for empty, stamp) and three lists of
  // scap table indices (two frontiers, the inserted entries); big[ri] = 1 marks
  // a root predicted too large, which goes to spill[] unstarted
  uint64_t* sstate;
  uint32_t* skey;
  uint32_t* sstamp;
  uint32_t* slist;
  uint32_t scap;
  const uint8_t* big;
  // tree_kernel: frontiers of at least sort_min entries (0: none) and at most
  // the kernel's LDS bound are sorted by device id before they are processed
  uint32_t sort_min;
};

// One workgroup per slot of threads_per_root threads: 64 (a wavefront per root),
// 256 or 1024.
hipError_t launch_tree(const TreeParams& p, uint32_t flags, uint32_t slots, int threads_per_root, hipStream_t s);
// PLaNT trees of the batch's roots over `slots` hashed slots of scap entries
// (small.hip), `threads` per root: 64 (small tables) or 256 (medium tables); a
// root marked big, or whose tree outgrows half its table, goes to spill[]
// (counted in CTR_SKIPPED) for a later launch.
hipError_t launch_small_trees(const TreeParams& p, uint32_t slots, int threads, hipStream_t s);
// scap-entry tables of `slots` hashed slots set to empty
hipError_t launch_init_small(const TreeParams& p, uint32_t slots, hipStream_t s);
// count cells set to the unreached state (state all ones, stamp 0) with the
// rank of their vertex (cell i: device id i mod n); the slots' seq words must be
// zeroed with them
hipError_t launch_init_cells(uint64_t* cell, size_t count, const uint32_t* rank, uint32_t n, hipStream_t s);
hipError_t launch_fill_u64(uint64_t* p, uint64_t v, size_t count, hipStream_t s);
hipError_t launch_fill_u32(uint32_t* p, uint32_t v, size_t count, hipStream_t s);
// dst[idx[j]] = src[j] for j in [0, m)
hipError_t launch_scatter_u32(const uint32_t* src, const uint32_t* idx, uint32_t m, uint32_t* dst, hipStream_t s);

// Sampled-betweenness order (order.hip) over the trees the tree kernel left in
// S slots (TREE_KEEP): slot s holds the search of roots[s]. Per slot sized n:
// cell 2 n words (the tree kernel's), vis (reached vertices, touched[s] of
// them), parent, acc ((children << 32) | subtree), leaf (nleaf[s] of them,
// zeroed first). score: per device id, summed over the roots.
struct BcTrees {
  const uint64_t* cell;
  const uint32_t* vis;
  const uint32_t* touched;
  const uint32_t* roots;  // device ids
  uint32_t* parent;
  unsigned long long* acc;
  uint32_t* leaf;
  uint32_t* nleaf;
  const uint32_t* ioff;   // in-arcs, arc = (tail << 32) | length
  const uint64_t* iarc;
  const uint32_t* inv;    // device id -> vertex id (the parent tie rule)
  unsigned long long* score;
  uint32_t n, S;
};
// The four passes over the S trees; max_touched bounds touched[] (grid width).
hipError_t launch_bc_trees(const BcTrees& t, unsigned max_touched, hipStream_t s);
hipError_t launch_score_keys(const unsigned long long* score, const uint32_t* inv, uint32_t n, uint64_t* key,
                             uint64_t* val, hipStream_t s);
// Degree order keys over vertex ids in [0, n); off1 / off0 are the out / in arc
// offsets (n + 1 each) and every degree is below 2^32. reference_ties: key =
// (degree descending, id descending); otherwise key = the seeded tie key.
// val[v] = v.
hipError_t launch_degree_keys(const uint64_t* off1, const uint64_t* off0, uint32_t n, uint64_t seed, int reference_ties,
                              uint64_t* key, uint64_t* val, hipStream_t s);
// key[i] = the descending degree key of vertex val[i] (32 bits).
hipError_t launch_degree_rekey(const uint64_t* off1, const uint64_t* off0, const uint64_t* val, uint32_t n,
                               uint64_t* key, hipStream_t s);

// Device array allocated and freed directly (never cached): segment storage.
template <class T>
struct BigBuf {
  T* p = nullptr;
  size_t n = 0;
  BigBuf() = default;
  BigBuf(const BigBuf&) = delete;
  BigBuf& operator=(const BigBuf&) = delete;
  ~BigBuf() { release(); }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  hipError_t alloc(size_t count) {
    release();
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void**)&p, count * sizeof(T));
    if (e != hipSuccess) {
      (void)hipGetLastError();
      dev_cache_flush();
      e = hipMalloc((void**)&p, count * sizeof(T));
    }
    if (e != hipSuccess) {
      (void)hipGetLastError();
      p = nullptr;
      return e;
    }
    n = count;
    return hipSuccess;
  }
};

// Batch pipeline: the tree kernel writes, per label, its row id (row[k]) and
// (hub << 32) | dist (hd[k]).
// cnt[row[k]] += 1 for k in [0, m) (cnt zeroed first).
hipError_t launch_count_rows(const uint32_t* row, uint64_t m, uint32_t* cnt, hipStream_t s);
// Exclusive scan: out[i] = sum of in[0..i) for i in [0, count], so out[count]
// is the total. tmp: scan_tmp_words(count) words.
size_t scan_tmp_words(uint64_t count);
hipError_t launch_scan_u32(const uint32_t* in, uint64_t count, uint64_t* out, uint64_t* tmp, hipStream_t s);
// Each label to its row of a new segment: position = atomicAdd(cur[row], 1)
// (cur = a copy of the segment offsets); fmt SEG_F5 (base h0) or SEG_F8.
hipError_t launch_scatter_labels(const uint32_t* row, const uint64_t* hd, uint64_t m, uint64_t* cur, uint32_t fmt,
                                 uint32_t h0, uint32_t* key, uint8_t* lo, uint64_t* ent, hipStream_t s);
// Sorts every row of a segment by hub. longq: n + 1 words of scratch (a list
// of the rows too long for one thread).
hipError_t launch_sort_rows(const uint64_t* off, uint32_t rows, uint32_t fmt, uint32_t* key, uint8_t* lo,
                            uint64_t* ent, uint32_t* longq, hipStream_t s);
// In-batch cleaning (shared_memory/labeling.hpp:78-88 over the batch labels):
// keep[k] = 0 when entry k = (h, d) of row v of seg (this side) is covered by a
// common hub x < h of row v and row h of oseg (the other side of the batch):
// L(v)[x] + L(h)[x] <= d. One wavefront per row. ctr[CTR_PRUNE_ENT] += entries
// of the other rows read.
hipError_t launch_clean(const SegView& seg, const SegView& oseg, uint32_t rows, uint8_t* keep,
                        unsigned long long* ctr, hipStream_t s);
// cnt[r] = kept entries of row r.
hipError_t launch_count_kept(const uint64_t* off, const uint8_t* keep, uint32_t rows, uint32_t* cnt, hipStream_t s);
// The kept entries of every row into a new segment of the same format (offsets
// noff); row order and entry order are kept.
hipError_t launch_copy_kept(const SegView& seg, const uint8_t* keep, uint32_t rows, const uint64_t* noff,
                            uint32_t* key, uint8_t* lo, uint64_t* ent, hipStream_t s);
// Merge of adjacent segments (ascending hub ranges) into one: cnt[r] = sum of
// the row lengths (row_total), then rows concatenated in segment order
// (concat_rows), each entry re-encoded to the output format. views: device
// array of k views.
hipError_t launch_row_total(const SegView* views, uint32_t k, uint32_t rows, uint32_t* cnt, hipStream_t s);
hipError_t launch_concat_rows(const SegView* views, uint32_t k, uint32_t rows, const uint64_t* noff, uint32_t fmt,
                              uint32_t h0, uint32_t* key, uint8_t* lo, uint64_t* ent, hipStream_t s);
// out[0] = max over rows of the summed row length over the k segments (out zeroed).
hipError_t launch_max_row(const SegView* views, uint32_t k, uint32_t rows, unsigned long long* out, hipStream_t s);
// Switch test of the hybrid build (distributed_memory/akiba.hpp:297-303): sets
// *flag when some root's tree_size / label_count exceeds ratio.
hipError_t launch_switch_test(const uint32_t* tree_size, const uint32_t* label_count, uint32_t R, double ratio,
                              unsigned long long* flag, hipStream_t s);
// *out += sum over vertices v and entries (h, d) of row row_of[v] of a segment
// of mix64(((h << 32) | d) ^ v * golden) (out zeroed first): an order-
// independent digest; summed over the segments it is the side's digest.
hipError_t launch_checksum(const SegView& seg, const uint32_t* row_of, uint32_t n, unsigned long long* out,
                           hipStream_t s);
// rng[2 (v k + s)], rng[2 (v k + s) + 1] = the bounds of vertex v's row (row
// row_of[v]) in segment s, for v in [0, n) and s < k (views: device array).
hipError_t launch_row_ranges(const SegView* views, uint32_t k, const uint32_t* row_of, uint32_t n, uint64_t* rng,
                             hipStream_t s);
// Labels with hub in [hlo, hhi) per vertex id (row row_of[v]) over k segments:
// with noff null, cnt[v] = their count; with noff (the exclusive scan of cnt),
// they are written to hub[] / dist[] from noff[v] on, hub-ascending.
hipError_t launch_window(const SegView* views, uint32_t k, const uint32_t* row_of, uint32_t n, uint32_t hlo,
                         uint32_t hhi, uint32_t* cnt, const uint64_t* noff, uint32_t* hub, int32_t* dist,
                         hipStream_t s);
// (hub << 32) | dist of every entry of a segment, in entry order.
hipError_t launch_decode(const SegView& seg, uint64_t count, uint64_t* out, hipStream_t s);
// What a query reads of a label handle, passed by value as a kernel argument
// (so segment fields are uniform scalar loads): per side the vertex-major range
// table of launch_row_ranges and every segment's entry arrays; both sides have
// nseg segments with the same hub ranges. fmts: SEG_F5 or SEG_F8 when every
// segment of both sides has that format, else 0.
constexpr uint32_t kQuerySeg = 16;  // segments per side a query takes (a handle merges more)
struct QuerySide {
  const uint64_t* rng;
  const uint32_t* key[kQuerySeg];
  const uint8_t* lo[kQuerySeg];
  const uint64_t* ent[kQuerySeg];
  uint32_t h0[kQuerySeg];
  uint32_t fmt[kQuerySeg];
};
struct QueryView {
  QuerySide s1, s0;
  uint32_t nseg;
  uint32_t fmts;
};
// Label merge-join query (shared_memory/labeling.hpp:56-65) over segmented
// sides; pairs hold vertex ids; an id >= n gives -1.
hipError_t launch_query(const QueryView& v, const uint32_t* pairs, uint32_t n, uint64_t q, int32_t* out,
                        hipStream_t s);
hipError_t launch_copy_u64(const uint64_t* src, uint64_t count, uint64_t* dst, hipStream_t s);

// Library primitive (hipCUB, libsort.hip) used by the device orders: radix sort of
// 64-bit key/value pairs.
hipError_t sort_pairs_u64(void* tmp, size_t& tmp_bytes, const uint64_t* kin, uint64_t* kout, const uint64_t* vin,
                          uint64_t* vout, uint64_t m, int end_bit, hipStream_t s);

}  // namespace chl
