/* C ABI of the MI355X canonical hub labeling (CHL) engine, libchl.so.
 *
 * The reference (qingdo/plant) has no FFI: its boundary is a C++ source-level API
 * plus the `hhl` command line. Each entry point below names the reference
 * interface it replaces. Conventions: opaque handles; every function returns an
 * int status (CHL_OK = 0, negative CHL_E_* on error) unless it is a pure getter;
 * no C++ exceptions cross this boundary; chl_last_error() gives a message for the
 * most recent failure on the calling thread.
 *
 * Types follow shared_memory/graph.hpp:42-47: vertex ids and hub ranks are
 * uint32, distances int32, and CHL_INFTY (INT_MAX) is the "no common hub" result.
 * Sides follow the reference: side 0 of v holds (hub rank h, d(order[h] -> v)),
 * side 1 holds (h, d(v -> order[h])) (shared_memory/akiba.hpp:112,116).
 */
#ifndef CHL_H_
#define CHL_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CHL_INFTY 2147483647

enum {
  CHL_OK = 0,
  CHL_E_INVAL = -1,   /* bad argument */
  CHL_E_IO = -2,      /* file cannot be opened or written */
  CHL_E_FORMAT = -3,  /* no reader accepts the file / malformed text */
  CHL_E_WEIGHT = -4,  /* arc length <= 0 (the reference asserts, dijkstra.hpp:91) */
  CHL_E_ORDER = -5,   /* order is not a permutation of the graph's vertices */
  CHL_E_NOMEM = -6,   /* device or host allocation failed */
  CHL_E_DEVICE = -7,  /* HIP runtime error, or no usable gfx950 device */
  CHL_E_RANGE = -8,   /* id out of range / size beyond a supported maximum */
  CHL_E_BUSY = -9,    /* server admission refused: no slot or memory free */
  CHL_E_STATE = -10   /* call not valid in the handle's current state */
};

typedef struct chl_graph chl_graph;
typedef struct chl_labels chl_labels;
typedef struct chl_server chl_server;

const char* chl_strerror(int code);
const char* chl_last_error(void);
/* Library version string and the device it will use ("" before first use). */
const char* chl_version(void);
/* Number of visible gfx950 devices (0 on a host without one). */
int chl_device_count(void);

/* ---- graph ------------------------------------------------------------------
 * Replaces hl::Graph::read(char*, bool undirected) (shared_memory/graph.hpp:327-335):
 * tries DIMACS, METIS, SNAP, edgelist in that order; returns the detected format
 * (1..4) through *format if non-null. */
int chl_graph_load(const char* path, int undirected, chl_graph** g, int* format);
/* Replaces Graph::set_n + add_arc + finalize (graph.hpp:349-353) for in-memory
 * arcs (0-based). undirected=1 adds every arc in both directions. */
int chl_graph_from_arcs(int64_t n, int64_t m, const uint32_t* src, const uint32_t* dst, const int32_t* len,
                        int undirected, chl_graph** g);
/* Replaces Graph::write (graph.hpp:338-343): DIMACS, one "a" line per out-arc. */
int chl_graph_write(const chl_graph* g, const char* path);
int64_t chl_graph_n(const chl_graph* g);
/* Replaces Graph::get_m (graph.hpp:314): arcs as the input listed them (an
 * undirected input edge counts twice), before parallel arcs are merged. */
int64_t chl_graph_m(const chl_graph* g);
/* Arcs in a range: 1 = out-arcs, 0 = in-arcs (graph.hpp:319-320). */
int64_t chl_graph_arcs(const chl_graph* g, int range);
/* 1 if every arc has an equal-length reverse arc (labels of both sides agree). */
int chl_graph_symmetric(const chl_graph* g);
/* Copies a range as CSR: off[n+1], head[arcs], len[arcs]; rows sorted by head. */
int chl_graph_csr(const chl_graph* g, int range, uint64_t* off, uint32_t* head, int32_t* len);
void chl_graph_free(chl_graph* g);

/* ---- workload generators (no reference counterpart: the reference reads
 * graphs made by external tools, README.md:24-25) ------------------------------ */
int chl_gen_grid(int64_t rows, int64_t cols, int32_t wmin, int32_t wmax, uint64_t seed, chl_graph** g);
int chl_gen_knn(int64_t n, int k, uint64_t seed, chl_graph** g);
int chl_gen_rmat(int scale, int64_t samples, double a, double b, double c, int32_t wmax, uint64_t seed,
                 chl_graph** g);

/* Locality-skewed query pairs (no reference counterpart; the serving workload
 * mixes them with uniform pairs): for each pair a uniform source u and a target
 * drawn uniformly from the vertices within `hops` out-arc hops of u. */
int chl_gen_local_pairs(const chl_graph* g, int64_t q, int hops, uint64_t seed, uint32_t* pairs);

/* ---- vertex order (order[rank] = vertex, rank 0 most important) --------------
 * Replaces hl::Order::read (shared_memory/ordering.hpp:52-67). Writes up to cap
 * ids and the file's count through *n. */
int chl_order_load(const char* path, uint32_t* order, int64_t cap, int64_t* n);
/* Replaces hl::Order::write (ordering.hpp:37-49). */
int chl_order_write(const char* path, const uint32_t* order, int64_t n);
/* Sampled-betweenness order (external tool in the reference, README.md:24). */
int chl_order_betweenness(const chl_graph* g, int64_t roots, uint64_t seed, uint32_t* order);
/* Same order, plus each vertex's summed descendant count (the score) if score != NULL. */
int chl_order_betweenness_scored(const chl_graph* g, int64_t roots, uint64_t seed, uint32_t* order, float* score);
/* The sampled-betweenness order computed on a gfx950 device: the same order
 * and scores as chl_order_betweenness_scored (score may be NULL). device_ms, if
 * not NULL, receives the device time (trees, scores and sort). No counterpart in
 * the reference, which reads the order from a file (shared_memory/hhl.cpp:59-62). */
int chl_order_betweenness_device(const chl_graph* g, int64_t roots, uint64_t seed, int device, uint32_t* order,
                                 float* score, double* device_ms);
/* Degree order, distributed_memory/degree.cpp:62-67 when reference_ties=1. */
int chl_order_degree(const chl_graph* g, uint64_t seed, int reference_ties, uint32_t* order);
/* The same degree order computed on a gfx950 device (keys and stable radix
 * sorts); replaces distributed_memory/degree.cpp:62-67 (reference_ties=1).
 * device_ms, if not NULL, receives the device time of the keys and sorts. */
int chl_order_degree_device(const chl_graph* g, uint64_t seed, int reference_ties, int device, uint32_t* order,
                            double* device_ms);

/* ---- label construction --------------------------------------------------- */
enum { CHL_MODE_GLL = 0, CHL_MODE_PLANT = 1, CHL_MODE_HYBRID = 2 };
enum { CHL_BATCH_STATIC = 0, CHL_BATCH_LEARNED = 1 };

typedef struct chl_build_opts {
  int mode;             /* CHL_MODE_*; HYBRID = PLaNT trees, then pruned trees: under static batching
                           from the batch where some root's tree/label ratio passes switch_ratio (the
                           device tests it); under learned batching where the trees predicted for the
                           next roots would cost more than pruned trees measured on a trial batch */
  int batching;         /* CHL_BATCH_STATIC (rank order) or CHL_BATCH_LEARNED (MLP-predicted packing) */
  int autotune;         /* 1 (PLaNT mode): pick per-root granularity and bucket width by timing trial trees
                           of sampled ranks before the first batch; ignored in the other modes */
  int64_t batch_roots;  /* roots per batch (0 = default) */
  int32_t delta;        /* bucket width of the delta-stepping relaxation (0 = eight times the mean arc length) */
  int32_t slots;        /* concurrent root slots = workgroups of the tree kernel (0 = default) */
  int granularity;      /* per-root parallelism: 0 = default (workgroup), 1 = wavefront, 2 = workgroup of
                           256 threads, 3 = workgroup of 1024 threads (a quarter of the roots in flight) */
  double switch_ratio;  /* ps of distributed_memory/hhl.cpp:23 (default 1024); rounded to float, as
                           hl::run's parameter is (distributed_memory/akiba.hpp:190) */
  int plant_prune;      /* 1: PLaNT trees also prune by the completed rank-prefix labels */
  int device;           /* HIP device ordinal */
  int verbose;          /* 1: per-batch lines on stderr */
  int64_t out_capacity; /* initial batch output buffer in labels (0 = automatic; it grows as needed) */
  const float* vertex_score; /* optional per-vertex importance score (e.g. sampled betweenness), a
                                feature of the learned batching predictor; NULL = not used */
  int keep_segments;    /* 1: keep the label store's segments as the build left them (no final merge);
                           0: merge them into one where device memory allows */
  int small_trees;      /* learned batching, PLaNT batches of a memory-limited build: 1 (default) runs
                           the roots predicted to grow small trees one wavefront each with a hash-table
                           search state, the others with the tree kernel; 0: the tree kernel only.
                           Labels are the same. */
  int medium_trees;     /* with small_trees: 1 also runs the roots predicted to grow medium trees 256
                           threads each with a larger hash table (7.5 GB); 0 (default): those go to the
                           tree kernel. Labels are the same. */
  int slot_memory_mb;   /* cap on the device memory of the search slots in MiB (0 = automatic: the free
                           memory beyond a reserve); a cap below what the default slot count needs makes
                           the build memory-limited, as on a large graph */
  int reserved[4];
} chl_build_opts;

typedef struct chl_build_stats {
  double lt, st, ct, total_s;   /* tree, sort, clean+append, wall seconds (akiba.hpp:214) */
  double avg_label;             /* (|L0|+|L1|)/n/2, labeling.hpp:153-158 */
  int64_t labels;               /* total entries, both sides */
  int64_t batches;
  int64_t trees_plant, trees_pruned;
  int64_t switch_rank;          /* first rank built with pruned trees (-1: none) */
  int64_t settled;              /* vertices processed over all trees (re-processing included) */
  int64_t arcs_scanned;
  int64_t candidates;           /* labels emitted before cleaning */
  int64_t prune_entries;        /* label entries read by prune checks */
  int64_t clean_entries;        /* label entries read by the cleaning merge-join */
  int64_t delta_used, slots_used, batch_used;
  double relax_ms, clean_ms;    /* summed device time of the tree kernels / cleaning + append */
  double relax_bytes;           /* algorithmic bytes of the tree kernels: the minimal work, 16 B per distinct
                                   settled vertex and per arc it scans, 8 B per prune probe and per label */
  double mlp_ms;                /* device time of MLP training/inference */
  /* learned batching: mean absolute error of log1p(tree size) per root, of the MLP
     prediction and of the static estimate (the previous batch's mean), over the
     batches the MLP predicted */
  double pred_mae_learned, pred_mae_static;
  int64_t pred_roots;           /* roots those errors cover */
  int64_t granularity_used;     /* 1 wavefront, 2 workgroup, 3 wide workgroup per root */
  double autotune_ms;           /* wall time spent by the autotuner (its trial trees included) */
  int64_t switch_trials;        /* hybrid under learned batching: pruned trial batches run */
  double policy_ms;             /* learned batching: wall time inside the policy (features, MLP, hooks) */
  int64_t tree_launches;        /* tree kernel launches (relax_ms is their summed device time) */
  double clean_bytes;           /* algorithmic bytes of the cleaning merge-join */
  double clean_kernel_ms;       /* its device time */
  int64_t settled_distinct;     /* distinct vertices settled over all trees (each tree's final bucket) */
  int64_t settled_arcs;         /* arcs those vertices scan once each */
  double relax_bytes_processed; /* relax_bytes counted over every processing instead (re-processing included) */
  double segment_ms;            /* device time building, cleaning and merging label segments */
  double store_bytes;           /* device bytes of the finished label store */
  int64_t segments;             /* segments per side of the finished store */
  int64_t slot_shrinks;         /* memory-limited builds: times search slots went back to the device */
  int64_t slots_final;          /* search slots at the last batch (slots_used: at the first) */
  double slot_s;                /* wall seconds spent giving search slots back (memory-limited builds) */
  int64_t small_roots;          /* roots whose trees the small-tree kernel completed */
  int64_t medium_roots;         /* roots whose trees the medium-tree kernel completed */
} chl_build_stats;

void chl_build_opts_default(chl_build_opts* o);
/* Replaces hl::run_paraPLL (shared_memory/akiba.hpp:132) and the hybrid hl::run
 * (distributed_memory/akiba.hpp:190) at one device: builds the canonical labels of
 * (g, order) on the GPU. Labels stay device-resident in *labels. */
int chl_build(const chl_graph* g, const uint32_t* order, int64_t n, const chl_build_opts* opts,
              chl_labels** labels, chl_build_stats* stats);

/* ---- labels ----------------------------------------------------------------- */
/* Replaces Labeling::write (shared_memory/labeling.hpp:167-183), same text;
 * rows are formatted on all host threads. */
int chl_labels_write(const chl_labels* l, const char* path);
/* Replaces Labeling::read (labeling.hpp:186-208), the same token rules (only
 * whitespace after the last label); a one-row-per-line file is parsed on all
 * host threads. The file is parsed before the device is touched, so a malformed
 * file is CHL_E_FORMAT on any host; uploads to the device. */
int chl_labels_read(const char* path, int device, chl_labels** l);
/* Binary sidecar of a label handle: the device store as it is (rows, row map,
 * shared sides), for fast save and load. No counterpart in the reference.
 * chl_labels_load checks the header, the file size, offsets, hub order and the
 * row map (CHL_E_FORMAT) before it touches the device. */
int chl_labels_save(const chl_labels* l, const char* path);
int chl_labels_load(const char* path, int device, chl_labels** l);
/* Labels from host CSR arrays (rows must be sorted by hub). */
int chl_labels_from_csr(int64_t n, const uint64_t* off0, const uint32_t* hub0, const int32_t* d0,
                        const uint64_t* off1, const uint32_t* hub1, const int32_t* d1, int device,
                        chl_labels** l);
int64_t chl_labels_n(const chl_labels* l);
int64_t chl_labels_total(const chl_labels* l, int side);
/* Replaces Labeling::get_avg (labeling.hpp:153-158). */
double chl_labels_avg(const chl_labels* l);
/* Host copy of one side: off[n+1], hub[total], dist[total]. */
int chl_labels_copy(const chl_labels* l, int side, uint64_t* off, uint32_t* hub, int32_t* dist);
/* Order-independent digest of one side, computed on the device: the sum (mod 2^64)
 * over every vertex v and entry (hub h, dist d) of v's row of
 * splitmix64_finalizer(((uint64_t)h << 32 | (uint32_t)d) ^ v * 0x9E3779B97F4A7C15).
 * Equal label sets give equal digests whatever the batch schedule. */
int chl_labels_checksum(const chl_labels* l, int side, uint64_t* sum);
/* len[v] = entries in vertex v's row of one side (no entry is copied). */
int chl_labels_row_lengths(const chl_labels* l, int side, uint32_t* len);
/* The labels of one side whose hub lies in [hub_lo, hub_hi), rows by vertex
 * id: off[n+1] always; hub[off[n]] and dist[off[n]] too when both are non-NULL
 * (call first with NULL to learn off[n]). A window of a store too large to copy
 * whole (no reference counterpart). */
int chl_labels_copy_hubs(const chl_labels* l, int side, uint32_t hub_lo, uint32_t hub_hi, uint64_t* off,
                         uint32_t* hub, int32_t* dist);
/* The segments of one side (no reference counterpart: the store's layout). A
 * side is a list of segments with ascending, disjoint hub-rank ranges
 * [h0, h1); each holds its entries in 5 bytes (fmt 5: hub - h0 < 2^24, dist <=
 * 65535) or 8 bytes (fmt 8). Writes up to cap segments; returns the count, or a
 * negative CHL_E_*. */
int chl_labels_segments(const chl_labels* l, int side, int64_t cap, uint32_t* h0, uint32_t* h1, uint32_t* fmt,
                        uint64_t* entries);
void chl_labels_free(chl_labels* l);

/* ---- queries ----------------------------------------------------------------
 * Replaces Labeling::query(u, v) (labeling.hpp:56-65) batched as batchLocalQuery
 * (distributed_memory/query_funcs.hpp:23-30): out[k] = d(pairs[2k] -> pairs[2k+1]),
 * CHL_INFTY when no hub is shared. Host pointers; synchronous. */
int chl_query(const chl_labels* l, const uint32_t* pairs, int64_t q, int32_t* out);
/* Same on device pointers, enqueued on `stream` (a hipStream_t; NULL = default);
 * returns once enqueued. A pair with an id >= n gives -1 (chl_query rejects it
 * with CHL_E_RANGE instead). */
int chl_query_device(const chl_labels* l, const uint32_t* d_pairs, int64_t q, int32_t* d_out, void* stream);
/* chl_query without any device allocation. No counterpart in the reference: it
 * is for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc). Pairs and distances are
 * staged in a fixed device array of the library (12 bytes per pair, one array
 * per device, held until the process ends) instead of buffers the handle
 * allocates; at most chl_query_noalloc_max() pairs per call (CHL_E_RANGE
 * beyond). Calls that overlap on one device share that array and race: the
 * caller serialises them. Same results and errors as chl_query otherwise. */
int chl_query_noalloc(const chl_labels* l, const uint32_t* pairs, int64_t q, int32_t* out);
int64_t chl_query_noalloc_max(void);

/* Query sources and results of distributed_memory/query.hpp:18-51:
 * chl_query_generate = generate_query: 2q values rand() % range from the C
 * library generator in its default (unseeded, seed 1) state, without touching
 * the caller's rand() state; chl_query_load = load_query: one "u v" pair per
 * line (writes up to cap pairs, the file's count through *q; a malformed or
 * over-long line is CHL_E_FORMAT); chl_query_write = write_result: one distance
 * per line. */
int chl_query_generate(int64_t q, uint32_t range, uint32_t* pairs);
int chl_query_load(const char* path, uint32_t* pairs, int64_t cap, int64_t* q);
int chl_query_write(const char* path, const int32_t* out, int64_t q);

/* ---- tree-size predictor of learned batching (no reference counterpart; it
 * predicts the (label_count, tree_size) pair of distributed_memory/akiba.hpp:99-103).
 * An 8 -> 64 -> 64 -> 2 ReLU MLP, BF16 weights, FP32 accumulation, on MFMA.
 * params: FP32 vector of chl_mlp_param_count() values laid out W1[8][64], b1[64],
 * W2[64][64], b2[64], W3[64][2], b3[2]; x: rows of 8 features; y/t: rows of 2. */
int chl_mlp_param_count(void);
int chl_mlp_forward(const float* params, const float* x, int64_t n, float* y);
/* grad = d/dparams of mean((MLP(x) - t)^2) over the n x 2 outputs; *loss = that mean */
int chl_mlp_grad(const float* params, const float* x, const float* t, int64_t n, float* grad, float* loss);
/* full-batch Adam (0.9, 0.999, 1e-8) for `steps` steps; params updated in place */
int chl_mlp_train(float* params, const float* x, const float* t, int64_t n, int steps, float lr, float* loss_curve);
/* one training step captured as a HIP graph, replayed `replays` times; before each
   replay the eager run's state is copied in, and the replay is checked against one
   eager step from that state: *max_rel = largest norm-relative difference of the
   gradient or of the updated parameters */
int chl_mlp_graph_check(const float* params, const float* x, const float* t, int64_t n, int replays, float lr,
                        float* max_rel);
/* one 32x32x8 BF16 matrix-core product D = A B (row-major; A 32x8, B 8x32 as BF16 bits) */
int chl_mfma_probe(const uint16_t* a, const uint16_t* b, float* d);

/* ---- serving (continuous batching over resident labels; extends the QLSN
 * latency/throughput modes of distributed_memory/query.hpp:62-91) ------------- */
typedef struct chl_server_opts {
  int streams;             /* concurrent device streams (default 4) */
  int64_t max_inflight_pairs; /* admission budget: pairs in flight at once (queued requests beyond twice
                                 this are refused with CHL_E_BUSY) */
  int64_t max_batch_pairs;    /* largest single request, and the most pairs one launch takes */
  int reserved[8];
} chl_server_opts;

typedef struct chl_server_stats {
  int64_t completed, pairs, rejected;
  double p50_us, p99_us, mean_us, max_us;
  double qps;              /* pairs per second since creation / last reset */
  double p90_us, p999_us;  /* more latency percentiles */
  int64_t late;            /* replayed requests submitted more than 100 us after their scheduled arrival */
  double lag_p99_us, lag_max_us; /* replayed requests: submission minus scheduled arrival */
  double reserved[3];
} chl_server_stats;

void chl_server_opts_default(chl_server_opts* o);
int chl_server_create(const chl_labels* l, const chl_server_opts* o, chl_server** s);
/* Admits a request when memory allows: copies pairs, returns a ticket at once;
 * out must stay valid until the ticket completes. CHL_E_BUSY when over budget. */
int chl_server_submit(chl_server* s, const uint32_t* pairs, int64_t q, int32_t* out, uint64_t* ticket);
/* 1 if done, 0 if pending; wait=1 blocks until done. */
int chl_server_poll(chl_server* s, uint64_t ticket, int wait);
int chl_server_drain(chl_server* s);
int chl_server_stats_get(chl_server* s, chl_server_stats* st);
void chl_server_stats_reset(chl_server* s);
/* Replays a request trace from a native thread, open loop: request i (sizes[i]
 * pairs, taken in order from `pairs`) is due arrival_us[i] microseconds after
 * the call starts and is submitted then, whatever is still in flight (a request
 * refused by admission is retried every 10 us; each refusal counts in
 * `rejected`). Latency is measured from the due time, so a late submission and
 * the wait behind refusals are part of it. Every admitted request has ended
 * when the call returns, also on an error. The results land in `out` in the
 * same order; latencies in the server statistics. */
int chl_server_replay(chl_server* s, const uint32_t* pairs, const int64_t* sizes, const double* arrival_us,
                      int64_t count, int32_t* out);
void chl_server_free(chl_server* s);

#ifdef __cplusplus
}
#endif
#endif /* CHL_H_ */
