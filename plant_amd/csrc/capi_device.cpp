// C ABI: label construction, label handles and distance queries.
#include <sys/stat.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <vector>

#include "../../include/chl.h"
#include "dev.hpp"
#include "engine.hpp"
#include "internal.hpp"

using namespace chl;

struct chl_labels {
  LabelStore store;
  QueryView qv;                  // what queries read (set once the range tables exist)
  DBuf<uint32_t> rank_of;        // vertex id -> row id
  std::vector<uint32_t> row_of;  // host copy of rank_of
  // query staging for host-pointer calls
  DBuf<uint32_t> qpairs;
  DBuf<int32_t> qout;
  hipStream_t stream = nullptr;
  ~chl_labels() {
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

// Row map and the per-segment row-range tables the queries read (per vertex id
// the (begin, end) of its row in each segment).
int make_handle(chl_labels* l, const std::vector<uint32_t>& row_of) {
  l->row_of = row_of;
  CHL_HIP(l->rank_of.reserve(std::max<size_t>(row_of.size(), 1)));
  if (!row_of.empty())
    CHL_HIP(hipMemcpy(l->rank_of.p, row_of.data(), sizeof(uint32_t) * row_of.size(), hipMemcpyHostToDevice));
  CHL_HIP(hipStreamCreateWithFlags(&l->stream, hipStreamNonBlocking));
  int rc = l->store.build_ranges(l->rank_of.p, l->stream);
  if (rc) return rc;
  l->qv = l->store.query_view();
  return CHL_OK;
}

// Host rows in vertex order from the device store.
int to_host(const chl_labels* l, int side, std::vector<uint64_t>& off, std::vector<uint64_t>& ent) {
  std::vector<uint64_t> roff, rent;
  int rc = l->store.download(side, roff, rent);
  if (rc) return rc;
  const int64_t n = l->store.n;
  off.assign(n + 1, 0);
  for (int64_t v = 0; v < n; ++v) {
    uint32_t r = l->row_of[v];
    off[v + 1] = off[v] + (roff[r + 1] - roff[r]);
  }
  ent.resize(off[n]);
  parallel_for(n, host_threads(), 4096, [&](int64_t v) {
    uint32_t r = l->row_of[v];
    std::copy(rent.begin() + roff[r], rent.begin() + roff[r + 1], ent.begin() + off[v]);
  });
  return CHL_OK;
}

bool rows_sorted(const std::vector<uint64_t>& off, const std::vector<uint64_t>& ent) {
  for (size_t v = 0; v + 1 < off.size(); ++v)
    for (uint64_t k = off[v] + 1; k < off[v + 1]; ++k)
      if ((ent[k] >> 32) <= (ent[k - 1] >> 32)) return false;
  return true;
}

int from_host_labels(const Labels& H, int device, chl_labels** out) {
  for (int s = 0; s < 2; ++s)
    if (!rows_sorted(H.off[s], H.ent[s])) return set_error(CHL_E_FORMAT, "label rows must be hub-ascending");
  int rc = ensure_device(device);
  if (rc) return rc;
  chl_labels* l = new chl_labels();
  rc = l->store.init(H.n, false, device);
  for (int s = 0; s < 2 && rc == CHL_OK; ++s) rc = l->store.upload(s, H.off[s], H.ent[s]);
  if (rc == CHL_OK) rc = l->store.sync_views();
  std::vector<uint32_t> ident(H.n);
  for (int64_t v = 0; v < H.n; ++v) ident[v] = (uint32_t)v;
  if (rc == CHL_OK) rc = make_handle(l, ident);
  if (rc) {
    delete l;
    return rc;
  }
  *out = l;
  return CHL_OK;
}

}  // namespace

namespace chl {
void labels_views(const chl_labels* l, QueryView* qv, uint32_t* n, int* device) {
  *qv = l->qv;
  *n = (uint32_t)l->store.n;
  *device = l->store.device;
}
hipStream_t labels_stream(const chl_labels* l) { return l->stream; }
}  // namespace chl

extern "C" {

const char* chl_version(void) { return "chl-mi355x 0.1 (gfx950)"; }

int chl_device_count(void) {
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess) return 0;
  return c;
}

void chl_build_opts_default(chl_build_opts* o) {
  if (!o) return;
  memset(o, 0, sizeof(*o));
  o->mode = CHL_MODE_PLANT;
  o->batching = CHL_BATCH_STATIC;
  o->autotune = 0;
  o->batch_roots = 0;
  o->delta = 0;
  o->slots = 0;
  o->small_trees = 1;
  o->medium_trees = 0;
  o->switch_ratio = 1024.0;
  o->plant_prune = 0;
  o->device = 0;
}

int chl_build(const chl_graph* g, const uint32_t* order, int64_t n, const chl_build_opts* opts, chl_labels** labels,
              chl_build_stats* stats) {
  if (!g || !order || !labels) return set_error(CHL_E_INVAL, "null argument");
  *labels = nullptr;
  if (n != g->g.n)
    return set_error(CHL_E_ORDER, "order has %lld entries, graph has %lld vertices", (long long)n, (long long)g->g.n);
  std::vector<uint32_t> ord(order, order + n), rank(n, 0xFFFFFFFFu);
  for (int64_t i = 0; i < n; ++i) {
    if (ord[i] >= (uint64_t)n || rank[ord[i]] != 0xFFFFFFFFu)
      return set_error(CHL_E_ORDER, "order is not a permutation (entry %lld = %u)", (long long)i, ord[i]);
    rank[ord[i]] = (uint32_t)i;
  }
  for (int r = 0; r < 2; ++r)
    for (int32_t w : g->g.len[r])
      if (w <= 0) return set_error(CHL_E_WEIGHT, "arc length %d <= 0", w);
  chl_build_opts o;
  if (opts) o = *opts;
  else chl_build_opts_default(&o);
  if (o.mode < CHL_MODE_GLL || o.mode > CHL_MODE_HYBRID) return set_error(CHL_E_INVAL, "bad mode %d", o.mode);
  if (o.switch_ratio <= 0) o.switch_ratio = 1024.0;
  // hl::run takes the ratio as a float and widens it to double for the test
  // (distributed_memory/akiba.hpp:190, 206): a ratio of 0.1 is 0.100000001490116.
  o.switch_ratio = (double)(float)o.switch_ratio;
  int rc = ensure_device(o.device);
  if (rc) return rc;
  if (stats) memset(stats, 0, sizeof(*stats));
  chl_labels* l = new chl_labels();
  chl_graph* gm = const_cast<chl_graph*>(g);  // the device copy is a cache inside the handle
  if (!gm->dev) gm->dev.reset(new DeviceGraph());
  rc = run_build(g->g, *gm->dev, ord, o, stats, l->store);
  if (rc == CHL_OK) rc = make_handle(l, rank);
  if (rc) {
    delete l;
    return rc;
  }
  *labels = l;
  return CHL_OK;
}

int chl_labels_write(const chl_labels* l, const char* path) {
  if (!l || !path) return set_error(CHL_E_INVAL, "null argument");
  Labels H;
  H.n = l->store.n;
  for (int s = 0; s < 2; ++s) {
    int rc = to_host(l, s, H.off[s], H.ent[s]);
    if (rc) return rc;
  }
  int rc = labels_write(H, path);
  return rc == CHL_OK ? rc : set_error(rc, "cannot write labels to %s", path);
}

int chl_labels_read(const char* path, int device, chl_labels** out) {
  if (!path || !out) return set_error(CHL_E_INVAL, "null argument");
  *out = nullptr;
  Labels H;
  int rc = labels_read(H, path);
  if (rc) return set_error(rc, "cannot read labels from %s", path);
  return from_host_labels(H, device, out);
}

// ---- binary sidecar ----------------------------------------------------------
// Little-endian: "CHLLABv1", u64 n, u32 shared sides, u32 0, u64 total[2], then
// per stored side off u64[n+1] and ent u64[total], then row_of u32[n] (vertex
// -> row). The store is saved as it sits on the device, rows in row-id order.
namespace {
constexpr char kBinMagic[8] = {'C', 'H', 'L', 'L', 'A', 'B', 'v', '1'};
struct BinHeader {
  char magic[8];
  uint64_t n;
  uint32_t shared;
  uint32_t zero;
  uint64_t total[2];
};
static_assert(sizeof(BinHeader) == 40, "binary label header layout");

bool write_all(FILE* f, const void* p, size_t bytes) { return bytes == 0 || fwrite(p, 1, bytes, f) == bytes; }
bool read_all(FILE* f, void* p, size_t bytes) { return bytes == 0 || fread(p, 1, bytes, f) == bytes; }
}  // namespace

int chl_labels_save(const chl_labels* l, const char* path) {
  if (!l || !path) return set_error(CHL_E_INVAL, "null argument");
  const LabelStore& S = l->store;
  BinHeader h;
  memcpy(h.magic, kBinMagic, 8);
  h.n = (uint64_t)S.n;
  h.shared = S.shared_sides ? 1u : 0u;
  h.zero = 0;
  h.total[0] = S.side_total(0);
  h.total[1] = S.shared_sides ? 0 : S.side_total(1);
  FILE* f = fopen(path, "wb");
  if (!f) return set_error(CHL_E_IO, "cannot open %s", path);
  bool ok = write_all(f, &h, sizeof(h));
  for (int s = 0; s < (S.shared_sides ? 1 : 2) && ok; ++s) {
    std::vector<uint64_t> off, ent;
    int rc = S.download(s, off, ent);
    if (rc) {
      fclose(f);
      return rc;
    }
    ok = write_all(f, off.data(), sizeof(uint64_t) * off.size()) && write_all(f, ent.data(), sizeof(uint64_t) * ent.size());
  }
  ok = ok && write_all(f, l->row_of.data(), sizeof(uint32_t) * l->row_of.size());
  if (fclose(f) != 0) ok = false;
  return ok ? CHL_OK : set_error(CHL_E_IO, "cannot write labels to %s", path);
}

int chl_labels_load(const char* path, int device, chl_labels** out) {
  if (!path || !out) return set_error(CHL_E_INVAL, "null argument");
  *out = nullptr;
  FILE* f = fopen(path, "rb");
  if (!f) return set_error(CHL_E_IO, "cannot open %s", path);
  struct Closer {
    FILE* f;
    ~Closer() { fclose(f); }
  } closer{f};
  BinHeader h;
  if (!read_all(f, &h, sizeof(h)) || memcmp(h.magic, kBinMagic, 8) != 0 || h.shared > 1 || h.zero != 0 ||
      h.n >= 0xFFFFFFFFull || (h.shared && h.total[1] != 0))
    return set_error(CHL_E_FORMAT, "%s is not a binary label file", path);
  const int sides = h.shared ? 1 : 2;
  // the file size must be exactly what the header announces (counts checked
  // against the size first, so that the sum cannot wrap)
  struct stat st;
  if (fstat(fileno(f), &st) != 0) return set_error(CHL_E_IO, "cannot stat %s", path);
  const unsigned long long fsize = (unsigned long long)st.st_size;
  if (h.n > fsize / 4 || h.total[0] > fsize / 8 || h.total[1] > fsize / 8)
    return set_error(CHL_E_FORMAT, "%s: size does not match its header", path);
  unsigned long long want = sizeof(h) + 4ull * h.n;
  for (int s = 0; s < sides; ++s) want += 8ull * (h.n + 1) + 8ull * h.total[s];
  if (fsize != want) return set_error(CHL_E_FORMAT, "%s: size does not match its header", path);
  const int64_t n = (int64_t)h.n;
  std::vector<uint64_t> off[2], ent[2];
  for (int s = 0; s < sides; ++s) {
    off[s].resize(n + 1);
    ent[s].resize(h.total[s]);
    if (!read_all(f, off[s].data(), 8 * off[s].size()) || !read_all(f, ent[s].data(), 8 * ent[s].size()))
      return set_error(CHL_E_IO, "cannot read %s", path);
    if (off[s][0] != 0 || off[s][n] != h.total[s])
      return set_error(CHL_E_FORMAT, "%s: row offsets do not span the entries", path);
    for (int64_t v = 0; v < n; ++v)
      if (off[s][v + 1] < off[s][v]) return set_error(CHL_E_FORMAT, "%s: row offsets decrease", path);
    if (!rows_sorted(off[s], ent[s])) return set_error(CHL_E_FORMAT, "%s: label rows must be hub-ascending", path);
  }
  std::vector<uint32_t> row_of(n);
  if (!read_all(f, row_of.data(), 4 * row_of.size())) return set_error(CHL_E_IO, "cannot read %s", path);
  {
    std::vector<uint8_t> seen(n, 0);
    for (uint32_t r : row_of) {
      if ((int64_t)r >= n || seen[r]) return set_error(CHL_E_FORMAT, "%s: the row map is not a permutation", path);
      seen[r] = 1;
    }
  }
  int rc = ensure_device(device);
  if (rc) return rc;
  chl_labels* l = new chl_labels();
  rc = l->store.init(n, h.shared != 0, device);
  for (int s = 0; s < sides && rc == CHL_OK; ++s) rc = l->store.upload(s, off[s], ent[s]);
  if (rc == CHL_OK) rc = l->store.sync_views();
  if (rc == CHL_OK) rc = make_handle(l, row_of);
  if (rc) {
    delete l;
    return rc;
  }
  *out = l;
  return CHL_OK;
}

int chl_labels_from_csr(int64_t n, const uint64_t* off0, const uint32_t* hub0, const int32_t* d0, const uint64_t* off1,
                        const uint32_t* hub1, const int32_t* d1, int device, chl_labels** out) {
  if (!out || n <= 0 || !off0 || !off1) return set_error(CHL_E_INVAL, "bad arguments");
  *out = nullptr;
  Labels H;
  H.n = n;
  const uint64_t* offs[2] = {off0, off1};
  const uint32_t* hubs[2] = {hub0, hub1};
  const int32_t* ds[2] = {d0, d1};
  for (int s = 0; s < 2; ++s) {
    H.off[s].assign(offs[s], offs[s] + n + 1);
    const uint64_t T = H.off[s][n];
    if (T && (!hubs[s] || !ds[s])) return set_error(CHL_E_INVAL, "null label arrays");
    H.ent[s].resize(T);
    for (uint64_t k = 0; k < T; ++k) {
      if (ds[s][k] < 0) return set_error(CHL_E_INVAL, "negative label distance");
      H.ent[s][k] = ((uint64_t)hubs[s][k] << 32) | (uint32_t)ds[s][k];
    }
  }
  return from_host_labels(H, device, out);
}

int64_t chl_labels_n(const chl_labels* l) { return l ? l->store.n : -1; }
int64_t chl_labels_total(const chl_labels* l, int side) {
  if (!l || side < 0 || side > 1) return -1;
  return (int64_t)l->store.side_total(side);
}
double chl_labels_avg(const chl_labels* l) {
  if (!l || l->store.n == 0) return 0.0;
  return (double)(l->store.side_total(0) + l->store.side_total(1)) / (double)l->store.n / 2.0;
}

int chl_labels_copy(const chl_labels* l, int side, uint64_t* off, uint32_t* hub, int32_t* dist) {
  if (!l || side < 0 || side > 1 || !off) return set_error(CHL_E_INVAL, "bad arguments");
  std::vector<uint64_t> o, e;
  int rc = to_host(l, side, o, e);
  if (rc) return rc;
  memcpy(off, o.data(), sizeof(uint64_t) * o.size());
  parallel_for((int64_t)e.size(), host_threads(), (int64_t)1 << 20, [&](int64_t k) {
    if (hub) hub[k] = (uint32_t)(e[k] >> 32);
    if (dist) dist[k] = (int32_t)(uint32_t)e[k];
  });
  return CHL_OK;
}

int chl_labels_row_lengths(const chl_labels* l, int side, uint32_t* len) {
  if (!l || side < 0 || side > 1 || !len) return set_error(CHL_E_INVAL, "bad arguments");
  const int64_t n = l->store.n;
  std::vector<uint64_t> roff(n + 1), tot(n, 0);
  CHL_HIP(hipSetDevice(l->store.device));
  for (auto& g : l->store.side(side)) {
    CHL_HIP(hipMemcpy(roff.data(), g->off.p, sizeof(uint64_t) * (n + 1), hipMemcpyDeviceToHost));
    for (int64_t r = 0; r < n; ++r) tot[r] += roff[r + 1] - roff[r];
  }
  for (int64_t v = 0; v < n; ++v) len[v] = (uint32_t)tot[l->row_of[v]];
  return CHL_OK;
}

int chl_labels_copy_hubs(const chl_labels* l, int side, uint32_t hub_lo, uint32_t hub_hi, uint64_t* off,
                         uint32_t* hub, int32_t* dist) {
  if (!l || side < 0 || side > 1 || !off || hub_hi < hub_lo) return set_error(CHL_E_INVAL, "bad arguments");
  const int64_t n = l->store.n;
  CHL_HIP(hipSetDevice(l->store.device));
  DBuf<uint32_t> cnt;
  DBuf<uint64_t> noff, tmp;
  CHL_HIP(cnt.reserve(std::max<int64_t>(n, 1)));
  CHL_HIP(noff.reserve(n + 1));
  CHL_HIP(tmp.reserve(scan_tmp_words((uint64_t)n)));
  const SegView* views = l->store.d_views(side);
  const uint32_t k = l->store.nseg();
  hipStream_t st = l->stream;
  if (k == 0) CHL_HIP(launch_fill_u32(cnt.p, 0, (size_t)n, st));
  CHL_HIP(launch_window(views, k, l->rank_of.p, (uint32_t)n, hub_lo, hub_hi, cnt.p, nullptr, nullptr, nullptr, st));
  CHL_HIP(launch_scan_u32(cnt.p, (uint64_t)n, noff.p, tmp.p, st));
  CHL_HIP(hipMemcpyAsync(off, noff.p, sizeof(uint64_t) * (n + 1), hipMemcpyDeviceToHost, st));
  CHL_HIP(hipStreamSynchronize(st));
  const uint64_t total = off[n];
  if (!hub || !dist || total == 0) return CHL_OK;  // sizes only
  DBuf<uint32_t> dh;
  DBuf<int32_t> dd;
  CHL_HIP(dh.reserve(total));
  CHL_HIP(dd.reserve(total));
  CHL_HIP(launch_window(views, k, l->rank_of.p, (uint32_t)n, hub_lo, hub_hi, cnt.p, noff.p, dh.p, dd.p, st));
  CHL_HIP(hipMemcpyAsync(hub, dh.p, sizeof(uint32_t) * total, hipMemcpyDeviceToHost, st));
  CHL_HIP(hipMemcpyAsync(dist, dd.p, sizeof(int32_t) * total, hipMemcpyDeviceToHost, st));
  CHL_HIP(hipStreamSynchronize(st));
  return CHL_OK;
}

int chl_labels_segments(const chl_labels* l, int side, int64_t cap, uint32_t* h0, uint32_t* h1, uint32_t* fmt,
                        uint64_t* entries) {
  if (!l || side < 0 || side > 1 || cap < 0) return set_error(CHL_E_INVAL, "bad arguments");
  const auto& S = l->store.side(side);
  for (size_t i = 0; i < S.size() && (int64_t)i < cap; ++i) {
    if (h0) h0[i] = S[i]->h0;
    if (h1) h1[i] = S[i]->h1;
    if (fmt) fmt[i] = S[i]->fmt;
    if (entries) entries[i] = S[i]->total;
  }
  return (int)S.size();
}

int chl_labels_checksum(const chl_labels* l, int side, uint64_t* sum) {
  if (!l || side < 0 || side > 1 || !sum) return set_error(CHL_E_INVAL, "bad arguments");
  CHL_HIP(hipSetDevice(l->store.device));
  DBuf<unsigned long long> acc;
  CHL_HIP(acc.reserve(1));
  CHL_HIP(launch_fill_u64((uint64_t*)acc.p, 0, 1, nullptr));
  for (auto& g : l->store.side(side))
    CHL_HIP(launch_checksum(g->view(), l->rank_of.p, (uint32_t)l->store.n, acc.p, nullptr));
  unsigned long long h = 0;
  CHL_HIP(hipMemcpy(&h, acc.p, sizeof(h), hipMemcpyDeviceToHost));
  *sum = h;
  return CHL_OK;
}

void chl_labels_free(chl_labels* l) { delete l; }

int chl_order_betweenness_device(const chl_graph* g, int64_t roots, uint64_t seed, int device, uint32_t* order,
                                 float* score, double* device_ms) {
  if (!g || !order || roots <= 0) return set_error(CHL_E_INVAL, "bad arguments");
  chl_graph* gm = const_cast<chl_graph*>(g);  // the device copy is a cache inside the handle
  if (!gm->dev) gm->dev.reset(new DeviceGraph());
  std::vector<uint32_t> o;
  std::vector<uint64_t> sc;
  int rc = order_betweenness_device(g->g, *gm->dev, device, roots, seed, o, score ? &sc : nullptr, device_ms);
  if (rc) return rc;
  if (!o.empty()) memcpy(order, o.data(), sizeof(uint32_t) * o.size());
  if (score)
    for (size_t v = 0; v < sc.size(); ++v) score[v] = (float)sc[v];
  return CHL_OK;
}

int chl_order_degree_device(const chl_graph* g, uint64_t seed, int reference_ties, int device, uint32_t* order,
                            double* device_ms) {
  if (!g || !order) return set_error(CHL_E_INVAL, "bad arguments");
  std::vector<uint32_t> o;
  int rc = order_degree_device(g->g, device, seed, reference_ties != 0, o, device_ms);
  if (rc) return rc;
  if (!o.empty()) memcpy(order, o.data(), sizeof(uint32_t) * o.size());
  return CHL_OK;
}

int chl_query_device(const chl_labels* l, const uint32_t* d_pairs, int64_t q, int32_t* d_out, void* stream) {
  if (!l || q < 0 || (q > 0 && (!d_pairs || !d_out))) return set_error(CHL_E_INVAL, "bad arguments");
  if (q == 0) return CHL_OK;
  CHL_HIP(hipSetDevice(l->store.device));
  CHL_HIP(launch_query(l->qv, d_pairs, (uint32_t)l->store.n, (uint64_t)q, d_out, (hipStream_t)stream));
  return CHL_OK;
}

int chl_query(const chl_labels* lc, const uint32_t* pairs, int64_t q, int32_t* out) {
  chl_labels* l = const_cast<chl_labels*>(lc);
  if (!l || q < 0 || (q > 0 && (!pairs || !out))) return set_error(CHL_E_INVAL, "bad arguments");
  if (q == 0) return CHL_OK;
  for (int64_t k = 0; k < 2 * q; ++k)
    if ((int64_t)pairs[k] >= l->store.n) return set_error(CHL_E_RANGE, "vertex %u out of range", pairs[k]);
  CHL_HIP(hipSetDevice(l->store.device));
  CHL_HIP(l->qpairs.reserve(2 * q));
  CHL_HIP(l->qout.reserve(q));
  CHL_HIP(hipMemcpyAsync(l->qpairs.p, pairs, sizeof(uint32_t) * 2 * q, hipMemcpyHostToDevice, l->stream));
  CHL_HIP(launch_query(l->qv, l->qpairs.p, (uint32_t)l->store.n, (uint64_t)q, l->qout.p, l->stream));
  CHL_HIP(hipMemcpyAsync(out, l->qout.p, sizeof(int32_t) * q, hipMemcpyDeviceToHost, l->stream));
  CHL_HIP(hipStreamSynchronize(l->stream));
  return CHL_OK;
}

}  // extern "C"
