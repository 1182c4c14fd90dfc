// Continuous-batching query server over device-resident labels.
//
// It extends the latency/throughput query modes of
// distributed_memory/query.hpp:62-91 (QLSN: all labels on one node, batched
// Labeling::query) into a request loop. Callers submit batches of (u, v) pairs of
// any size up to max_batch_pairs and get a ticket back at once. A server thread
// admits queued requests in arrival order whenever a stream slot is free: the
// slot takes every queued request that fits its pinned staging (max_batch_pairs
// in all) and the in-flight pair budget (max_inflight_pairs: admission by
// memory), and runs one merge-join launch over all of them on its stream; the
// kernel reads the pairs from and writes the distances to the slot's pinned,
// device-mapped staging, so a launch issues no copy command (asynchronous
// copies were seen to block the submitting thread for milliseconds now and
// then). Requests that arrive while others are on the GPU
// therefore join the next launch and overlap with the running ones, and the
// host cost per launch is shared by every request in it. Completion is observed
// by the same thread, which timestamps it; latency = completion - submission.
// Results are per pair, so a request gets the same distances alone or in any
// group.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include <sys/prctl.h>

#include "../../include/chl.h"
#include "dev.hpp"
#include "engine.hpp"
#include "internal.hpp"

using namespace chl;

namespace {

// A replayed request submitted this long after its scheduled arrival counts as late.
constexpr double kLateUs = 100.0;

double now_us() {
  return std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Request {
  uint64_t ticket = 0;
  uint64_t arena_off = 0;  // its pairs in the server's queue arena
  int64_t q = 0;
  int32_t* out = nullptr;
  double t_submit = 0, t_done = 0;
  uint64_t pool_off = 0;
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  // pinned, device-mapped staging: the kernel reads the pairs from h_pairs and
  // writes the distances to h_out over the bus (d_pairs / d_out are the same
  // memory as the device sees it), so a launch needs no copy commands
  uint32_t* h_pairs = nullptr;
  int32_t* h_out = nullptr;
  uint32_t* d_pairs = nullptr;
  int32_t* d_out = nullptr;
  int64_t cap = 0;                // pairs the staging holds
  bool busy = false;              // touched by the server thread only
  std::vector<Request*> group;    // requests of the running launch, in order
  uint64_t off = 0, total = 0;    // their range in the in-flight pair budget
};

// First-fit allocator of pair ranges (the in-flight budget, the queue arena).
struct Pool {
  uint64_t size = 0;
  std::map<uint64_t, uint64_t> free_;  // offset -> length
  void init(uint64_t s) {
    size = s;
    free_.clear();
    free_[0] = s;
  }
  bool alloc(uint64_t len, uint64_t& off) {
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= len) {
        off = it->first;
        uint64_t rest = it->second - len;
        free_.erase(it);
        if (rest) free_[off + len] = rest;
        return true;
      }
    }
    return false;
  }
  void release(uint64_t off, uint64_t len) {
    auto it = free_.emplace(off, len).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
};

}  // namespace

struct chl_server {
  const chl_labels* labels = nullptr;
  QueryView qv;  // the label handle's segments and range tables
  uint32_t n = 0;
  int device = 0;
  chl_server_opts o;
  std::vector<Slot> slots;

  Pool pool;

  std::mutex mu;
  std::condition_variable cv;
  std::deque<Request*> pending;
  std::map<uint64_t, Request*> all;  // ticket -> request until collected
  std::vector<Request*> spare;       // collected requests, reused
  // queued requests' pairs: a preallocated arena (no allocation per request,
  // so no page faults on the submission path)
  std::vector<uint32_t> arena;
  Pool arena_pool;
  int64_t queued_pairs = 0;
  std::map<uint64_t, bool> done;
  uint64_t next_ticket = 1;
  bool stop = false;
  std::atomic<int> device_error{0};
  int inflight = 0;  // launches on the device (under mu)
  std::thread worker;

  // statistics since creation / last reset
  std::vector<double> lat_us;
  std::vector<double> lag_us;  // replayed requests: submission minus scheduled arrival
  int64_t completed = 0, pair_count = 0, rejected = 0, late = 0;
  double t_first = 0, t_last = 0;

  void loop();
  int launch(Slot& s);
  // stops the worker (finishing what is queued and in flight) and frees every
  // slot's stream, event and pinned staging; also after a failed creation
  ~chl_server();
};

chl_server::~chl_server() {
  {
    std::lock_guard<std::mutex> lk(mu);
    stop = true;
  }
  cv.notify_all();
  if (worker.joinable()) worker.join();
  (void)hipSetDevice(device);
  for (auto& sl : slots) {
    if (sl.stream) (void)hipStreamSynchronize(sl.stream);
    if (sl.done) (void)hipEventDestroy(sl.done);
    if (sl.stream) (void)hipStreamDestroy(sl.stream);
    if (sl.h_pairs) (void)hipHostFree(sl.h_pairs);
    if (sl.h_out) (void)hipHostFree(sl.h_out);
  }
  for (auto& kv : all) delete kv.second;
  for (Request* r : spare) delete r;
}

namespace chl {
// Device views of a label handle (defined in capi_device.cpp).
void labels_views(const chl_labels* l, QueryView* qv, uint32_t* n, int* device);
}

int chl_server::launch(Slot& s) {
  uint64_t pos = 0;
  for (Request* r : s.group) {
    memcpy(s.h_pairs + 2 * pos, arena.data() + 2 * r->arena_off, sizeof(uint32_t) * 2 * r->q);
    pos += r->q;
  }
  {
    std::lock_guard<std::mutex> lk(mu);
    for (Request* r : s.group) arena_pool.release(r->arena_off, (uint64_t)r->q);
  }
  CHL_HIP(launch_query(qv, s.d_pairs, n, s.total, s.d_out, s.stream));
  CHL_HIP(hipEventRecord(s.done, s.stream));
  return CHL_OK;
}

void chl_server::loop() {
  (void)hipSetDevice(device);
  // short timed waits below: ask for 1 us timer slack instead of the default 50
  (void)prctl(PR_SET_TIMERSLACK, 1000UL, 0, 0, 0);
  for (;;) {
    bool progress = false;
    // admission: a free slot takes the queued requests, in arrival order, that
    // fit its staging and one free range of the in-flight budget
    if (device_error) {
      // nothing more is admitted: queued requests end with the error
      std::lock_guard<std::mutex> lk(mu);
      for (Request* r : pending) {
        arena_pool.release(r->arena_off, (uint64_t)r->q);
        done[r->ticket] = true;
      }
      queued_pairs = 0;
      if (!pending.empty()) progress = true;
      pending.clear();
    }
    for (auto& s : slots) {
      if (s.busy || device_error) continue;
      {
        std::lock_guard<std::mutex> lk(mu);
        if (pending.empty()) break;
        size_t take = 0;
        uint64_t tot = 0;
        for (Request* r : pending) {
          if (tot + (uint64_t)r->q > (uint64_t)s.cap) break;
          tot += (uint64_t)r->q;
          ++take;
        }
        uint64_t off = 0;
        while (take > 0 && !pool.alloc(tot, off)) tot -= (uint64_t)pending[--take]->q;
        if (take == 0) break;  // the budget is spent: wait for completions
        s.group.assign(pending.begin(), pending.begin() + take);
        pending.erase(pending.begin(), pending.begin() + take);
        queued_pairs -= (int64_t)tot;
        s.off = off;
        s.total = tot;
        ++inflight;
      }
      s.busy = true;
      if (launch(s) != CHL_OK && !device_error) device_error = CHL_E_DEVICE;
      progress = true;
    }
    // completion
    for (auto& s : slots) {
      if (!s.busy) continue;
      hipError_t e = hipEventQuery(s.done);
      if (e == hipErrorNotReady) {
        if (!device_error) continue;
        // after an error elsewhere the slot is released only once its launch
        // has ended: its kernel may still read and write the staging
        e = hipStreamSynchronize(s.stream);
      }
      const double t = now_us();
      uint64_t pos = 0;
      for (Request* r : s.group) {
        memcpy(r->out, s.h_out + pos, sizeof(int32_t) * r->q);
        pos += (uint64_t)r->q;
        r->t_done = t;
      }
      {
        std::lock_guard<std::mutex> lk(mu);
        if (e != hipSuccess && !device_error) device_error = set_error(CHL_E_DEVICE, "query stream failed");
        pool.release(s.off, s.total);
        for (Request* r : s.group) {
          done[r->ticket] = true;
          lat_us.push_back(r->t_done - r->t_submit);
          completed++;
          pair_count += r->q;
          if (t_first == 0 || r->t_submit < t_first) t_first = r->t_submit;
          t_last = r->t_done;
        }
        --inflight;
      }
      cv.notify_all();
      s.busy = false;
      s.group.clear();
      progress = true;
    }
    std::unique_lock<std::mutex> lk(mu);
    if (stop && pending.empty() && inflight == 0) break;
    if (!progress) {
      // nothing to do now: sleep until a submission, or briefly while launches
      // are in flight (no spinning: a busy host thread can get its process
      // throttled for milliseconds)
      if (inflight > 0) cv.wait_for(lk, std::chrono::microseconds(10));
      else cv.wait(lk, [&] { return stop || !pending.empty(); });
    }
  }
}

extern "C" {

void chl_server_opts_default(chl_server_opts* o) {
  if (!o) return;
  memset(o, 0, sizeof(*o));
  o->streams = 4;
  o->max_inflight_pairs = (int64_t)1 << 20;
  o->max_batch_pairs = 65536;
}

int chl_server_create(const chl_labels* l, const chl_server_opts* opts, chl_server** out) {
  if (!l || !out) return set_error(CHL_E_INVAL, "null argument");
  *out = nullptr;
  chl_server_opts o;
  if (opts) o = *opts;
  else chl_server_opts_default(&o);
  if (o.streams < 1 || o.streams > 32 || o.max_batch_pairs < 1 || o.max_inflight_pairs < o.max_batch_pairs)
    return set_error(CHL_E_INVAL, "bad server options");
  std::unique_ptr<chl_server> s(new chl_server());
  s->labels = l;
  s->o = o;
  labels_views(l, &s->qv, &s->n, &s->device);
  CHL_HIP(hipSetDevice(s->device));
  s->pool.init((uint64_t)o.max_inflight_pairs);
  // queue arena: twice the pool (the back-pressure limit) plus one request;
  // touched once here so that later submissions do not fault pages in
  const uint64_t arena_pairs = 2 * (uint64_t)o.max_inflight_pairs + (uint64_t)o.max_batch_pairs;
  s->arena.assign(2 * arena_pairs, 0u);
  s->arena_pool.init(arena_pairs);
  s->slots.resize(o.streams);
  for (auto& sl : s->slots) {
    CHL_HIP(hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking));
    CHL_HIP(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
    // staging for the requests of one launch: up to max_batch_pairs in all (or
    // one request of that size), so a launch stays short and queued requests
    // wait at most about one launch
    sl.cap = o.max_batch_pairs;
    // coherent: the kernel's reads and writes go to host memory uncached, so
    // the host sees the distances once the launch's event has completed
    CHL_HIP(hipHostMalloc((void**)&sl.h_pairs, sizeof(uint32_t) * 2 * sl.cap, hipHostMallocMapped | hipHostMallocCoherent));
    CHL_HIP(hipHostMalloc((void**)&sl.h_out, sizeof(int32_t) * sl.cap, hipHostMallocMapped | hipHostMallocCoherent));
    CHL_HIP(hipHostGetDevicePointer((void**)&sl.d_pairs, sl.h_pairs, 0));
    CHL_HIP(hipHostGetDevicePointer((void**)&sl.d_out, sl.h_out, 0));
  }
  // one tiny launch on every stream now, so that no request pays for a
  // stream's first use (device queue set-up takes milliseconds)
  for (auto& sl : s->slots) {
    sl.h_pairs[0] = sl.h_pairs[1] = 0;
    CHL_HIP(launch_query(s->qv, sl.d_pairs, s->n, 1, sl.d_out, sl.stream));
    CHL_HIP(hipStreamSynchronize(sl.stream));
  }
  chl_server* raw = s.release();
  raw->worker = std::thread([raw] { raw->loop(); });
  *out = raw;
  return CHL_OK;
}

}  // extern "C"

namespace {
// Queues a request whose latency is measured from t_arrival (now for a direct
// submission, the scheduled arrival for a replayed one).
int submit_at(chl_server* s, const uint32_t* pairs, int64_t q, int32_t* out, uint64_t* ticket, double t_arrival) {
  if (!s || !ticket || q < 0 || (q > 0 && (!pairs || !out))) return set_error(CHL_E_INVAL, "bad arguments");
  if (q > s->o.max_batch_pairs) return set_error(CHL_E_RANGE, "request of %lld pairs over max_batch_pairs", (long long)q);
  for (int64_t k = 0; k < 2 * q; ++k)
    if (pairs[k] >= s->n) return set_error(CHL_E_RANGE, "vertex %u out of range", pairs[k]);
  const double t_submit = t_arrival;
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->device_error) return s->device_error;
  // queued pairs beyond twice the device pool are refused (back-pressure)
  uint64_t aoff = 0;
  if (s->queued_pairs + q > 2 * s->o.max_inflight_pairs || (q > 0 && !s->arena_pool.alloc((uint64_t)q, aoff))) {
    s->rejected++;
    return set_error(CHL_E_BUSY, "admission queue full");
  }
  Request* r = nullptr;
  if (!s->spare.empty()) {
    r = s->spare.back();
    s->spare.pop_back();
  } else {
    r = new Request();
  }
  *r = Request();
  r->arena_off = aoff;
  r->q = q;
  r->out = out;
  if (q) memcpy(s->arena.data() + 2 * aoff, pairs, sizeof(uint32_t) * 2 * q);
  r->ticket = s->next_ticket++;
  r->t_submit = t_submit;
  *ticket = r->ticket;
  s->all[r->ticket] = r;
  if (q == 0) {
    r->t_done = r->t_submit;
    s->done[r->ticket] = true;
    return CHL_OK;
  }
  s->queued_pairs += q;
  s->pending.push_back(r);
  s->cv.notify_all();
  return CHL_OK;
}
}  // namespace

extern "C" {

int chl_server_submit(chl_server* s, const uint32_t* pairs, int64_t q, int32_t* out, uint64_t* ticket) {
  return submit_at(s, pairs, q, out, ticket, now_us());
}

int chl_server_poll(chl_server* s, uint64_t ticket, int wait) {
  if (!s) return set_error(CHL_E_INVAL, "null server");
  std::unique_lock<std::mutex> lk(s->mu);
  if (!s->all.count(ticket)) return set_error(CHL_E_INVAL, "unknown ticket %llu", (unsigned long long)ticket);
  if (wait) s->cv.wait(lk, [&] { return s->done.count(ticket) > 0 || s->device_error; });
  if (s->device_error) return s->device_error;
  if (!s->done.count(ticket)) return 0;
  s->done.erase(ticket);
  auto it = s->all.find(ticket);
  s->spare.push_back(it->second);
  s->all.erase(it);
  return 1;
}

int chl_server_drain(chl_server* s) {
  if (!s) return set_error(CHL_E_INVAL, "null server");
  std::unique_lock<std::mutex> lk(s->mu);
  s->cv.wait(lk, [&] { return s->pending.empty() && s->inflight == 0; });
  return s->device_error;
}

int chl_server_stats_get(chl_server* s, chl_server_stats* st) {
  if (!s || !st) return set_error(CHL_E_INVAL, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  memset(st, 0, sizeof(*st));
  st->completed = s->completed;
  st->pairs = s->pair_count;
  st->rejected = s->rejected;
  st->late = s->late;
  if (!s->lag_us.empty()) {
    std::vector<double> v = s->lag_us;
    std::sort(v.begin(), v.end());
    st->lag_p99_us = v[(size_t)std::min<double>(v.size() - 1, std::floor(0.99 * (v.size() - 1) + 0.5))];
    st->lag_max_us = v.back();
  }
  if (!s->lat_us.empty()) {
    std::vector<double> v = s->lat_us;
    std::sort(v.begin(), v.end());
    auto pct = [&](double p) {
      size_t i = (size_t)std::min<double>(v.size() - 1, std::floor(p * (v.size() - 1) + 0.5));
      return v[i];
    };
    st->p50_us = pct(0.50);
    st->p99_us = pct(0.99);
    st->p90_us = pct(0.90);
    st->p999_us = pct(0.999);
    st->max_us = v.back();
    double sum = 0;
    for (double x : v) sum += x;
    st->mean_us = sum / v.size();
    if (s->t_last > s->t_first) st->qps = (double)s->pair_count / ((s->t_last - s->t_first) * 1e-6);
  }
  return CHL_OK;
}

void chl_server_stats_reset(chl_server* s) {
  if (!s) return;
  std::lock_guard<std::mutex> lk(s->mu);
  s->lat_us.clear();
  s->completed = 0;
  s->pair_count = 0;
  s->rejected = 0;
  s->late = 0;
  s->lag_us.clear();
  s->t_first = 0;
  s->t_last = 0;
}

int chl_server_replay(chl_server* s, const uint32_t* pairs, const int64_t* sizes, const double* arrival_us,
                      int64_t count, int32_t* out) {
  if (!s || count < 0 || (count > 0 && (!pairs || !sizes || !arrival_us || !out)))
    return set_error(CHL_E_INVAL, "bad arguments");
  std::vector<uint64_t> tickets;
  tickets.reserve((size_t)count);
  const unsigned long slack = (unsigned long)prctl(PR_GET_TIMERSLACK, 0, 0, 0, 0);
  (void)prctl(PR_SET_TIMERSLACK, 1000UL, 0, 0, 0);
  const double t0 = now_us();
  int64_t pos = 0;
  int rc = CHL_OK;
  std::vector<double> lags;
  lags.reserve((size_t)count);
  int64_t late = 0;
  for (int64_t i = 0; i < count && rc == CHL_OK; ++i) {
    // open loop: wait for the arrival time, not for earlier completions (sleep,
    // then yield for the last few microseconds)
    const double due = t0 + arrival_us[i];
    for (double ahead; (ahead = due - now_us()) > 0;) {
      if (ahead > 30) std::this_thread::sleep_for(std::chrono::microseconds((int64_t)(ahead - 20)));
      else std::this_thread::yield();
    }
    uint64_t t = 0;
    for (;;) {
      // latency counts from the scheduled arrival, so time spent behind a
      // late submission or a refused admission is part of it
      rc = submit_at(s, pairs + 2 * pos, sizes[i], out + pos, &t, due);
      if (rc != CHL_E_BUSY) break;
      std::this_thread::sleep_for(std::chrono::microseconds(10));
    }
    if (rc != CHL_OK) break;
    const double lag = now_us() - due;
    lags.push_back(lag);
    if (lag > kLateUs) ++late;
    tickets.push_back(t);
    pos += sizes[i];
  }
  (void)prctl(PR_SET_TIMERSLACK, slack, 0, 0, 0);
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->lag_us.insert(s->lag_us.end(), lags.begin(), lags.end());
    s->late += late;
  }
  // every admitted request ends before the call returns, also on an error, so
  // nothing writes into `out` afterwards
  const int drc = chl_server_drain(s);
  if (rc == CHL_OK) rc = drc;
  for (uint64_t t : tickets) {
    const int prc = chl_server_poll(s, t, 1);
    if (prc < 0 && rc == CHL_OK) rc = prc;
  }
  return rc;
}

void chl_server_free(chl_server* s) { delete s; }

}  // extern "C"
