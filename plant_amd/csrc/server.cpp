// Continuous-batching query server over device-resident labels.
//
// It extends the latency/throughput query modes of
// distributed_memory/query.hpp:62-91 (QLSN: all labels on one node, batched
// Labeling::query) into a request loop. Callers submit batches of (u, v) pairs of
// any size up to max_batch_pairs and get a ticket back at once. A server thread
// admits queued requests in arrival order as soon as a stream slot and room in
// the device pair pool are free (admission by memory), stages each request in its
// own pinned buffers, and runs copy-in, the merge-join kernel and copy-out on the
// slot's stream, so requests that arrive while others are on the GPU overlap
// with them. Completion is observed by the same thread, which timestamps it;
// latency = completion - submission.
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

#include "../../include/chl.h"
#include "dev.hpp"
#include "engine.hpp"
#include "internal.hpp"

using namespace chl;

namespace {

double now_us() {
  return std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

struct Request {
  uint64_t ticket = 0;
  std::vector<uint32_t> pairs;
  int64_t q = 0;
  int32_t* out = nullptr;
  double t_submit = 0, t_done = 0;
  uint64_t pool_off = 0;
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  uint32_t* h_pairs = nullptr;
  int32_t* h_out = nullptr;
  int64_t cap = 0;
  Request* req = nullptr;
};

// First-fit allocator of pair ranges in the device pool.
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
  const uint64_t* off1 = nullptr;
  const uint64_t* ent1 = nullptr;
  const uint64_t* off0 = nullptr;
  const uint64_t* ent0 = nullptr;
  const uint32_t* rank_of = nullptr;
  uint32_t n = 0;
  int device = 0;
  chl_server_opts o;
  std::vector<Slot> slots;
  DBuf<uint32_t> d_pairs;
  DBuf<int32_t> d_out;
  Pool pool;

  std::mutex mu;
  std::condition_variable cv;
  std::deque<Request*> pending;
  std::map<uint64_t, std::unique_ptr<Request>> all;  // ticket -> request until collected
  std::map<uint64_t, bool> done;
  uint64_t next_ticket = 1;
  bool stop = false;
  int device_error = 0;
  std::thread worker;

  // statistics since creation / last reset
  std::vector<double> lat_us;
  int64_t completed = 0, pair_count = 0, rejected = 0;
  double t_first = 0, t_last = 0;

  void loop();
  int launch(Slot& s, Request* r);
};

namespace chl {
// Device views of a label handle (defined in capi_device.cpp).
void labels_views(const chl_labels* l, const uint64_t** off1, const uint64_t** ent1, const uint64_t** off0,
                  const uint64_t** ent0, const uint32_t** rank_of, uint32_t* n, int* device);
}

int chl_server::launch(Slot& s, Request* r) {
  memcpy(s.h_pairs, r->pairs.data(), sizeof(uint32_t) * 2 * r->q);
  uint32_t* dp = d_pairs.p + 2 * r->pool_off;
  int32_t* dq = d_out.p + r->pool_off;
  CHL_HIP(hipMemcpyAsync(dp, s.h_pairs, sizeof(uint32_t) * 2 * r->q, hipMemcpyHostToDevice, s.stream));
  CHL_HIP(launch_query(off1, ent1, off0, ent0, dp, rank_of, n, (uint64_t)r->q, dq, s.stream));
  CHL_HIP(hipMemcpyAsync(s.h_out, dq, sizeof(int32_t) * r->q, hipMemcpyDeviceToHost, s.stream));
  CHL_HIP(hipEventRecord(s.done, s.stream));
  s.req = r;
  return CHL_OK;
}

void chl_server::loop() {
  (void)hipSetDevice(device);
  std::unique_lock<std::mutex> lk(mu);
  while (true) {
    bool progress = false;
    // admit in arrival order while a slot and pool room are free
    while (!pending.empty()) {
      Request* r = pending.front();
      Slot* free_slot = nullptr;
      for (auto& s : slots)
        if (!s.req) {
          free_slot = &s;
          break;
        }
      uint64_t off = 0;
      if (!free_slot || !pool.alloc((uint64_t)r->q, off)) break;
      r->pool_off = off;
      pending.pop_front();
      int rc = launch(*free_slot, r);
      if (rc) {
        device_error = rc;
        free_slot->req = r;  // completes (with the error recorded) below
      }
      progress = true;
    }
    // collect finished slots
    bool busy = false;
    for (auto& s : slots) {
      if (!s.req) continue;
      hipError_t e = device_error ? hipSuccess : hipEventQuery(s.done);
      if (e == hipErrorNotReady) {
        busy = true;
        continue;
      }
      Request* r = s.req;
      if (e != hipSuccess && !device_error) device_error = set_error(CHL_E_DEVICE, "query stream failed");
      memcpy(r->out, s.h_out, sizeof(int32_t) * r->q);
      r->t_done = now_us();
      pool.release(r->pool_off, (uint64_t)r->q);
      s.req = nullptr;
      done[r->ticket] = true;
      lat_us.push_back(r->t_done - r->t_submit);
      completed++;
      pair_count += r->q;
      if (t_first == 0) t_first = r->t_submit;
      t_last = r->t_done;
      progress = true;
    }
    if (progress) cv.notify_all();
    if (stop && pending.empty() && !busy) break;
    if (!progress) {
      if (busy) {
        lk.unlock();
        std::this_thread::yield();
        lk.lock();
      } else {
        cv.wait(lk, [&] { return stop || !pending.empty(); });
      }
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
  labels_views(l, &s->off1, &s->ent1, &s->off0, &s->ent0, &s->rank_of, &s->n, &s->device);
  CHL_HIP(hipSetDevice(s->device));
  CHL_HIP(s->d_pairs.reserve(2 * (size_t)o.max_inflight_pairs));
  CHL_HIP(s->d_out.reserve((size_t)o.max_inflight_pairs));
  s->pool.init((uint64_t)o.max_inflight_pairs);
  s->slots.resize(o.streams);
  for (auto& sl : s->slots) {
    CHL_HIP(hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking));
    CHL_HIP(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming));
    CHL_HIP(hipHostMalloc((void**)&sl.h_pairs, sizeof(uint32_t) * 2 * o.max_batch_pairs, hipHostMallocDefault));
    CHL_HIP(hipHostMalloc((void**)&sl.h_out, sizeof(int32_t) * o.max_batch_pairs, hipHostMallocDefault));
    sl.cap = o.max_batch_pairs;
  }
  chl_server* raw = s.release();
  raw->worker = std::thread([raw] { raw->loop(); });
  *out = raw;
  return CHL_OK;
}

int chl_server_submit(chl_server* s, const uint32_t* pairs, int64_t q, int32_t* out, uint64_t* ticket) {
  if (!s || !ticket || q < 0 || (q > 0 && (!pairs || !out))) return set_error(CHL_E_INVAL, "bad arguments");
  if (q > s->o.max_batch_pairs) return set_error(CHL_E_RANGE, "request of %lld pairs over max_batch_pairs", (long long)q);
  for (int64_t k = 0; k < 2 * q; ++k)
    if (pairs[k] >= s->n) return set_error(CHL_E_RANGE, "vertex %u out of range", pairs[k]);
  std::unique_ptr<Request> r(new Request());
  r->pairs.assign(pairs, pairs + 2 * q);
  r->q = q;
  r->out = out;
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->device_error) return s->device_error;
  // queued pairs beyond twice the device pool are refused (back-pressure)
  int64_t queued = 0;
  for (auto* p : s->pending) queued += p->q;
  if (queued + q > 2 * s->o.max_inflight_pairs) {
    s->rejected++;
    return set_error(CHL_E_BUSY, "admission queue full");
  }
  r->ticket = s->next_ticket++;
  r->t_submit = now_us();
  *ticket = r->ticket;
  if (q == 0) {
    r->t_done = r->t_submit;
    s->done[r->ticket] = true;
    s->all[r->ticket] = std::move(r);
    return CHL_OK;
  }
  s->pending.push_back(r.get());
  s->all[r->ticket] = std::move(r);
  s->cv.notify_all();
  return CHL_OK;
}

int chl_server_poll(chl_server* s, uint64_t ticket, int wait) {
  if (!s) return set_error(CHL_E_INVAL, "null server");
  std::unique_lock<std::mutex> lk(s->mu);
  if (!s->all.count(ticket)) return set_error(CHL_E_INVAL, "unknown ticket %llu", (unsigned long long)ticket);
  if (wait) s->cv.wait(lk, [&] { return s->done.count(ticket) > 0 || s->device_error; });
  if (s->device_error) return s->device_error;
  if (!s->done.count(ticket)) return 0;
  s->done.erase(ticket);
  s->all.erase(ticket);
  return 1;
}

int chl_server_drain(chl_server* s) {
  if (!s) return set_error(CHL_E_INVAL, "null server");
  std::unique_lock<std::mutex> lk(s->mu);
  s->cv.wait(lk, [&] {
    if (s->device_error) return true;
    if (!s->pending.empty()) return false;
    for (auto& sl : s->slots)
      if (sl.req) return false;
    return true;
  });
  return s->device_error;
}

int chl_server_stats_get(chl_server* s, chl_server_stats* st) {
  if (!s || !st) return set_error(CHL_E_INVAL, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  memset(st, 0, sizeof(*st));
  st->completed = s->completed;
  st->pairs = s->pair_count;
  st->rejected = s->rejected;
  if (!s->lat_us.empty()) {
    std::vector<double> v = s->lat_us;
    std::sort(v.begin(), v.end());
    auto pct = [&](double p) {
      size_t i = (size_t)std::min<double>(v.size() - 1, std::floor(p * (v.size() - 1) + 0.5));
      return v[i];
    };
    st->p50_us = pct(0.50);
    st->p99_us = pct(0.99);
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
  s->t_first = 0;
  s->t_last = 0;
}

void chl_server_free(chl_server* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv.notify_all();
  if (s->worker.joinable()) s->worker.join();
  (void)hipSetDevice(s->device);
  for (auto& sl : s->slots) {
    if (sl.stream) (void)hipStreamSynchronize(sl.stream);
    if (sl.done) (void)hipEventDestroy(sl.done);
    if (sl.stream) (void)hipStreamDestroy(sl.stream);
    if (sl.h_pairs) (void)hipHostFree(sl.h_pairs);
    if (sl.h_out) (void)hipHostFree(sl.h_out);
  }
  delete s;
}

}  // extern "C"
