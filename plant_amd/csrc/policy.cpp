// Batching policies of the label build.
//
// Static (CHL_BATCH_STATIC): the builder's own schedule, contiguous rank ranges
// of doubling width taken in rank order, the GPU analogue of the reference's
// supersteps (shared_memory/akiba.hpp:152-185, distributed_memory/akiba.hpp:259,480).
//
// Learned (CHL_BATCH_LEARNED): a small MLP (mlp.hip) predicts each remaining
// root's tree size and label count from cheap per-root features and is trained
// online on the sizes every finished batch measures. The predictions choose the
// batch boundary (enough predicted work that the largest tree is a small part
// of it), the order of roots inside the batch (largest predicted tree first, so
// the persistent workgroups finish together) and the batch output capacity
// (predicted labels, so batches rarely spill). In hybrid mode they also place
// the PLaNT-to-pruned switch (predicted tree/label ratio, confirmed by a timed
// trial batch). Batches stay contiguous rank
// ranges, so the labels are the same under either policy. Until the first batch
// has been measured the static schedule is used, and the static estimate (the
// previous batch's mean) is kept beside the MLP's to report both errors.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <random>
#include <vector>

#include "dev.hpp"
#include "engine.hpp"
#include "internal.hpp"
#include "mlp.hpp"

namespace chl {
namespace {

struct Learned {
  const Graph& g;
  const std::vector<uint32_t>& order;
  chl_build_opts o;
  chl_build_stats* stats;
  int64_t n;
  hipStream_t stream = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  DBuf<float> X, P, M, V, G, Y, loss, T;
  DBuf<uint16_t> wb;
  DBuf<uint32_t> tidx, mb;
  std::vector<float> hT;
  std::vector<uint32_t> hidx;
  uint64_t tcap = 65536, tcount = 0;
  int step = 0;
  bool trained = false;
  int64_t pred_r0 = -1;
  uint32_t pred_cnt = 0;
  std::vector<float> pred;  // [pred_cnt][2], log1p units
  std::vector<float> lin;   // [pred_cnt][2], expm1 of pred
  double last_mean_log = -1.0;
  double mae_l = 0, mae_s = 0;
  int64_t mae_n = 0;
  double mlp_ms = 0;
  std::mt19937_64 rng{12345};
  const BuildHooks* hooks = nullptr;  // the builder writes its slot count there
  // labels one batch may write: the free device memory at 32 bytes per label
  // (the tree output's 12, the batch segment's 5 to 8, cleaning flags and the
  // compacted copy), at most 2^31 - 1; next_width keeps a 2x margin on the
  // predicted total
  double label_budget = 268435456.0;

  // Hybrid switch point. PLaNT work per root grows with its tree/label ratio,
  // while a pruned tree costs about its labels times the label merge-join; the
  // GPU ratio of the two costs depends on the graph (road-like graphs keep PLaNT
  // cheaper to the end, scale-free graphs want pruned trees after the first
  // thousand roots). A batch whose predicted tree/label ratio (sum of predicted
  // tree sizes over sum of predicted labels) reaches theta runs pruned trees as a
  // trial; batches are cut where a block of kBlock roots crosses theta. The
  // trial's measured device time against the PLaNT time predicted for its roots
  // (measured time per tree vertex of the last PLaNT batch times the predicted
  // tree sizes) makes the switch final, or doubles theta.
  static constexpr uint32_t kBlock = 256;
  bool hybrid = false, switched = false, trial_pending = false;
  double theta = 16.0, ms_per_vertex = 0.0, trial_plant_ms = 0.0, trial_ratio = 0.0;
  int64_t decided_r0 = -1;
  bool decided_trial = false;
  uint32_t last_width = 0;  // width of the last batch this policy sized
  int64_t trials = 0;
  double host_ms = 0;  // wall time inside the policy
  // Sums of many per-root predictions made in log space come out low (the mean
  // of exp is above exp of the mean); the measured over predicted sums of the
  // pilot and of every finished batch scale the predicted sums back.
  double calib_t = 1.0, calib_l = 1.0;
  std::vector<float> hxs;    // host copy of the features, rows by rank
  std::vector<float> lrank;  // log1p(rank), the trend's abscissa
  DBuf<float> Xp, Yp;      // gathered rows and their predictions

  Learned(const Graph& gg, const std::vector<uint32_t>& ord, const chl_build_opts& opts, chl_build_stats* s)
      : g(gg), order(ord), o(opts), stats(s), n(gg.n), hybrid(opts.mode == CHL_MODE_HYBRID) {}
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int graph_B = -1;
  DBuf<int> stepd;  // Adam step counter in device memory

  ~Learned() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (stream) (void)hipStreamDestroy(stream);
  }

  int init() {
    CHL_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    CHL_HIP(hipEventCreate(&e0));
    CHL_HIP(hipEventCreate(&e1));
    {
      size_t freeb = 0, totalb = 0;
      CHL_HIP(hipMemGetInfo(&freeb, &totalb));
      update_budget(freeb);
    }
    // features by rank: log rank, log degree, log 2-hop degree sum, log score
    // (on all host threads; sums in fixed chunks added in order, so the
    // features do not depend on the thread schedule)
    const int nt = host_threads();
    std::vector<double> f1(n), f2(n), f3(n, 0.0);
    std::vector<uint64_t> deg(n);
    parallel_for(n, nt, 16384, [&](int64_t v) {
      deg[v] = (g.off[1][v + 1] - g.off[1][v]) + (g.off[0][v + 1] - g.off[0][v]);
    });
    parallel_for(n, nt, 4096, [&](int64_t v) {
      uint64_t two = 0;
      for (uint64_t a = g.off[1][v]; a < g.off[1][v + 1]; ++a) two += deg[g.head[1][a]];
      f1[v] = std::log1p((double)deg[v]);
      f2[v] = std::log1p((double)two);
      if (o.vertex_score) f3[v] = std::log1p(std::max(0.0, (double)o.vertex_score[v]));
    });
    constexpr int64_t kChunk = 65536;
    const int64_t nchunks = (n + kChunk - 1) / kChunk;
    auto chunk_sum = [&](auto term) {
      std::vector<double> part(nchunks, 0.0);
      parallel_for(nchunks, nt, 1, [&](int64_t c) {
        double a = 0;
        for (int64_t v = c * kChunk; v < std::min<int64_t>(n, (c + 1) * kChunk); ++v) a += term(v);
        part[c] = a;
      });
      double total = 0;
      for (double x : part) total += x;
      return total;
    };
    auto standardize = [&](std::vector<double>& f) {
      const double mu = chunk_sum([&](int64_t v) { return f[v]; }) / (double)n;
      const double var = chunk_sum([&](int64_t v) { return (f[v] - mu) * (f[v] - mu); });
      const double sd = std::sqrt(var / (double)n);
      parallel_for(n, nt, 16384, [&](int64_t v) { f[v] = sd > 0 ? (f[v] - mu) / sd : 0.0; });
    };
    standardize(f1);
    standardize(f2);
    standardize(f3);
    std::vector<float> hx((size_t)n * kMlpF, 0.0f);
    const double lr = std::log1p((double)n);
    parallel_for(n, nt, 16384, [&](int64_t r) {
      const uint32_t v = order[r];
      float* row = &hx[(size_t)r * kMlpF];
      row[0] = (float)(std::log1p((double)r) / lr);
      row[1] = (float)f1[v];
      row[2] = (float)f2[v];
      row[3] = (float)f3[v];
    });
    CHL_HIP(X.reserve(hx.size()));
    CHL_HIP(hipMemcpy(X.p, hx.data(), sizeof(float) * hx.size(), hipMemcpyHostToDevice));
    hxs.swap(hx);
    lrank.resize(n);
    parallel_for(n, host_threads(), 16384, [&](int64_t r) { lrank[r] = (float)std::log1p((double)r); });
    // parameters: uniform(+-sqrt(6/(fan_in+fan_out))) weights, zero biases
    std::vector<float> hp(kMlpParams, 0.0f);
    auto fill = [&](int off, int fin, int fout) {
      std::uniform_real_distribution<float> u(-std::sqrt(6.0f / (fin + fout)), std::sqrt(6.0f / (fin + fout)));
      for (int i = 0; i < fin * fout; ++i) hp[off + i] = u(rng);
    };
    fill(kMlpOffW1, 4, kMlpH);
    for (int k = 4; k < kMlpF; ++k)
      for (int c = 0; c < kMlpH; ++c) hp[kMlpOffW1 + k * kMlpH + c] = 0.0f;
    fill(kMlpOffW2, kMlpH, kMlpH);
    fill(kMlpOffW3, kMlpH, kMlpO);
    CHL_HIP(P.reserve(kMlpParams));
    CHL_HIP(M.reserve(kMlpParams));
    CHL_HIP(V.reserve(kMlpParams));
    CHL_HIP(G.reserve(kMlpParams));
    CHL_HIP(wb.reserve(kMlpParams));
    CHL_HIP(loss.reserve(1));
    CHL_HIP(hipMemcpy(P.p, hp.data(), sizeof(float) * kMlpParams, hipMemcpyHostToDevice));
    CHL_HIP(hipMemset(M.p, 0, sizeof(float) * kMlpParams));
    CHL_HIP(hipMemset(V.p, 0, sizeof(float) * kMlpParams));
    CHL_HIP(mlp_sync(P.p, wb.p, stream));
    CHL_HIP(stepd.reserve(1));
    CHL_HIP(launch_fill_u32((uint32_t*)stepd.p, 0, 1, stream));
    CHL_HIP(tidx.reserve(tcap));
    CHL_HIP(T.reserve(tcap * kMlpO));
    CHL_HIP(mb.reserve(1024));
    CHL_HIP(Y.reserve(65536 * kMlpO));
    hT.resize(tcap * kMlpO);
    hidx.resize(tcap);
    CHL_HIP(hipStreamSynchronize(stream));
    return CHL_OK;
  }

  void timed_begin() { (void)hipEventRecord(e0, stream); }
  void timed_end() {
    (void)hipEventRecord(e1, stream);
    (void)hipEventSynchronize(e1);
    float t = 0;
    (void)hipEventElapsedTime(&t, e0, e1);
    mlp_ms += t;
  }

  // Predictions for ranks [r0, r0 + cnt).
  int predict(int64_t r0, uint32_t cnt) {
    if (pred_r0 == r0 && pred_cnt >= cnt) return CHL_OK;
    CHL_HIP(Y.reserve((size_t)cnt * kMlpO));
    timed_begin();
    CHL_HIP(mlp_forward(X.p + (size_t)r0 * kMlpF, (int)cnt, wb.p, P.p, Y.p, stream));
    pred.resize((size_t)cnt * kMlpO);
    CHL_HIP(hipMemcpyAsync(pred.data(), Y.p, sizeof(float) * cnt * kMlpO, hipMemcpyDeviceToHost, stream));
    timed_end();
    // add the trend; keep the linear values (tree size, labels) beside the logs
    lin.resize((size_t)cnt * kMlpO);
    const float a0 = (float)ta[0], a1 = (float)ta[1], b0 = (float)tb[0], b1 = (float)tb[1];
    parallel_for(cnt, host_threads(), 16384, [&](int64_t k) {
      const float x = lrank[r0 + k];
      float* y = &pred[(size_t)k * kMlpO];
      y[0] += a0 + b0 * x;
      y[1] += a1 + b1 * x;
      lin[(size_t)k * kMlpO] = std::expm1(std::max(0.0f, y[0]));
      lin[(size_t)k * kMlpO + 1] = std::expm1(std::max(0.0f, y[1]));
    });
    pred_r0 = r0;
    pred_cnt = cnt;
    return CHL_OK;
  }

  float pred_tree(uint32_t rank) const { return pred[(size_t)(rank - pred_r0) * kMlpO]; }
  float pred_labels(uint32_t rank) const { return pred[(size_t)(rank - pred_r0) * kMlpO + 1]; }
  float lin_tree(uint32_t rank) const { return lin[(size_t)(rank - pred_r0) * kMlpO]; }
  float lin_labels(uint32_t rank) const { return lin[(size_t)(rank - pred_r0) * kMlpO + 1]; }

  // Predicted sums of tree sizes and labels over ranks [a, b) (inside the
  // predicted range).
  void pred_sums(int64_t a, int64_t b, double& tree, double& labels) const {
    tree = 0;
    labels = 0;
    for (int64_t r = a; r < b; ++r) {
      tree += lin_tree((uint32_t)r);
      labels += lin_labels((uint32_t)r);
    }
  }
  double pred_ratio(int64_t a, int64_t b) const {
    double t, l;
    pred_sums(a, b, t, l);
    return t * calib_t / std::max(1.0, l * calib_l);
  }

  void update_budget(size_t freeb) {
    label_budget = std::min<double>(2147483647.0, std::max<double>(1 << 22, (double)freeb / 32.0));
  }

  uint32_t next_width(int64_t r0, uint32_t static_width, uint32_t max_width) {
    (void)static_width;
    decided_r0 = -1;
    {
      // the budget follows the free memory as the label store grows
      size_t freeb = 0, totalb = 0;
      if (hipMemGetInfo(&freeb, &totalb) == hipSuccess) update_budget(freeb);
    }
    if (!trained || switched) return 0;
    // at most a quarter of the vertices and eight times the previous width per
    // batch, so that an error of the predicted label total (heavy-tailed sums
    // from a small sample) costs a few more batches, not the device memory
    int64_t cap = std::min<int64_t>(max_width, std::max<int64_t>(65536, n / 4));
    if (last_width) cap = std::min<int64_t>(cap, std::max<int64_t>(8 * (int64_t)last_width, 4096));
    const uint32_t cnt = (uint32_t)std::min<int64_t>(cap, n - r0);
    if (predict(r0, cnt)) return 0;
    // the widest batch whose predicted labels fit the output budget (every batch
    // boundary is a device-wide wait for the batch's last tree), and at least one
    // root per slot
    const uint32_t slots = std::max<uint32_t>(1, hooks ? hooks->slots : 1);
    if (hybrid && r0 > 0 && ms_per_vertex > 0 &&
        pred_ratio(r0, r0 + std::min<uint32_t>(kBlock, cnt)) >= theta) {
      // a trial batch of pruned trees: two roots per slot
      decided_r0 = r0;
      decided_trial = true;
      return std::min<uint32_t>(cnt, 2 * slots);
    }
    double labels = 0;
    uint32_t R = 0;
    while (R < cnt) {
      labels += calib_l * lin_labels((uint32_t)(r0 + R));
      if (2.0 * labels > label_budget && R >= slots) break;  // 2x margin on the prediction
      // hybrid: end the batch before the first block predicted to cross theta
      if (hybrid && R >= slots && R % kBlock == 0 && R + kBlock <= cnt &&
          pred_ratio(r0 + R, r0 + R + kBlock) >= theta)
        break;
      ++R;
    }
    decided_r0 = r0;
    decided_trial = false;
    last_width = R;
    return R;
  }

  // Hybrid: the kind of trees of the batch [r0, r0 + R).
  int plant_next(int64_t r0, uint32_t R, int* final) {
    *final = 0;
    if (switched) {
      *final = 1;
      return 0;
    }
    if (!trained || ms_per_vertex <= 0) return 1;
    if (predict(r0, R)) return 1;
    double t, l;
    pred_sums(r0, r0 + R, t, l);
    t *= calib_t;
    l *= calib_l;
    const bool trial = decided_r0 == r0 ? decided_trial : t / std::max(1.0, l) >= theta;
    if (!trial) return 1;
    trial_pending = true;
    trial_plant_ms = ms_per_vertex * t;
    trial_ratio = t / std::max(1.0, l);
    ++trials;
    return 0;
  }

  // Stable counting sort of roots by predicted output o (log1p units, 1/32
  // steps), largest first: an approximate longest-processing-time order in O(R).
  void lpt_order(uint32_t* roots, uint32_t R, int o) {
    constexpr int kBuckets = 1024;
    std::vector<uint32_t> cnt(kBuckets + 1, 0), key(R), out(R);
    for (uint32_t k = 0; k < R; ++k) {
      const float y = pred[(size_t)(roots[k] - pred_r0) * kMlpO + o];
      const int b = (int)std::min<float>(kBuckets - 1, std::max(0.0f, y * 32.0f));
      key[k] = (uint32_t)(kBuckets - 1 - b);  // descending
      ++cnt[key[k] + 1];
    }
    for (int b = 0; b < kBuckets; ++b) cnt[b + 1] += cnt[b];
    for (uint32_t k = 0; k < R; ++k) out[cnt[key[k]]++] = roots[k];
    std::memcpy(roots, out.data(), sizeof(uint32_t) * R);
  }

  void order_roots(uint32_t* roots, uint32_t R) {
    if (!trained || pred_r0 < 0 || roots[0] < pred_r0 || roots[0] + R > pred_r0 + pred_cnt) return;
    if (switched || trial_pending)  // pruned trees: their work follows their labels
      lpt_order(roots, R, 1);
    else
      lpt_order(roots, R, 0);
  }

  // The whole budget once the batch is predicted to need more than a few million
  // labels (the buffer is kept across batches), else twice the prediction.
  uint64_t out_estimate(const uint32_t* roots, uint32_t R) {
    if (!trained || pred_r0 < 0) return 0;
    double s = 0;
    for (uint32_t k = 0; k < R; ++k) {
      if (roots[k] < pred_r0 || roots[k] >= pred_r0 + pred_cnt) return 0;
      s += lin_labels(roots[k]);
    }
    s *= calib_l;
    return (uint64_t)std::min(label_budget, std::max(2.0 * s, s > 4e6 ? label_budget : 0.0)) + 4096;
  }

  // Small-tree routing: a root goes to the small-tree kernel when its predicted
  // tree (processed vertices) is at most `limit`. The caller's limit leaves room
  // for the prediction's error (its table holds several times the limit); a root
  // predicted small whose tree outgrows the table runs again on the tree kernel.
  uint32_t route_small(const uint32_t* roots, uint32_t R, uint32_t limit, uint8_t* big) {
    if (!trained || switched || trial_pending || pred_r0 < 0) return 0;
    uint32_t small = 0;
    for (uint32_t k = 0; k < R; ++k) {
      if (roots[k] < pred_r0 || roots[k] >= pred_r0 + pred_cnt) return 0;
      const bool b = lin_tree(roots[k]) > (float)limit;
      big[k] = b ? 1 : 0;
      small += b ? 0 : 1;
    }
    return small;
  }

  // Pilot sample: 512 ranks log-uniform over [max(1024, n/32), n) (trees there are
  // cheap on road-like graphs), so the first real batch is already packed by
  // prediction; the top ranks are left to the trend. The launch lasts as long as
  // its largest tree: on the 1M road-like graph 42 ms from n/128, 17 ms from n/32
  // (build 1.033-1.051 against 1.020-1.031 s, interleaved on one device; labels and
  // batches the same; neutral at 16M). Not in hybrid mode: on
  // scale-free graphs low-ranked PLaNT trees cover most of the graph, and the
  // pilot's slowest tree would cost as much as the first batch; there the first
  // batch (the top ranks, static width) trains the network.
#ifndef CHL_PILOT_DIV
#define CHL_PILOT_DIV 32
#endif
  static constexpr int64_t kPilotDiv = CHL_PILOT_DIV;
  uint32_t pilot(uint32_t* roots, uint32_t cap) {
    if (hybrid) return 0;
    const int64_t lo = std::min<int64_t>(std::max<int64_t>(1024, n / kPilotDiv), n / 2);
    if (n - lo < 64) return 0;
    const uint32_t R = (uint32_t)std::min<int64_t>(std::min<int64_t>(cap, 512), n - lo);
    std::uniform_real_distribution<double> u(std::log((double)lo + 1.0), std::log((double)n));
    for (uint32_t k = 0; k < R; ++k) {
      int64_t r = (int64_t)std::floor(std::exp(u(rng))) - 1;
      roots[k] = (uint32_t)std::min<int64_t>(n - 1, std::max<int64_t>(lo, r));
    }
    return R;
  }

  static double clamp_calib(double c) { return std::isfinite(c) ? std::min(100.0, std::max(0.5, c)) : 1.0; }

  // Predicted tree and label sums (linear units, before calibration) of the
  // roots at the given ranks.
  int pred_sums_at(const uint32_t* ranks, uint32_t R, double& tsum, double& lsum) {
    tsum = 0;
    lsum = 0;
    if (!R) return CHL_OK;
    std::vector<float> rows((size_t)R * kMlpF);
    for (uint32_t k = 0; k < R; ++k)
      std::memcpy(&rows[(size_t)k * kMlpF], &hxs[(size_t)ranks[k] * kMlpF], sizeof(float) * kMlpF);
    CHL_HIP(Xp.reserve(rows.size()));
    CHL_HIP(Yp.reserve((size_t)R * kMlpO));
    CHL_HIP(hipMemcpyAsync(Xp.p, rows.data(), sizeof(float) * rows.size(), hipMemcpyHostToDevice, stream));
    timed_begin();
    CHL_HIP(mlp_forward(Xp.p, (int)R, wb.p, P.p, Yp.p, stream));
    timed_end();
    std::vector<float> y((size_t)R * kMlpO);
    CHL_HIP(hipMemcpy(y.data(), Yp.p, sizeof(float) * y.size(), hipMemcpyDeviceToHost));
    for (uint32_t k = 0; k < R; ++k) {
      tsum += std::expm1(std::max(0.0, (double)y[(size_t)k * kMlpO] + trend(0, ranks[k])));
      lsum += std::expm1(std::max(0.0, (double)y[(size_t)k * kMlpO + 1] + trend(1, ranks[k])));
    }
    return CHL_OK;
  }

  int on_pilot(const uint32_t* roots, const uint32_t* ts, const uint32_t* lc, uint32_t R) {
    // the static estimate of the first batch: the sample mean
    double mean_log = 0;
    for (uint32_t k = 0; k < R; ++k) mean_log += std::log1p((double)ts[k]);
    last_mean_log = R ? mean_log / R : -1.0;
    for (uint32_t k = 0; k < R; ++k) {
      const uint64_t pos = tcount % tcap;
      hidx[pos] = roots[k];
      hT[pos * kMlpO] = (float)std::log1p((double)ts[k]);
      hT[pos * kMlpO + 1] = (float)std::log1p((double)lc[k]);
      ++tcount;
    }
    fit_trend();
    for (int rep = 0; rep < 3; ++rep) {
      int rc = train();
      if (rc) return rc;
    }
    double pt, pl, mt = 0, ml = 0;
    int rc = pred_sums_at(roots, R, pt, pl);
    if (rc) return rc;
    for (uint32_t k = 0; k < R; ++k) {
      mt += ts[k];
      ml += lc[k];
    }
    calib_t = clamp_calib(mt / pt);
    calib_l = clamp_calib(ml / pl);
    return CHL_OK;
  }

  int on_batch(const uint32_t* roots, const uint32_t* ts, const uint32_t* lc, uint32_t R, uint32_t flags, double ms) {
    if (!(flags & TREE_FLAGS)) {
      // pruned trees measure no PLaNT sizes; a trial decides the switch
      if (trial_pending) {
        trial_pending = false;
        if (ms < trial_plant_ms) switched = true;
        else theta = 2.0 * std::max(theta, trial_ratio);
        if (o.verbose)
          fprintf(stderr, "switch trial: pruned %.3f ms, PLaNT predicted %.3f ms (ratio %.1f): %s\n", ms,
                  trial_plant_ms, trial_ratio, switched ? "switch" : "stay");
      }
      return CHL_OK;
    }
    // per-batch sums over the roots on all host threads (fixed chunks added in
    // order: the same result on any schedule)
    std::vector<double> y(R);
    parallel_for(R, host_threads(), 16384, [&](int64_t k) { y[k] = std::log1p((double)ts[k]); });
    auto sums = [&](auto term, int w, double* out) {
      constexpr uint32_t kChunk = 65536;
      const uint32_t nc = (R + kChunk - 1) / kChunk;
      std::vector<double> part((size_t)nc * w, 0.0);
      parallel_for(nc, host_threads(), 1, [&](int64_t c) {
        for (uint32_t k = (uint32_t)c * kChunk; k < std::min<uint32_t>(R, ((uint32_t)c + 1) * kChunk); ++k)
          term(k, &part[(size_t)c * w]);
      });
      for (int i = 0; i < w; ++i) out[i] = 0;
      for (uint32_t c = 0; c < nc; ++c)
        for (int i = 0; i < w; ++i) out[i] += part[(size_t)c * w + i];
    };
    double base[2];
    sums([&](uint32_t k, double* a) { a[0] += ts[k]; a[1] += y[k]; }, 2, base);
    if (base[0] > 0) ms_per_vertex = ms / base[0];
    // prediction error before learning from this batch
    const double mean_log = base[1] / std::max<uint32_t>(R, 1);
    if (trained && last_mean_log >= 0) {
      bool covered = true;
      for (uint32_t k = 0; k < R && covered; ++k)
        covered = roots[k] >= pred_r0 && roots[k] < pred_r0 + pred_cnt;
      if (covered) {
        double a[6];  // |error| learned, |error| static, predicted tree, predicted labels, measured both
        const double lm = last_mean_log;
        sums([&](uint32_t k, double* p) {
          p[0] += std::fabs(pred_tree(roots[k]) - y[k]);
          p[1] += std::fabs(lm - y[k]);
          p[2] += lin_tree(roots[k]);
          p[3] += lin_labels(roots[k]);
          p[4] += ts[k];
          p[5] += lc[k];
        }, 6, a);
        mae_l += a[0];
        mae_s += a[1];
        mae_n += R;
        calib_t = clamp_calib(a[4] / a[2]);
        calib_l = clamp_calib(a[5] / a[3]);
      }
    }
    last_mean_log = mean_log;
    // append (rank, targets) to the sample ring; training draws from the newest
    // 16384 samples only, so a wider batch contributes an even stride of its roots
    const uint32_t take = std::min<uint32_t>(R, 16384);
    for (uint32_t i = 0; i < take; ++i) {
      const uint32_t k = (uint32_t)((uint64_t)i * R / take);
      const uint64_t pos = tcount % tcap;
      hidx[pos] = roots[k];
      hT[pos * kMlpO] = (float)y[k];
      hT[pos * kMlpO + 1] = (float)std::log1p((double)lc[k]);
      ++tcount;
    }
    fit_trend();
    return train();
  }

  // Least-squares line of each target against log1p(rank) over the most recent
  // samples. The MLP learns the residual from this trend, so that predictions for
  // ranks beyond the measured ones follow the trend instead of extrapolating the
  // network.
  double ta[kMlpO] = {0, 0}, tb[kMlpO] = {0, 0};
  double trend(int o, uint32_t rank) const { return ta[o] + tb[o] * std::log1p((double)rank); }
  void fit_trend() {
    const uint64_t avail = std::min<uint64_t>(tcount, tcap);
    const uint64_t recent = std::min<uint64_t>(avail, 16384);
    for (int o = 0; o < kMlpO; ++o) {
      double sx = 0, sy = 0, sxx = 0, sxy = 0;
      for (uint64_t b = 0; b < recent; ++b) {
        const uint64_t pos = (tcount - 1 - b) % tcap;
        const double x = std::log1p((double)hidx[pos]), y = hT[pos * kMlpO + o];
        sx += x;
        sy += y;
        sxx += x * x;
        sxy += x * y;
      }
      const double m = (double)recent, den = m * sxx - sx * sx;
      tb[o] = den > 1e-9 ? (m * sxy - sx * sy) / den : 0.0;
      ta[o] = (sy - tb[o] * sx) / m;
    }
  }

  // kSteps Adam steps on minibatches of up to 512 samples drawn from the most
  // recent 16384; the minibatches (sample ranks and targets) are drawn on the
  // host from the mirror of the sample ring and uploaded in one copy each.
  static constexpr int kSteps = 60;
  int train() {
    const uint64_t avail = std::min<uint64_t>(tcount, tcap);
    const uint64_t recent = std::min<uint64_t>(avail, 16384);
    const int B = (int)std::min<uint64_t>(512, avail);
    std::vector<uint32_t> rk((size_t)kSteps * B);
    std::vector<float> tv((size_t)kSteps * B * kMlpO);
    for (size_t k = 0; k < rk.size(); ++k) {
      const uint64_t back = rng() % recent;  // 0 = newest
      const uint64_t pos = (tcount - 1 - back) % tcap;
      rk[k] = hidx[pos];
      for (int o = 0; o < kMlpO; ++o) tv[k * kMlpO + o] = (float)(hT[pos * kMlpO + o] - trend(o, hidx[pos]));
    }
    CHL_HIP(mb.reserve((size_t)kSteps * 512));
    CHL_HIP(T.reserve(std::max<size_t>((size_t)kSteps * 512 * kMlpO, tcap * kMlpO)));
    CHL_HIP(hipMemcpyAsync(mb.p, rk.data(), sizeof(uint32_t) * rk.size(), hipMemcpyHostToDevice, stream));
    CHL_HIP(hipMemcpyAsync(T.p, tv.data(), sizeof(float) * tv.size(), hipMemcpyHostToDevice, stream));
    // the kSteps steps are one captured graph, replayed for every training
    // call with the same minibatch size (inputs live in fixed device buffers)
    if (!exec || graph_B != B) {
      if (exec) (void)hipGraphExecDestroy(exec);
      if (graph) (void)hipGraphDestroy(graph);
      exec = nullptr;
      graph = nullptr;
      CHL_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
      for (int s = 0; s < kSteps; ++s) {
        CHL_HIP(launch_fill_u32((uint32_t*)G.p, 0, kMlpParams, stream));
        CHL_HIP(launch_fill_u32((uint32_t*)loss.p, 0, 1, stream));
        CHL_HIP(mlp_grad(X.p, T.p + (size_t)s * B * kMlpO, mb.p + (size_t)s * B, B, wb.p, P.p, G.p, loss.p, stream));
        CHL_HIP(mlp_adam(P.p, G.p, M.p, V.p, wb.p, 3e-3f, stepd.p, stream));
      }
      CHL_HIP(hipStreamEndCapture(stream, &graph));
      CHL_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
      graph_B = B;
    }
    timed_begin();
    CHL_HIP(hipGraphLaunch(exec, stream));
    timed_end();
    step += kSteps;
    trained = true;
    pred_r0 = -1;  // weights changed
    return CHL_OK;
  }
};

// wall time spent inside the policy (host work, copies and the MLP launches)
struct HostTimer {
  double& acc;
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  explicit HostTimer(void* c) : acc(((Learned*)c)->host_ms) {}
  ~HostTimer() { acc += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count(); }
};

uint32_t hk_next_width(void* c, int64_t r0, uint32_t sw, uint32_t mw) {
  HostTimer t(c);
  return ((Learned*)c)->next_width(r0, sw, mw);
}
void hk_order(void* c, uint32_t* roots, uint32_t R) {
  HostTimer t(c);
  ((Learned*)c)->order_roots(roots, R);
}
uint64_t hk_out(void* c, const uint32_t* roots, uint32_t R) {
  HostTimer t(c);
  return ((Learned*)c)->out_estimate(roots, R);
}
uint32_t hk_small(void* c, const uint32_t* roots, uint32_t R, uint32_t limit, uint8_t* big) {
  HostTimer t(c);
  return ((Learned*)c)->route_small(roots, R, limit, big);
}
uint32_t hk_pilot(void* c, uint32_t* roots, uint32_t cap) {
  HostTimer t(c);
  return ((Learned*)c)->pilot(roots, cap);
}
void hk_on_pilot(void* c, const uint32_t* roots, const uint32_t* ts, const uint32_t* lc, uint32_t R) {
  HostTimer t(c);
  (void)((Learned*)c)->on_pilot(roots, ts, lc, R);
}
void hk_batch(void* c, const uint32_t* roots, const uint32_t* ts, const uint32_t* lc, uint32_t R, uint32_t flags,
              double ms, uint64_t) {
  HostTimer t(c);
  (void)((Learned*)c)->on_batch(roots, ts, lc, R, flags, ms);
}
int hk_plant_next(void* c, int64_t r0, uint32_t R, int* final) {
  HostTimer t(c);
  return ((Learned*)c)->plant_next(r0, R, final);
}

}  // namespace

int run_build(const Graph& g, DeviceGraph& dg, const std::vector<uint32_t>& order, const chl_build_opts& opts,
              chl_build_stats* stats, LabelStore& L) {
  if (opts.batching != CHL_BATCH_LEARNED) return build_labels(g, dg, order, opts, stats, nullptr, L);
  Learned pol(g, order, opts, stats);
  int rc;
  {
    HostTimer t(&pol);
    rc = pol.init();
  }
  if (rc) return rc;
  BuildHooks h;
  h.ctx = &pol;
  h.next_width = hk_next_width;
  h.order_roots = hk_order;
  h.out_estimate = hk_out;
  if (opts.small_trees) h.route_small = hk_small;
  h.on_batch = hk_batch;
  h.pilot = hk_pilot;
  h.on_pilot = hk_on_pilot;
  if (opts.mode == CHL_MODE_HYBRID) h.plant_next = hk_plant_next;
  pol.hooks = &h;
  rc = build_labels(g, dg, order, opts, stats, &h, L);
  if (stats) {
    stats->mlp_ms = pol.mlp_ms;
    stats->pred_roots = pol.mae_n;
    stats->pred_mae_learned = pol.mae_n ? pol.mae_l / pol.mae_n : 0.0;
    stats->pred_mae_static = pol.mae_n ? pol.mae_s / pol.mae_n : 0.0;
    stats->switch_trials = pol.trials;
    stats->policy_ms = pol.host_ms;
  }
  return rc;
}

}  // namespace chl

// ---- C ABI of the predictor kernels (host pointers; used by the tests) --------
using namespace chl;

namespace {
struct MlpDev {
  DBuf<float> P, X, T, Y, G, loss, M, V;
  DBuf<uint16_t> wb;
  DBuf<uint32_t> idx;
  DBuf<int> step;
  int load(const float* params, const float* x, int64_t N) {
    CHL_HIP(P.reserve(kMlpParams));
    CHL_HIP(wb.reserve(kMlpParams));
    CHL_HIP(hipMemcpy(P.p, params, sizeof(float) * kMlpParams, hipMemcpyHostToDevice));
    CHL_HIP(mlp_sync(P.p, wb.p, nullptr));
    CHL_HIP(X.reserve((size_t)std::max<int64_t>(N, 1) * kMlpF));
    if (N) CHL_HIP(hipMemcpy(X.p, x, sizeof(float) * N * kMlpF, hipMemcpyHostToDevice));
    return CHL_OK;
  }
};
}  // namespace

extern "C" {

int chl_mlp_param_count(void) { return kMlpParams; }

int chl_mlp_forward(const float* params, const float* x, int64_t n, float* y) {
  if (!params || n < 0 || (n && (!x || !y)) || n > (1 << 30)) return set_error(CHL_E_INVAL, "bad arguments");
  int rc = ensure_device(0);
  if (rc) return rc;
  MlpDev d;
  if ((rc = d.load(params, x, n))) return rc;
  CHL_HIP(d.Y.reserve((size_t)std::max<int64_t>(n, 1) * kMlpO));
  CHL_HIP(mlp_forward(d.X.p, (int)n, d.wb.p, d.P.p, d.Y.p, nullptr));
  if (n) CHL_HIP(hipMemcpy(y, d.Y.p, sizeof(float) * n * kMlpO, hipMemcpyDeviceToHost));
  return CHL_OK;
}

int chl_mlp_grad(const float* params, const float* x, const float* t, int64_t n, float* grad, float* loss) {
  if (!params || !x || !t || !grad || n <= 0 || n > (1 << 20)) return set_error(CHL_E_INVAL, "bad arguments");
  int rc = ensure_device(0);
  if (rc) return rc;
  MlpDev d;
  if ((rc = d.load(params, x, n))) return rc;
  std::vector<uint32_t> id(n);
  std::iota(id.begin(), id.end(), 0u);
  CHL_HIP(d.idx.reserve(n));
  CHL_HIP(hipMemcpy(d.idx.p, id.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice));
  CHL_HIP(d.T.reserve((size_t)n * kMlpO));
  CHL_HIP(hipMemcpy(d.T.p, t, sizeof(float) * n * kMlpO, hipMemcpyHostToDevice));
  CHL_HIP(d.G.reserve(kMlpParams));
  CHL_HIP(d.loss.reserve(1));
  CHL_HIP(hipMemset(d.G.p, 0, sizeof(float) * kMlpParams));
  CHL_HIP(hipMemset(d.loss.p, 0, sizeof(float)));
  CHL_HIP(mlp_grad(d.X.p, d.T.p, d.idx.p, (int)n, d.wb.p, d.P.p, d.G.p, d.loss.p, nullptr));
  CHL_HIP(hipMemcpy(grad, d.G.p, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost));
  float l = 0;
  CHL_HIP(hipMemcpy(&l, d.loss.p, sizeof(float), hipMemcpyDeviceToHost));
  if (loss) *loss = l / (float)(n * kMlpO);
  return CHL_OK;
}

// Full-batch Adam for `steps` steps at learning rate lr; params updated in place,
// loss_curve[s] = mean squared error before step s (if non-null).
int chl_mlp_train(float* params, const float* x, const float* t, int64_t n, int steps, float lr, float* loss_curve) {
  if (!params || !x || !t || n <= 0 || n > (1 << 20) || steps < 0) return set_error(CHL_E_INVAL, "bad arguments");
  int rc = ensure_device(0);
  if (rc) return rc;
  MlpDev d;
  if ((rc = d.load(params, x, n))) return rc;
  std::vector<uint32_t> id(n);
  std::iota(id.begin(), id.end(), 0u);
  CHL_HIP(d.idx.reserve(n));
  CHL_HIP(hipMemcpy(d.idx.p, id.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice));
  CHL_HIP(d.T.reserve((size_t)n * kMlpO));
  CHL_HIP(hipMemcpy(d.T.p, t, sizeof(float) * n * kMlpO, hipMemcpyHostToDevice));
  CHL_HIP(d.G.reserve(kMlpParams));
  CHL_HIP(d.M.reserve(kMlpParams));
  CHL_HIP(d.V.reserve(kMlpParams));
  CHL_HIP(d.step.reserve(1));
  CHL_HIP(hipMemset(d.step.p, 0, sizeof(int)));
  CHL_HIP(d.loss.reserve(1));
  CHL_HIP(hipMemset(d.M.p, 0, sizeof(float) * kMlpParams));
  CHL_HIP(hipMemset(d.V.p, 0, sizeof(float) * kMlpParams));
  for (int s = 0; s < steps; ++s) {
    CHL_HIP(hipMemset(d.G.p, 0, sizeof(float) * kMlpParams));
    CHL_HIP(hipMemset(d.loss.p, 0, sizeof(float)));
    CHL_HIP(mlp_grad(d.X.p, d.T.p, d.idx.p, (int)n, d.wb.p, d.P.p, d.G.p, d.loss.p, nullptr));
    CHL_HIP(mlp_adam(d.P.p, d.G.p, d.M.p, d.V.p, d.wb.p, lr, d.step.p, nullptr));
    if (loss_curve) {
      float l = 0;
      CHL_HIP(hipMemcpy(&l, d.loss.p, sizeof(float), hipMemcpyDeviceToHost));
      loss_curve[s] = l / (float)(n * kMlpO);
    }
  }
  CHL_HIP(hipMemcpy(params, d.P.p, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost));
  return CHL_OK;
}

// One full-batch training step captured as a HIP graph and replayed `replays`
// times on a parameter copy, each replay followed by the same step run eagerly on
// a second copy; *max_rel gets the largest |graph - eager| / (|eager| + 1e-3)
// over all parameters after every replay.
int chl_mlp_graph_check(const float* params, const float* x, const float* t, int64_t n, int replays, float lr,
                        float* max_rel) {
  if (!params || !x || !t || !max_rel || n <= 0 || n > (1 << 20) || replays < 1)
    return set_error(CHL_E_INVAL, "bad arguments");
  int rc = ensure_device(0);
  if (rc) return rc;
  MlpDev e, g;
  if ((rc = e.load(params, x, n)) || (rc = g.load(params, x, n))) return rc;
  std::vector<uint32_t> id(n);
  std::iota(id.begin(), id.end(), 0u);
  for (MlpDev* d : {&e, &g}) {
    CHL_HIP(d->idx.reserve(n));
    CHL_HIP(hipMemcpy(d->idx.p, id.data(), sizeof(uint32_t) * n, hipMemcpyHostToDevice));
    CHL_HIP(d->T.reserve((size_t)n * kMlpO));
    CHL_HIP(hipMemcpy(d->T.p, t, sizeof(float) * n * kMlpO, hipMemcpyHostToDevice));
    CHL_HIP(d->G.reserve(kMlpParams));
    CHL_HIP(d->M.reserve(kMlpParams));
    CHL_HIP(d->V.reserve(kMlpParams));
    CHL_HIP(d->loss.reserve(1));
    CHL_HIP(d->step.reserve(1));
    CHL_HIP(hipMemset(d->M.p, 0, sizeof(float) * kMlpParams));
    CHL_HIP(hipMemset(d->V.p, 0, sizeof(float) * kMlpParams));
    CHL_HIP(hipMemset(d->step.p, 0, sizeof(int)));
  }
  hipStream_t s = nullptr;
  CHL_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  auto step = [&](MlpDev& d, hipStream_t st) -> int {
    CHL_HIP(launch_fill_u32((uint32_t*)d.G.p, 0, kMlpParams, st));
    CHL_HIP(launch_fill_u32((uint32_t*)d.loss.p, 0, 1, st));
    CHL_HIP(mlp_grad(d.X.p, d.T.p, d.idx.p, (int)n, d.wb.p, d.P.p, d.G.p, d.loss.p, st));
    CHL_HIP(mlp_adam(d.P.p, d.G.p, d.M.p, d.V.p, d.wb.p, lr, d.step.p, st));
    return CHL_OK;
  };
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  rc = CHL_OK;
  if (hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) != hipSuccess) rc = set_error(CHL_E_DEVICE, "capture");
  if (!rc) rc = step(g, s);
  if (!rc && hipStreamEndCapture(s, &graph) != hipSuccess) rc = set_error(CHL_E_DEVICE, "end capture");
  if (!rc && hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess)
    rc = set_error(CHL_E_DEVICE, "instantiate");
  // Every replay starts from the eager run's state (parameters, BF16 copy,
  // Adam moments and step), copied in by eager work between replays, and is
  // compared with one eager step from that state: gradients and updated
  // parameters, relative to their norms. (Over many steps the two trajectories
  // drift apart by design: FP32 atomics sum gradients in any order, and Adam
  // turns sign flips of near-zero gradients into full-size steps.)
  std::vector<float> pg(kMlpParams), pe(kMlpParams), gg(kMlpParams), ge(kMlpParams);
  float worst = 0.0f;
  auto rel = [](const std::vector<float>& a, const std::vector<float>& b) {
    double num = 0, den = 0;
    for (size_t i = 0; i < a.size(); ++i) {
      num += (double)(a[i] - b[i]) * (a[i] - b[i]);
      den += (double)b[i] * b[i];
    }
    return (float)std::sqrt(num / std::max(den, 1e-30));
  };
  for (int r = 0; r < replays && !rc; ++r) {
    if (hipMemcpy(g.P.p, e.P.p, sizeof(float) * kMlpParams, hipMemcpyDeviceToDevice) != hipSuccess ||
        hipMemcpy(g.wb.p, e.wb.p, sizeof(uint16_t) * kMlpParams, hipMemcpyDeviceToDevice) != hipSuccess ||
        hipMemcpy(g.M.p, e.M.p, sizeof(float) * kMlpParams, hipMemcpyDeviceToDevice) != hipSuccess ||
        hipMemcpy(g.V.p, e.V.p, sizeof(float) * kMlpParams, hipMemcpyDeviceToDevice) != hipSuccess ||
        hipMemcpy(g.step.p, e.step.p, sizeof(int), hipMemcpyDeviceToDevice) != hipSuccess ||
        hipDeviceSynchronize() != hipSuccess) {
      rc = set_error(CHL_E_DEVICE, "state copy");
      break;
    }
    if (hipGraphLaunch(exec, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess) {
      rc = set_error(CHL_E_DEVICE, "graph replay");
      break;
    }
    if ((rc = step(e, nullptr))) break;
    if (hipDeviceSynchronize() != hipSuccess ||
        hipMemcpy(pg.data(), g.P.p, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(pe.data(), e.P.p, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(gg.data(), g.G.p, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(ge.data(), e.G.p, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost) != hipSuccess) {
      rc = set_error(CHL_E_DEVICE, "readback");
      break;
    }
    worst = std::max(worst, std::max(rel(pg, pe), rel(gg, ge)));
  }
  if (exec) (void)hipGraphExecDestroy(exec);
  if (graph) (void)hipGraphDestroy(graph);
  (void)hipStreamDestroy(s);
  *max_rel = worst;
  return rc;
}

int chl_mfma_probe(const uint16_t* a, const uint16_t* b, float* d) {
  if (!a || !b || !d) return set_error(CHL_E_INVAL, "null argument");
  int rc = ensure_device(0);
  if (rc) return rc;
  DBuf<uint16_t> A, B;
  DBuf<float> D;
  CHL_HIP(A.reserve(32 * 8));
  CHL_HIP(B.reserve(8 * 32));
  CHL_HIP(D.reserve(32 * 32));
  CHL_HIP(hipMemcpy(A.p, a, sizeof(uint16_t) * 256, hipMemcpyHostToDevice));
  CHL_HIP(hipMemcpy(B.p, b, sizeof(uint16_t) * 256, hipMemcpyHostToDevice));
  CHL_HIP(mfma_probe(A.p, B.p, D.p, nullptr));
  CHL_HIP(hipMemcpy(d, D.p, sizeof(float) * 1024, hipMemcpyDeviceToHost));
  return CHL_OK;
}

}  // extern "C"
