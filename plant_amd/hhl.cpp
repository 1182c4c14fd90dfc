// Drop-in driver for shared_memory/hhl.cpp:16-79: builds the canonical hub
// labels of a graph under a vertex order on the GPU and optionally writes them.
//
//   hhl -g <graph> -o <order> [-p <host threads>] [-s <sf>] [-paraPLL] [-l <labels out>]
//       [-mode plant|gll|hybrid] [-b <roots per batch>] [-delta <d>] [-ps <ratio>]
//       [-learned] [-prune] [-device <id>]
//       [-rq <query file> | -nq <count>] [-qo <result file>]
//
// The reference flags keep their meaning where it carries over: -g and -o name
// the inputs, -l the label file (labeling.hpp:167-183 text). -p sets host
// threads used for ingest (default 16, as shared_memory/README.md:14-15 states);
// -s and -paraPLL are accepted for compatibility (the GPU batches by rank range,
// not by sf*N labels). stdout repeats the argv line and then the two result lines
// of shared_memory/akiba.hpp:214-215; stderr carries "Order size: N |V|: n".
// Queries follow the distributed driver (distributed_memory/hhl.cpp:47-48,82-91):
// -rq reads "u v" lines (load_query), -nq generates pairs as generate_query
// does; the distances go one per line to -qo (write_result) or to stdout.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../include/chl.h"

int main(int argc, char** argv) {
  for (int i = 0; i < argc; ++i) printf("%s ", argv[i]);
  printf("\n");
  const char* gpath = nullptr;
  const char* opath = nullptr;
  const char* lpath = nullptr;
  const char* qpath = nullptr;
  const char* qopath = nullptr;
  long long nq = 0;
  chl_build_opts o;
  chl_build_opts_default(&o);
  int threads = 16;
  for (int i = 1; i < argc; ++i) {
    auto arg = [&](const char* f) { return strcmp(argv[i], f) == 0 && i + 1 < argc; };
    if (arg("-g")) gpath = argv[++i];
    else if (arg("-o")) opath = argv[++i];
    else if (arg("-l")) lpath = argv[++i];
    else if (arg("-p")) threads = atoi(argv[++i]);
    else if (arg("-s")) ++i;
    else if (strcmp(argv[i], "-paraPLL") == 0) {
    } else if (arg("-mode")) {
      std::string m = argv[++i];
      o.mode = m == "gll" ? CHL_MODE_GLL : m == "hybrid" ? CHL_MODE_HYBRID : CHL_MODE_PLANT;
    } else if (arg("-b")) o.batch_roots = atoll(argv[++i]);
    else if (arg("-delta")) o.delta = atoi(argv[++i]);
    else if (arg("-ps")) o.switch_ratio = atof(argv[++i]);
    else if (arg("-device")) o.device = atoi(argv[++i]);
    else if (strcmp(argv[i], "-learned") == 0) o.batching = CHL_BATCH_LEARNED;
    else if (strcmp(argv[i], "-prune") == 0) o.plant_prune = 1;
    else if (arg("-rq")) qpath = argv[++i];
    else if (arg("-nq")) nq = atoll(argv[++i]);
    else if (arg("-qo")) qopath = argv[++i];
  }
  if (threads > 0) {
    char buf[32];
    snprintf(buf, sizeof(buf), "%d", threads);
    setenv("CHL_HOST_THREADS", buf, 0);
  }
  if (!opath) {
    fprintf(stderr, "hhl: -o <order> is required\n");
    return 1;
  }
  int64_t N = 0;
  if (chl_order_load(opath, nullptr, 0, &N) != CHL_OK) {
    printf("ERROR: couldn't open order file named %s\n", opath);
    return 1;
  }
  std::vector<uint32_t> order((size_t)N);
  chl_order_load(opath, order.data(), N, &N);
  chl_graph* g = nullptr;
  if (!gpath || chl_graph_load(gpath, 0, &g, nullptr) != CHL_OK) {
    fprintf(stderr, "Unable to read graph from raw %s\n", gpath ? gpath : "(none)");
    return 1;
  }
  fprintf(stderr, "Order size: %lld |V|: %lld\n", (long long)N, (long long)chl_graph_n(g));
  chl_labels* L = nullptr;
  chl_build_stats st;
  int rc = chl_build(g, order.data(), N, &o, &L, &st);
  if (rc != CHL_OK) {
    fprintf(stderr, "hhl: build failed: %s (%s)\n", chl_strerror(rc), chl_last_error());
    chl_graph_free(g);
    return 1;
  }
  const double tot = st.lt + st.st + st.ct;
  printf("Average label size after gll is %g, labeling time = %g, sorting time = %g, cleaning time = %g, total time = %g\n",
         st.avg_label, st.lt, st.st, st.ct, tot);
  printf("%g %g\n", st.avg_label, tot);
  if (lpath && chl_labels_write(L, lpath) != CHL_OK) {
    fprintf(stderr, "hhl: cannot write labels: %s\n", chl_last_error());
    rc = 1;
  }
  if (rc == CHL_OK && (qpath || nq > 0)) {
    std::vector<uint32_t> pairs;
    int64_t q = 0;
    if (qpath) {
      if (chl_query_load(qpath, nullptr, 0, &q) != CHL_OK) {
        fprintf(stderr, "hhl: cannot read queries: %s\n", chl_last_error());
        rc = 1;
      } else {
        pairs.resize(2 * (size_t)q);
        chl_query_load(qpath, pairs.data(), q, &q);
      }
    } else {
      q = nq;
      pairs.resize(2 * (size_t)q);
      chl_query_generate(q, (uint32_t)N, pairs.data());
    }
    if (rc == CHL_OK) {
      std::vector<int32_t> dist((size_t)q);
      if (chl_query(L, pairs.data(), q, dist.data()) != CHL_OK) {
        fprintf(stderr, "hhl: query failed: %s\n", chl_last_error());
        rc = 1;
      } else if (qopath) {
        if (chl_query_write(qopath, dist.data(), q) != CHL_OK) {
          fprintf(stderr, "hhl: cannot write results: %s\n", chl_last_error());
          rc = 1;
        }
      } else {
        for (int64_t k = 0; k < q; ++k) printf("%d\n", dist[k]);
      }
    }
  }
  chl_labels_free(L);
  chl_graph_free(g);
  return rc == CHL_OK ? 0 : 1;
}
