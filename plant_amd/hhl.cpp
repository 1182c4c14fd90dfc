// Drop-in driver for shared_memory/hhl.cpp:16-79: builds the canonical hub
// labels of a graph under a vertex order on the GPU and optionally writes them.
//
//   hhl -g <graph> -o <order> [-p <host threads>] [-s <sf>] [-paraPLL] [-l <labels out>]
//       [-mode plant|gll|hybrid] [-b <roots per batch>] [-delta <d>] [-ps <ratio>]
//       [-learned] [-prune] [-device <id>]
//
// The reference flags keep their meaning where it carries over: -g and -o name
// the inputs, -l the label file (labeling.hpp:167-183 text). -p sets host
// threads used for ingest (default 16, as shared_memory/README.md:14-15 states);
// -s and -paraPLL are accepted for compatibility (the GPU batches by rank range,
// not by sf*N labels). stdout repeats the argv line and then the two result lines
// of shared_memory/akiba.hpp:214-215; stderr carries "Order size: N |V|: n".
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
  chl_labels_free(L);
  chl_graph_free(g);
  return rc == CHL_OK ? 0 : 1;
}
