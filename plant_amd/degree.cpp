// Drop-in for the reference's degree-order tool (distributed_memory/degree.cpp:
// 32-70): reads a graph, prints "Graph has <n> vertices and <m> arcs", and
// writes the vertices by degree (out-arcs plus in-arcs) descending, ties by the
// higher id first, as an order file (ordering.hpp:37-49 format). The order is
// computed on the GPU (chl_order_degree_device).
//
//   degree [-device <id>] -o <ordering> <graph>
//
// The command line keeps the reference's rules: options before the graph, "--"
// ends them, -h or anything unknown prints the usage and exits with status 1, as
// does a missing -o or graph or an extra argument. -device is this tool's own.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../include/chl.h"

static void usage(const char* argv0) {
  printf("Usage: %s -o ordering graph\n", argv0);
  printf("  -o ordering\tFile with the vertex order\n");
  exit(1);
}

int main(int argc, char** argv) {
  const char* graph_file = nullptr;
  const char* order_file = nullptr;
  int device = 0;
  int argi;
  for (argi = 1; argi < argc; ++argi) {
    if (argv[argi][0] == '-') {
      if (!strcmp("--", argv[argi])) {
        ++argi;
        break;
      } else if (!strcmp("-o", argv[argi])) {
        if (++argi >= argc) usage(argv[0]);
        order_file = argv[argi];
      } else if (!strcmp("-device", argv[argi])) {
        if (++argi >= argc) usage(argv[0]);
        device = atoi(argv[argi]);
      } else {
        usage(argv[0]);  // -h and unknown options
      }
    } else if (!graph_file) {
      graph_file = argv[argi];
    } else {
      break;
    }
  }
  if (argi != argc || !graph_file || !order_file) usage(argv[0]);

  chl_graph* g = nullptr;
  if (chl_graph_load(graph_file, 0, &g, nullptr) != CHL_OK) {
    fprintf(stderr, "Unable to read graph from file %s\n", graph_file);
    return 1;
  }
  const int64_t n = chl_graph_n(g);
  printf("Graph has %lld vertices and %lld arcs\n", (long long)n, (long long)chl_graph_m(g));
  fflush(stdout);
  std::vector<uint32_t> order((size_t)n);
  const int rc = chl_order_degree_device(g, 0, /*reference_ties=*/1, device, order.data(), nullptr);
  if (rc != CHL_OK) {
    fprintf(stderr, "degree: %s (%s)\n", chl_strerror(rc), chl_last_error());
    chl_graph_free(g);
    return 1;
  }
  if (chl_order_write(order_file, order.data(), n) != CHL_OK)
    fprintf(stderr, "Unable to write order to file %s\n", order_file);
  chl_graph_free(g);
  return 0;
}
