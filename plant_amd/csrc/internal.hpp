// Definitions shared by the host C ABI and the device engine.
#pragma once
#include "../../include/chl.h"
#include "host.hpp"

struct chl_graph {
  chl::Graph g;
};

namespace chl {
int set_error(int code, const char* fmt, ...);
}
