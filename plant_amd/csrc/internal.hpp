// Definitions shared by the host C ABI and the device engine.
#pragma once
#include <memory>

#include "../../include/chl.h"
#include "engine.hpp"
#include "host.hpp"

struct chl_graph {
  chl::Graph g;
  // device copy in locality order, made by the first build on a device
  std::unique_ptr<chl::DeviceGraph> dev;
};

namespace chl {
int set_error(int code, const char* fmt, ...);
}
