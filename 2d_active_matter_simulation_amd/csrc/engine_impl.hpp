// Internals of the ensemble engine shared with the serving loop.
#pragma once
#include <map>
#include <tuple>
#include <vector>

#include "run_impl.hpp"

namespace af {

struct Slot {
  bool used = false;
  af_params p;
  Geometry g;
  unsigned char* base = nullptr;
  long long steps = 0;
};

struct Class {
  hipStream_t stream = nullptr;
  void* descs = nullptr;   // device array of per-run descriptors
  int cap = 0;             // capacity of descs (entries)
  size_t desc_bytes = 0;
};

using ClassKey = std::tuple<int, int, int, const Schedule*>;  // precision, nx, ny, schedule

}  // namespace af

struct af_engine {
  size_t budget = 0, in_use = 0;
  std::vector<af::Slot> slots;
  std::map<af::ClassKey, af::Class> classes;
  hipStream_t stream = nullptr;
  // cached graphs (one step, kGraphSteps steps) for the last stepped slot list
  std::vector<int> graph_slots;
  hipGraph_t graph[2] = {nullptr, nullptr};
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  std::vector<hipEvent_t> events;
};
