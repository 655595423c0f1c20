// The opaque handle types of the C ABI, shared by the translation units that
// implement it.
#pragma once
#include <memory>
#include <vector>

#include "../../include/pfcore.h"
#include "pf_engine_api.hpp"

struct pf_handle {
  std::unique_ptr<pf::EngineBase> eng;
  int precision = PF_FP64;
};

struct pf_batch {
  std::vector<pf_handle*> h;
  int threads = 1;
  // launches of the last fused call: batched launches issued, and the
  // launches the instances recorded (what one-by-one execution would issue)
  long long fused_launches = 0, fused_recorded = 0;
  // graphs captured and graph launches of the last fused call; whether fused
  // calls capture their phases into graphs (pf_batch_set_graphs)
  long long fused_graphs_built = 0, fused_graph_launches = 0;
  int graphs = 1;
};
