// v1_view.h — the uploaded V1 view, shared by the batch and single-route paths.
#pragma once

#include <cstdint>
#include <mutex>

#include "apgx.h"

struct apgx_v1_view {
  apgx_v1_header header{};
  int device = 0;
  uint64_t layer_count = 0, node_count = 0, run_count = 0, run_node_count = 0, legal_edges = 0;
  apgx_v1_node* d_nodes = nullptr;
  apgx_v1_run* d_runs = nullptr;
  uint32_t* d_run_nodes = nullptr;
  uint64_t persistent_bytes = 0;
  // Execution lease of single routes: taken by apgx_v1_route_execute and held
  // by the pending route until it is released, one route at a time per view
  // (the reference's per-view execution mutex, gpu/cuda/planar_router.cu:232-237).
  std::mutex route_lease;
  // Host-mapped cancellation word (allocated on the first cancellable route):
  // the host mirrors the caller's flag into it while the route runs and the
  // search kernel reads it once per round.
  uint32_t* cancel_word = nullptr;
};
