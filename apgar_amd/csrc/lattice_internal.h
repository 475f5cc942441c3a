// lattice_internal.h — host-side lattice object and internal search entry
// shared by the search (lattice.hip) and the negotiated allocator
// (negotiate.hip).  Not part of the C-ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <vector>

#include "apgx.h"

struct apgx_lattice {
  apgx_lattice_desc desc{};
  void* d_cost = nullptr;
  int device = 0;
  uint32_t min_edge = 0;      // smallest edge weight (free cell cost or via)
  uint32_t min_cell_cost = 0; // smallest free cell cost (0 when every cell is blocked)
  uint32_t max_cell_cost = 0; // largest free cell cost
  uint64_t free_cells = 0;
  double mean_cost = 1.0;     // mean free-cell cost
  std::mutex cache_mutex;
  void* cached_labels = nullptr;  // reusable label pool (returned by searches)
  size_t cached_label_bytes = 0;
  void* cached_block = nullptr;   // reusable search workspace (returned by searches)
  size_t cached_block_bytes = 0;
  // One search at a time per lattice owns the stream and events (the
  // reference's per-view execution lease).
  std::mutex exec_mutex;
  hipStream_t stream = nullptr;
  hipEvent_t events[2] = {nullptr, nullptr};
  size_t lds_limit = 0;       // device attributes, queried once at creation
  uint32_t cu_count = 0;
  // Landmark fields (apgx_lattice_build_landmarks): lm_count x cells
  // distances and lm_count x column maxima; replaced under exec_mutex.
  uint32_t lm_count = 0;
  uint32_t* d_lm_dist = nullptr;
  uint32_t* d_lm_colmax = nullptr;
  std::vector<uint8_t> host_costs;  // host copy for request preflight
  uint32_t HostCost(uint64_t cell) const {
    return desc.cost_bytes == 2 ? reinterpret_cast<const uint16_t*>(host_costs.data())[cell]
                                : reinterpret_cast<const uint32_t*>(host_costs.data())[cell];
  }
};


namespace apgx {

// A device cost plane that replaces the lattice's own for one search: u32
// costs with the same blocked cells as the lattice plane and every free cost
// at least the lattice's smallest free cost (so its goal bound stays a lower
// bound).  max_cost selects the tile kernel.
struct CostOverride {
  const uint32_t* plane;
  uint32_t max_cost;
  // 1: via moves also pay the entered cell's cost (via + plane[v]).
  uint32_t via_pays_cell;
  // Optional, one per net of the search: net i reads the plane at
  // plane + net_offset[i] (whole-lattice planes laid end to end), so each net
  // can search its own snapshot; nullptr: every net reads `plane`.
  const uint64_t* net_offset = nullptr;
};

// explore: every net searches until its queue is empty (no pin is ever
// settled; status DISCONNECTED at the end) and keeps its labels for
// ExtractLandmarks.  Used to build landmark fields.
int RouteNets(apgx_lattice* lat, const CostOverride* override_plane, const apgx_net_query* nets,
              uint32_t net_count, const uint32_t* pins, uint64_t pin_count,
              const apgx_route_options* options, apgx_search** out, bool explore = false);

}  // namespace apgx
