// negotiate.hip — negotiated-congestion allocator on the GPU.
//
// Canonical Sequential Rip-Up-and-Reroute v1
// (src/allocator/cpu_sequential_negotiated_routing.cc:993-1329) restated on the
// dense lattice with one resource per cell.  Every plane-wide step is a
// coalesced HBM-bound kernel over the resident planes (no host copies of
// them during the run):
//   rip-up / install   atomic +-1 scatter of the batch's tree cells into usage
//   snapshot           cost = base + history + pf * max(0, usage + 1 - cap),
//                      bound checks and the plane's largest cost in one pass
//   pass boundary      overuse = sum max(0, usage - cap); history update
// The per-batch rip-up + snapshot pair reads its parameters from device
// memory and is captured once as a hipGraph, replayed for every batch.  The
// searches run through the same entry as apgx_route_nets with the snapshot
// plane in place of the lattice's costs.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <memory>
#include <vector>

#include "apgx.h"
#include "common.h"
#include "lattice_internal.h"

struct apgx_negotiation {
  uint32_t net_count = 0;
  uint64_t cells = 0;
  apgx_negotiation_summary summary{};
  std::vector<uint64_t> overuse_per_pass;
  std::vector<uint32_t> history, usage;
  std::vector<std::vector<uint32_t>> routes;
  std::vector<uint8_t> routed;
};

namespace apgx {
namespace {

constexpr int kBlock = 256;
constexpr uint32_t kGrid = 2048;  // grid-stride kernels: 8 waves' worth per CU
// Searches of at most this many nets use a bucket of kLatencyDeltaCells mean
// cell costs (16 times the lattice search's default of 48) unless the caller
// sets frontier_delta, and check for their end every kLatencySyncRounds
// rounds unless the caller sets rounds_per_sync.
constexpr uint32_t kLatencyBatchNets = 8;
constexpr double kLatencyDeltaCells = 768.0;
constexpr uint32_t kLatencySyncRounds = 16;

// Parameters of the captured batch-prep graph, in device memory.
struct BatchParams {
  const uint32_t* list;   // cells to rip up (the batch's incumbents)
  uint64_t count;
  uint64_t present_factor;
};

// Flags: [0] bound exceeded, [1] largest snapshot cost, [2..3] overuse (u64).
struct Flags {
  uint32_t bound;
  uint32_t max_cost;
  unsigned long long overuse;
};

__global__ __launch_bounds__(kBlock) void ScatterKernel(uint32_t* usage, const BatchParams* p,
                                                        const uint32_t* list, uint64_t count, int delta) {
  const uint32_t* cells = p ? p->list : list;
  const uint64_t n = p ? p->count : count;
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x)
    atomicAdd(&usage[cells[i]], static_cast<uint32_t>(delta));
}

__device__ __forceinline__ uint32_t WaveMax(uint32_t v) {
  for (int o = 32; o > 0; o >>= 1) v = max(v, static_cast<uint32_t>(__shfl_xor(static_cast<int>(v), o)));
  return v;
}
__device__ __forceinline__ unsigned long long WaveSum(unsigned long long v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Snapshot cost plane.  Free cells: base + cong with
// cong = history + pf * max(0, usage + 1 - capacity); blocked cells stay 0.
template <typename CostT>
__global__ __launch_bounds__(kBlock) void SnapshotKernel(const CostT* base, const uint8_t* cap,
                                                         const uint32_t* history, const uint32_t* usage,
                                                         uint32_t* snap, uint64_t cells,
                                                         const BatchParams* p, uint64_t max_cong,
                                                         uint32_t max_cell, Flags* flags) {
  const uint64_t pf = p->present_factor;
  uint32_t local_max = 0, bound = 0;
  for (uint64_t c = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; c < cells;
       c += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint32_t b = base[c];
    uint32_t out = 0;
    if (b != 0) {
      const uint64_t capacity = cap ? cap[c] : 1u;
      const uint64_t predicted = static_cast<uint64_t>(usage[c]) + 1;
      const uint64_t over = predicted > capacity ? predicted - capacity : 0;
      uint64_t present = 0, cong = 0;
      const bool ovf = __builtin_mul_overflow(pf, over, &present) ||
                       __builtin_add_overflow(present, static_cast<uint64_t>(history[c]), &cong);
      if (ovf || cong > max_cong || b + cong > max_cell) {
        bound = 1;
      } else {
        out = static_cast<uint32_t>(b + cong);
      }
    }
    snap[c] = out;
    local_max = max(local_max, out);
  }
  local_max = WaveMax(local_max);
  if ((threadIdx.x & 63) == 0) atomicMax(&flags->max_cost, local_max);
  if (__any(bound) && (threadIdx.x & 63) == 0) atomicOr(&flags->bound, 1u);
}

__global__ __launch_bounds__(kBlock) void OveruseKernel(const uint8_t* cap, const uint32_t* usage,
                                                        uint64_t cells, Flags* flags) {
  unsigned long long sum = 0;
  for (uint64_t c = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; c < cells;
       c += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint32_t capacity = cap ? cap[c] : 1u;
    const uint32_t u = usage[c];
    sum += u > capacity ? u - capacity : 0u;
  }
  sum = WaveSum(sum);
  if ((threadIdx.x & 63) == 0 && sum) atomicAdd(&flags->overuse, sum);
}

__global__ __launch_bounds__(kBlock) void HistoryKernel(const uint8_t* cap, const uint32_t* usage,
                                                        uint32_t* history, uint64_t cells, uint64_t inc,
                                                        uint64_t max_value, Flags* flags) {
  uint32_t bound = 0;
  for (uint64_t c = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; c < cells;
       c += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint32_t capacity = cap ? cap[c] : 1u;
    const uint32_t u = usage[c];
    if (u > capacity) {
      uint64_t add = 0, h = 0;
      const bool ovf = __builtin_mul_overflow(inc, static_cast<uint64_t>(u - capacity), &add) ||
                       __builtin_add_overflow(add, static_cast<uint64_t>(history[c]), &h);
      if (ovf || h > max_value || h > 0xFFFFFFFFull) bound = 1;
      else history[c] = static_cast<uint32_t>(h);
    }
  }
  if (__any(bound) && (threadIdx.x & 63) == 0) atomicOr(&flags->bound, 1u);
}

// Snapshot of one net's window (every layer) into that net's own plane, at the
// lattice cell index: free cells get base + history + pf * max(0, usage + 1 -
// capacity), blocked cells 0.  The bound checks of the snapshot instant run on
// the host copy of the usage plane; here the cost is only clamped below the
// search's saturated distance (values that large fail those checks, and the
// run with them).  *max_cost receives the window's largest cost.
template <typename CostT>
__global__ __launch_bounds__(kBlock) void WindowSnapshotKernel(const CostT* base, const uint8_t* cap,
                                                               const uint32_t* history, const uint32_t* usage,
                                                               uint32_t* dst, uint32_t W, uint32_t H,
                                                               uint32_t x0, uint32_t y0, uint32_t w, uint32_t h,
                                                               uint32_t L, uint64_t pf, uint32_t* max_cost) {
  const uint64_t plane = static_cast<uint64_t>(w) * h, n = plane * L;
  uint32_t local_max = 0;
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < n;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint64_t l = i / plane, r = i % plane;
    const uint64_t c = (l * H + y0 + r / w) * W + x0 + r % w;
    const uint32_t b = base[c];
    uint32_t out = 0;
    if (b != 0) {
      const uint64_t capacity = cap ? cap[c] : 1u;
      const uint64_t predicted = static_cast<uint64_t>(usage[c]) + 1;
      const uint64_t over = predicted > capacity ? predicted - capacity : 0;
      uint64_t present = 0, cost = 0;
      const bool ovf = __builtin_mul_overflow(pf, over, &present) ||
                       __builtin_add_overflow(present, static_cast<uint64_t>(history[c]) + b, &cost);
      out = (ovf || cost > 0x1FFFFFFFull) ? 0x1FFFFFFFu : static_cast<uint32_t>(cost);
    }
    dst[c] = out;
    local_max = max(local_max, out);
  }
  local_max = WaveMax(local_max);
  if ((threadIdx.x & 63) == 0) atomicMax(max_cost, local_max);
}

struct DeviceBuffers {
  uint8_t* cap = nullptr;
  uint32_t* history = nullptr;
  uint32_t* usage = nullptr;
  uint32_t* snap = nullptr;
  uint32_t* list = nullptr;
  uint64_t list_capacity = 0;
  BatchParams* params = nullptr;
  Flags* flags = nullptr;
  uint32_t* work = nullptr;      // speculative groups: usage with the group's rip-ups applied
  uint32_t* slots = nullptr;     // speculative groups: one whole-lattice snapshot plane per member
  uint32_t* slot_max = nullptr;  // largest cost of each member's window snapshot
  ~DeviceBuffers() {
    for (void* p : {static_cast<void*>(cap), static_cast<void*>(history), static_cast<void*>(usage),
                    static_cast<void*>(snap), static_cast<void*>(list), static_cast<void*>(params),
                    static_cast<void*>(flags), static_cast<void*>(work), static_cast<void*>(slots),
                    static_cast<void*>(slot_max)})
      if (p) (void)hipFree(p);
  }
};

// Exact-order batches: level(j) = 1 + the largest level of an earlier net whose
// window overlaps j's (0 if none).  Nets of one level have pairwise disjoint
// windows, so none reads a cell another one's rip-up or new route changes, and
// routing a level at once equals routing its nets one after another; levels in
// increasing order keep every overlapping pair in net order.
std::vector<std::vector<uint32_t>> WindowLevels(const apgx_net_query* nets, uint32_t n,
                                                const apgx_lattice_desc& d) {
  struct Box { uint32_t x0, y0, x1, y1; };
  std::vector<Box> box(n);
  for (uint32_t i = 0; i < n; ++i) {
    const apgx_window& w = nets[i].window;
    box[i] = (w.width == 0 && w.height == 0) ? Box{0, 0, d.width, d.height}
                                             : Box{w.x0, w.y0, w.x0 + w.width, w.y0 + w.height};
  }
  std::vector<uint32_t> level(n, 0);
  uint32_t levels = 0;
  for (uint32_t j = 0; j < n; ++j) {
    uint32_t lv = 0;
    for (uint32_t i = 0; i < j; ++i)
      if (box[i].x0 < box[j].x1 && box[j].x0 < box[i].x1 && box[i].y0 < box[j].y1 && box[j].y0 < box[i].y1)
        lv = std::max(lv, level[i] + 1);
    level[j] = lv;
    levels = std::max(levels, lv + 1);
  }
  std::vector<std::vector<uint32_t>> out(levels);
  for (uint32_t j = 0; j < n; ++j) out[level[j]].push_back(j);
  return out;
}

}  // namespace
}  // namespace apgx

extern "C" void apgx_negotiation_policy_defaults(apgx_negotiation_policy* p) {
  if (p == nullptr) return;
  p->maximum_passes = 8;
  p->batch_nets = 1;
  p->initial_present_cost = 0;
  p->present_cost_increment = 100;
  p->historical_cost_increment = 10;
  p->maximum_congestion_cost_value = 10000000;
  p->on_pass = nullptr;
  p->on_pass_context = nullptr;
  p->group_nets = 0;
  p->reserved = 0;
}

extern "C" int apgx_negotiate(apgx_lattice* lat, const uint8_t* capacity, const apgx_net_query* nets,
                              uint32_t net_count, const uint32_t* pins, uint64_t pin_count,
                              const apgx_negotiation_policy* policy, const uint32_t* initial_history,
                              const apgx_route_options* search_options, apgx_negotiation** out) {
  using namespace apgx;
  const auto wall_start = std::chrono::steady_clock::now();
  if (lat == nullptr || out == nullptr || (net_count && (nets == nullptr || pins == nullptr)))
    return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_negotiate: null argument");
  *out = nullptr;
  apgx_negotiation_policy pol;
  apgx_negotiation_policy_defaults(&pol);
  if (policy) pol = *policy;
  if (pol.maximum_passes == 0)
    return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_negotiate: maximum_passes must be positive");
  const apgx_lattice_desc& d = lat->desc;
  const uint64_t cells = static_cast<uint64_t>(d.width) * d.height * d.layers;
  apgx_route_options sopt{};
  if (search_options) sopt = *search_options;

  DeviceBuffers b;
  APGX_HIP(hipMalloc(&b.history, 4 * cells), "hipMalloc");
  APGX_HIP(hipMalloc(&b.usage, 4 * cells), "hipMalloc");
  APGX_HIP(hipMalloc(&b.snap, 4 * cells), "hipMalloc");
  APGX_HIP(hipMalloc(&b.params, sizeof(BatchParams)), "hipMalloc");
  APGX_HIP(hipMalloc(&b.flags, sizeof(Flags)), "hipMalloc");
  if (capacity) {
    APGX_HIP(hipMalloc(&b.cap, cells), "hipMalloc");
    APGX_HIP(hipMemcpy(b.cap, capacity, cells, hipMemcpyHostToDevice), "capacity upload");
  }
  if (initial_history) {
    APGX_HIP(hipMemcpy(b.history, initial_history, 4 * cells, hipMemcpyHostToDevice), "history upload");
  } else {
    APGX_HIP(hipMemset(b.history, 0, 4 * cells), "history clear");
  }
  APGX_HIP(hipMemset(b.usage, 0, 4 * cells), "usage clear");
  hipStream_t stream = nullptr;
  APGX_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
  struct StreamGuard {
    hipStream_t s;
    hipGraphExec_t g = nullptr;
    ~StreamGuard() {
      if (g) (void)hipGraphExecDestroy(g);
      (void)hipStreamDestroy(s);
    }
  } guard{stream};

  auto* neg = new (std::nothrow) apgx_negotiation();
  if (neg == nullptr) return Fail(APGX_E_RESOURCE_EXHAUSTED, "apgx_negotiate: host allocation");
  std::unique_ptr<apgx_negotiation> owner(neg);
  neg->net_count = net_count;
  neg->cells = cells;
  neg->routes.assign(net_count, {});
  neg->routed.assign(net_count, 0);
  std::vector<std::vector<uint32_t>> previous;
  std::vector<uint8_t> previous_routed;
  bool have_previous = false;
  const uint64_t max_cong = pol.maximum_congestion_cost_value;
  const uint32_t max_cell = APGX_MAX_CELL_COST;
  Flags host_flags{};
  // Host staging of the scatter lists; the rip-up list is reused only after
  // the batch's synchronisation, the install list only after the next one.
  std::vector<uint32_t> rip_list, install_list;

  auto ensure_list = [&](uint64_t n) -> int {
    if (n <= b.list_capacity) return APGX_OK;
    APGX_HIP(hipStreamSynchronize(stream), "hipStreamSynchronize");  // pending scatters read the old list
    if (b.list) APGX_HIP(hipFree(b.list), "hipFree");
    b.list = nullptr;
    b.list_capacity = std::max<uint64_t>(n, 2 * b.list_capacity);
    APGX_HIP(hipMalloc(&b.list, 4 * b.list_capacity), "hipMalloc");
    return APGX_OK;
  };
  const uint32_t grid = kGrid;
  // Batch prep: rip up the batch's incumbents, then the snapshot plane.
  auto launch_prep = [&]() -> int {
    hipLaunchKernelGGL(ScatterKernel, dim3(grid), dim3(kBlock), 0, stream, b.usage,
                       static_cast<const BatchParams*>(b.params), static_cast<const uint32_t*>(nullptr),
                       uint64_t{0}, -1);
    if (d.cost_bytes == 2)
      hipLaunchKernelGGL(SnapshotKernel<uint16_t>, dim3(grid), dim3(kBlock), 0, stream,
                         static_cast<const uint16_t*>(lat->d_cost), static_cast<const uint8_t*>(b.cap),
                         static_cast<const uint32_t*>(b.history), static_cast<const uint32_t*>(b.usage),
                         b.snap, cells, static_cast<const BatchParams*>(b.params), max_cong, max_cell, b.flags);
    else
      hipLaunchKernelGGL(SnapshotKernel<uint32_t>, dim3(grid), dim3(kBlock), 0, stream,
                         static_cast<const uint32_t*>(lat->d_cost), static_cast<const uint8_t*>(b.cap),
                         static_cast<const uint32_t*>(b.history), static_cast<const uint32_t*>(b.usage),
                         b.snap, cells, static_cast<const BatchParams*>(b.params), max_cong, max_cell, b.flags);
    APGX_HIP(hipGetLastError(), "batch prep launch");
    return APGX_OK;
  };
  {
    hipGraph_t graph = nullptr;
    APGX_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    const int rc = launch_prep();
    const hipError_t end = hipStreamEndCapture(stream, &graph);
    if (rc != APGX_OK) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    APGX_HIP(end, "hipStreamEndCapture");
    const hipError_t inst = hipGraphInstantiate(&guard.g, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    APGX_HIP(inst, "hipGraphInstantiate");
  }

  // Batches in processing order: consecutive runs of batch_nets nets, or
  // (batch_nets == 0) the window-disjoint levels of the exact order.
  std::vector<std::vector<uint32_t>> batches, all_nets(1);
  for (uint32_t n = 0; n < net_count; ++n) all_nets[0].push_back(n);
  if (pol.batch_nets == 0) {
    if (pol.group_nets == 0) batches = WindowLevels(nets, net_count, d);
  } else {
    for (uint32_t b0 = 0; b0 < net_count; b0 += pol.batch_nets) {
      std::vector<uint32_t> v;
      for (uint32_t n = b0; n < std::min<uint64_t>(net_count, uint64_t{b0} + pol.batch_nets); ++n) v.push_back(n);
      batches.push_back(std::move(v));
    }
  }
  apgx_negotiation_summary& sum = neg->summary;
  std::vector<apgx_net_query> batch_nets;
  std::vector<apgx_net_result> results;
  std::vector<apgx_connection> conns;
  std::vector<uint32_t> path_cells;
  // Searches `members` on `ov` and reads the results back into results /
  // conns / path_cells (result i belongs to members[i]).
  auto search = [&](const std::vector<uint32_t>& members, const CostOverride& ov) -> int {
    const uint32_t count = static_cast<uint32_t>(members.size());
    batch_nets.clear();
    for (uint32_t n : members) batch_nets.push_back(nets[n]);
    // Bucket width: a batch of a few nets leaves the GPU nearly idle, so its
    // search time is rounds times the latency of one round; wide buckets cut
    // the rounds (tile visits they waste are nearly free there).  A large
    // batch keeps the lattice search's default (visits cost throughput).
    // six_layer_1536, an exact-order pass of 1-2 nets per search: 17.8 s at
    // the default width, 9.6 s at 16 times it (wider buckets change nothing
    // more: the rounds are then the propagation depth); the 2000-net first
    // pass: 0.33 s at the default, 0.75 s at 16 times it.
    // Such a search also checks for its end every kLatencySyncRounds rounds
    // rather than every 8: a check leaves the GPU idle while the host reads
    // the round's count, and the rounds launched past the end are empty
    // (six_layer_1536 pass 1: 8.05 s at 8, 7.89 s at 16, 8.32 s at 4; the
    // same trees).
    apgx_route_options opt = sopt;
    if (opt.frontier_delta == 0 && count <= kLatencyBatchNets)
      opt.frontier_delta = static_cast<uint32_t>(std::max(1.0, kLatencyDeltaCells * lat->mean_cost));
    if (opt.rounds_per_sync == 0 && count <= kLatencyBatchNets) opt.rounds_per_sync = kLatencySyncRounds;
    apgx_search* s = nullptr;
    if (int rc = RouteNets(lat, &ov, batch_nets.data(), count, pins, pin_count, &opt, &s); rc != APGX_OK)
      return rc;
    uint64_t nconn = 0, ncell = 0;
    int rc = apgx_search_sizes(s, &nconn, &ncell);
    results.resize(count);
    conns.resize(std::max<uint64_t>(nconn, 1));
    path_cells.resize(std::max<uint64_t>(ncell, 1));
    if (rc == APGX_OK) rc = apgx_search_readback(s, results.data(), conns.data(), nconn, path_cells.data(), ncell);
    apgx_route_stats st{};
    if (rc == APGX_OK) rc = apgx_search_stats(s, &st);
    apgx_search_release(s);
    if (rc != APGX_OK) return rc;
    sum.search_device_ms += st.device_ms;
    sum.tile_visits += st.tile_visits;
    sum.search_rounds += st.rounds;
    ++sum.batches;
    for (const apgx_net_result& r : results)
      if (r.status == APGX_NET_BUDGET_EXHAUSTED)
        return Fail(APGX_E_RESOURCE_EXHAUSTED, "apgx_negotiate: a search exhausted its round budget");
    return APGX_OK;
  };
  // The route a net installs after its search (result i): its new tree when
  // reached (pin 0, then each connection path from its pin up to, not
  // including, the tree cell it joins), else its incumbent.
  auto installed_route = [&](uint32_t i, uint32_t n, std::vector<uint32_t>* out) {
    const apgx_net_result& r = results[i];
    if (r.status != APGX_NET_REACHED) {
      *out = neg->routes[n];
      return false;
    }
    out->clear();
    out->push_back(pins[nets[n].pin_offset]);
    for (uint32_t k = 0; k < r.connection_count; ++k) {
      const apgx_connection& c = conns[r.connection_offset + k];
      for (uint32_t j = 0; j + 1 < c.cell_count; ++j) out->push_back(path_cells[c.cell_offset + j]);
    }
    return true;
  };
  auto launch_scatter = [&](uint32_t* plane, const uint32_t* list, uint64_t count, int delta) -> int {
    if (count == 0) return APGX_OK;
    hipLaunchKernelGGL(ScatterKernel, dim3(grid), dim3(kBlock), 0, stream, plane,
                       static_cast<const BatchParams*>(nullptr), list, count, delta);
    APGX_HIP(hipGetLastError(), "scatter launch");
    return APGX_OK;
  };
  auto launch_window_snapshot = [&](uint32_t n, const uint32_t* usage, uint32_t* dst, uint32_t* max_cost,
                                    uint64_t pf) -> int {
    const apgx_window& q = nets[n].window;
    const apgx_window win = (q.width == 0 && q.height == 0) ? apgx_window{0, 0, d.width, d.height} : q;
    if (win.width == 0 || win.height == 0 || win.x0 + static_cast<uint64_t>(win.width) > d.width ||
        win.y0 + static_cast<uint64_t>(win.height) > d.height)
      return APGX_OK;  // an invalid window: the search reports the net invalid without reading costs
    const uint64_t n_cells = static_cast<uint64_t>(win.width) * win.height * d.layers;
    const uint32_t blocks = static_cast<uint32_t>(std::min<uint64_t>(grid, (n_cells + kBlock - 1) / kBlock));
    if (d.cost_bytes == 2)
      hipLaunchKernelGGL(WindowSnapshotKernel<uint16_t>, dim3(blocks), dim3(kBlock), 0, stream,
                         static_cast<const uint16_t*>(lat->d_cost), static_cast<const uint8_t*>(b.cap),
                         static_cast<const uint32_t*>(b.history), usage, dst, d.width, d.height, win.x0, win.y0,
                         win.width, win.height, d.layers, pf, max_cost);
    else
      hipLaunchKernelGGL(WindowSnapshotKernel<uint32_t>, dim3(blocks), dim3(kBlock), 0, stream,
                         static_cast<const uint32_t*>(lat->d_cost), static_cast<const uint8_t*>(b.cap),
                         static_cast<const uint32_t*>(b.history), usage, dst, d.width, d.height, win.x0, win.y0,
                         win.width, win.height, d.layers, pf, max_cost);
    APGX_HIP(hipGetLastError(), "window snapshot launch");
    return APGX_OK;
  };

  // Speculative groups (batch_nets = 0, group_nets > 0): device planes for the
  // group's working usage and one snapshot plane per member; host copies of the
  // usage and history planes for the per-net bound checks; per-cell stamps of
  // the commit that last installed the cell.
  const uint32_t group = pol.batch_nets == 0 ? std::min(pol.group_nets, std::max(net_count, 1u)) : 0u;
  std::vector<uint32_t> host_usage, host_history, stamp;
  uint32_t seq = 0;
  if (group > 0) {
    APGX_HIP(hipMalloc(&b.work, 4 * cells), "hipMalloc");
    APGX_HIP(hipMalloc(&b.slots, 4 * cells * group), "hipMalloc");
    APGX_HIP(hipMalloc(&b.slot_max, 4 * group), "hipMalloc");
    host_usage.resize(cells);
    host_history.resize(cells);
    stamp.assign(cells, 0);
  }
  // The bound checks of the snapshot instant of a net, on the host planes:
  // congestion and cell cost of `cell` at usage host_usage[cell].
  auto cell_within_bounds = [&](uint64_t cell, uint64_t pf) -> bool {
    const uint64_t bcost = lat->HostCost(cell);
    if (bcost == 0) return true;
    const uint64_t capv = capacity ? capacity[cell] : 1u;
    const uint64_t predicted = static_cast<uint64_t>(host_usage[cell]) + 1;
    const uint64_t over = predicted > capv ? predicted - capv : 0;
    uint64_t present = 0, cong = 0;
    if (__builtin_mul_overflow(pf, over, &present) ||
        __builtin_add_overflow(present, static_cast<uint64_t>(host_history[cell]), &cong))
      return false;
    return cong <= max_cong && bcost + cong <= max_cell;
  };
  // One pass of the exact order in speculative groups.
  auto group_pass = [&](uint64_t pf) -> int {
    APGX_HIP(hipMemcpy(host_usage.data(), b.usage, 4 * cells, hipMemcpyDeviceToHost), "usage readback");
    APGX_HIP(hipMemcpy(host_history.data(), b.history, 4 * cells, hipMemcpyDeviceToHost), "history readback");
    // The first net's snapshot instant is checked over the whole plane on the
    // device (its rip-up, the plane-wide snapshot and its bound flag, then the
    // rip-up undone); every later instant differs from the one before only at
    // the cells the previous net installed.
    {
      std::vector<uint32_t> first = neg->routes[0];
      if (int rc = ensure_list(std::max<uint64_t>(first.size(), 1)); rc != APGX_OK) return rc;
      if (!first.empty())
        APGX_HIP(hipMemcpyAsync(b.list, first.data(), 4 * first.size(), hipMemcpyHostToDevice, stream),
                 "rip-up upload");
      const BatchParams params{b.list, first.size(), pf};
      APGX_HIP(hipMemcpyAsync(b.params, &params, sizeof(params), hipMemcpyHostToDevice, stream), "params upload");
      host_flags = Flags{0, 0, 0};
      APGX_HIP(hipMemcpyAsync(b.flags, &host_flags, sizeof(Flags), hipMemcpyHostToDevice, stream), "flags clear");
      APGX_HIP(hipGraphLaunch(guard.g, stream), "hipGraphLaunch");
      if (int rc = launch_scatter(b.usage, b.list, first.size(), 1); rc != APGX_OK) return rc;
      APGX_HIP(hipMemcpyAsync(&host_flags, b.flags, sizeof(Flags), hipMemcpyDeviceToHost, stream), "flags readback");
      APGX_HIP(hipStreamSynchronize(stream), "hipStreamSynchronize");
      if (host_flags.bound)
        return Fail(APGX_E_RESOURCE_EXHAUSTED, "apgx_negotiate: congestion cost exceeds the configured value bound");
    }
    std::vector<uint32_t> prev_installed, tree, upload;
    std::vector<uint32_t> pending_rip, pending_install;
    std::vector<uint64_t> offsets(group);
    for (uint32_t k = 0; k < group; ++k) offsets[k] = static_cast<uint64_t>(k) * cells;
    // Applies the pending commits to the committed usage plane.
    auto flush = [&]() -> int {
      if (pending_rip.empty() && pending_install.empty()) return APGX_OK;
      APGX_HIP(hipStreamSynchronize(stream), "hipStreamSynchronize");  // the list buffer is free again
      upload = pending_rip;
      upload.insert(upload.end(), pending_install.begin(), pending_install.end());
      if (int rc = ensure_list(upload.size()); rc != APGX_OK) return rc;
      APGX_HIP(hipMemcpyAsync(b.list, upload.data(), 4 * upload.size(), hipMemcpyHostToDevice, stream),
               "commit upload");
      if (int rc = launch_scatter(b.usage, b.list, pending_rip.size(), -1); rc != APGX_OK) return rc;
      if (int rc = launch_scatter(b.usage, b.list + pending_rip.size(), pending_install.size(), 1); rc != APGX_OK)
        return rc;
      APGX_HIP(hipStreamSynchronize(stream), "hipStreamSynchronize");
      pending_rip.clear();
      pending_install.clear();
      return APGX_OK;
    };
    std::vector<uint32_t> members;
    std::vector<uint32_t> slot_max(group);
    std::vector<std::vector<uint32_t>> spec_tree(group);
    std::vector<uint8_t> spec_reached(group);
    std::vector<uint64_t> rip_offset(group + 1);
    for (uint32_t g0 = 0; g0 < net_count; g0 += group) {
      const uint32_t g1 = static_cast<uint32_t>(std::min<uint64_t>(net_count, uint64_t{g0} + group));
      members.clear();
      for (uint32_t n = g0; n < g1; ++n) members.push_back(n);
      // Snapshots: the working plane starts as the committed usage; member k's
      // snapshot follows the rip-ups of members 0..k.
      upload.clear();
      for (uint32_t k = 0; k < members.size(); ++k) {
        rip_offset[k] = upload.size();
        upload.insert(upload.end(), neg->routes[members[k]].begin(), neg->routes[members[k]].end());
      }
      rip_offset[members.size()] = upload.size();
      if (int rc = ensure_list(std::max<uint64_t>(upload.size(), 1)); rc != APGX_OK) return rc;
      APGX_HIP(hipMemcpyAsync(b.work, b.usage, 4 * cells, hipMemcpyDeviceToDevice, stream), "usage copy");
      if (!upload.empty())
        APGX_HIP(hipMemcpyAsync(b.list, upload.data(), 4 * upload.size(), hipMemcpyHostToDevice, stream),
                 "rip-up upload");
      APGX_HIP(hipMemsetAsync(b.slot_max, 0, 4 * group, stream), "slot max clear");
      for (uint32_t k = 0; k < members.size(); ++k) {
        if (int rc = launch_scatter(b.work, b.list + rip_offset[k], rip_offset[k + 1] - rip_offset[k], -1);
            rc != APGX_OK)
          return rc;
        if (int rc = launch_window_snapshot(members[k], b.work, b.slots + offsets[k], b.slot_max + k, pf);
            rc != APGX_OK)
          return rc;
      }
      APGX_HIP(hipMemcpyAsync(slot_max.data(), b.slot_max, 4 * group, hipMemcpyDeviceToHost, stream),
               "slot max readback");
      APGX_HIP(hipStreamSynchronize(stream), "hipStreamSynchronize");
      uint32_t max_cost = lat->min_cell_cost;
      for (uint32_t k = 0; k < members.size(); ++k) max_cost = std::max(max_cost, slot_max[k]);
      CostOverride ov{b.slots, max_cost, 1u};
      ov.net_offset = offsets.data();
      if (int rc = search(members, ov); rc != APGX_OK) return rc;
      sum.group_routed_nets += members.size();
      for (uint32_t k = 0; k < members.size(); ++k)
        spec_reached[k] = installed_route(k, members[k], &spec_tree[k]) ? 1u : 0u;
      const uint32_t start_seq = seq;
      // Commit in order.
      for (uint32_t k = 0; k < members.size(); ++k) {
        const uint32_t n = members[k];
        // Snapshot instant of n: its rip-up, then the cells the previous net
        // installed are the only ones whose cost can have risen since the
        // previous instant.
        for (uint32_t c : neg->routes[n]) --host_usage[c];
        if (n > 0)
          for (uint32_t c : prev_installed)
            if (!cell_within_bounds(c, pf))
              return Fail(APGX_E_RESOURCE_EXHAUSTED,
                          "apgx_negotiate: congestion cost exceeds the configured value bound");
        // A net that is not reached keeps its incumbent: reachability does not
        // depend on the costs, so that holds on the exact snapshot too.
        bool reached = spec_reached[k] != 0;
        bool conflict = false;
        if (reached)
          for (uint32_t c : spec_tree[k]) conflict |= stamp[c] > start_seq;
        tree.swap(spec_tree[k]);
        pending_rip.insert(pending_rip.end(), neg->routes[n].begin(), neg->routes[n].end());
        if (conflict) {
          // Searched again on the exact snapshot: every earlier net committed.
          if (int rc = flush(); rc != APGX_OK) return rc;
          uint32_t redo_max = 0;
          APGX_HIP(hipMemsetAsync(b.slot_max, 0, 4, stream), "slot max clear");
          if (int rc = launch_window_snapshot(n, b.usage, b.slots, b.slot_max, pf); rc != APGX_OK) return rc;
          APGX_HIP(hipMemcpyAsync(&redo_max, b.slot_max, 4, hipMemcpyDeviceToHost, stream), "slot max readback");
          APGX_HIP(hipStreamSynchronize(stream), "hipStreamSynchronize");
          const std::vector<uint32_t> one{n};
          const CostOverride exact{b.slots, std::max(redo_max, lat->min_cell_cost), 1u};
          if (int rc = search(one, exact); rc != APGX_OK) return rc;
          ++sum.redone_nets;
          reached = installed_route(0, n, &tree);
        }
        if (reached) neg->routed[n] = 1;
        neg->routes[n] = tree;
        ++seq;
        for (uint32_t c : tree) {
          stamp[c] = seq;
          ++host_usage[c];
        }
        pending_install.insert(pending_install.end(), tree.begin(), tree.end());
        prev_installed = tree;
      }
      if (int rc = flush(); rc != APGX_OK) return rc;
    }
    return APGX_OK;
  };

  for (uint32_t pass = 0; pass < pol.maximum_passes; ++pass) {
    uint64_t pf = 0;
    if (__builtin_mul_overflow(static_cast<uint64_t>(pass), pol.present_cost_increment, &pf) ||
        __builtin_add_overflow(pf, pol.initial_present_cost, &pf))
      return Fail(APGX_E_RESOURCE_EXHAUSTED, "apgx_negotiate: present factor overflowed uint64");
    // A zero present factor prices no cell by its usage: the snapshot is base +
    // history for every batch of the pass, no route depends on another, and
    // one batch of every net gives the routes of any order.
    if (pf != 0 && group > 0 && net_count > 0) {
      if (int rc = group_pass(pf); rc != APGX_OK) return rc;
    }
    static const std::vector<std::vector<uint32_t>> kNoBatches;
    const std::vector<std::vector<uint32_t>>& pass_batches = pf == 0 ? all_nets : (group > 0 ? kNoBatches : batches);
    for (const std::vector<uint32_t>& members : pass_batches) {
      const uint32_t count = static_cast<uint32_t>(members.size());
      rip_list.clear();
      for (uint32_t n : members) rip_list.insert(rip_list.end(), neg->routes[n].begin(), neg->routes[n].end());
      if (int rc = ensure_list(std::max<uint64_t>(rip_list.size(), 1)); rc != APGX_OK) return rc;
      if (!rip_list.empty())
        APGX_HIP(hipMemcpyAsync(b.list, rip_list.data(), 4 * rip_list.size(), hipMemcpyHostToDevice, stream),
                 "rip-up upload");
      const BatchParams params{b.list, rip_list.size(), pf};
      APGX_HIP(hipMemcpyAsync(b.params, &params, sizeof(params), hipMemcpyHostToDevice, stream), "params upload");
      host_flags = Flags{0, 0, 0};
      APGX_HIP(hipMemcpyAsync(b.flags, &host_flags, sizeof(Flags), hipMemcpyHostToDevice, stream), "flags clear");
      APGX_HIP(hipGraphLaunch(guard.g, stream), "hipGraphLaunch");
      APGX_HIP(hipMemcpyAsync(&host_flags, b.flags, sizeof(Flags), hipMemcpyDeviceToHost, stream), "flags readback");
      APGX_HIP(hipStreamSynchronize(stream), "hipStreamSynchronize");
      if (host_flags.bound)
        return Fail(APGX_E_RESOURCE_EXHAUSTED,
                    "apgx_negotiate: congestion cost exceeds the configured value bound");
      // Route the batch on the snapshot plane.
      const CostOverride ov{b.snap, std::max(host_flags.max_cost, lat->min_cell_cost), 1u};
      if (int rc = search(members, ov); rc != APGX_OK) return rc;
      // Install new routes (a disconnected or invalid net keeps its incumbent).
      install_list.clear();
      std::vector<uint32_t> tree;
      for (uint32_t i = 0; i < count; ++i) {
        const uint32_t n = members[i];
        if (installed_route(i, n, &tree)) {
          neg->routes[n] = tree;
          neg->routed[n] = 1;
        }
        install_list.insert(install_list.end(), neg->routes[n].begin(), neg->routes[n].end());
      }
      if (!install_list.empty()) {
        if (int e = ensure_list(install_list.size()); e != APGX_OK) return e;
        APGX_HIP(hipMemcpyAsync(b.list, install_list.data(), 4 * install_list.size(), hipMemcpyHostToDevice,
                                stream), "install upload");
        if (int e = launch_scatter(b.usage, b.list, install_list.size(), 1); e != APGX_OK) return e;
      }
    }
    // Accounting at the pass boundary.
    host_flags = Flags{0, 0, 0};
    APGX_HIP(hipMemcpyAsync(b.flags, &host_flags, sizeof(Flags), hipMemcpyHostToDevice, stream), "flags clear");
    hipLaunchKernelGGL(OveruseKernel, dim3(grid), dim3(kBlock), 0, stream, static_cast<const uint8_t*>(b.cap),
                       static_cast<const uint32_t*>(b.usage), cells, b.flags);
    APGX_HIP(hipGetLastError(), "overuse launch");
    APGX_HIP(hipMemcpyAsync(&host_flags, b.flags, sizeof(Flags), hipMemcpyDeviceToHost, stream), "flags readback");
    APGX_HIP(hipStreamSynchronize(stream), "hipStreamSynchronize");
    const uint64_t overuse = host_flags.overuse;
    uint32_t routed = 0;
    for (uint32_t n = 0; n < net_count; ++n) routed += neg->routed[n];
    neg->overuse_per_pass.push_back(overuse);
    sum.completed_passes = pass + 1;
    sum.total_overuse_units = overuse;
    sum.routed_nets = routed;
    if (pol.on_pass)
      pol.on_pass(pol.on_pass_context, pass, overuse, routed,
                  std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - wall_start).count());
    if (routed == net_count && overuse == 0) {
      sum.termination = APGX_NEGOTIATION_FEASIBLE;
      break;
    }
    if (have_previous && previous_routed == neg->routed && previous == neg->routes) {
      sum.termination = APGX_NEGOTIATION_STALLED;
      break;
    }
    if (pass + 1 == pol.maximum_passes) {
      sum.termination = APGX_NEGOTIATION_PASS_LIMIT;
      break;
    }
    previous = neg->routes;
    previous_routed = neg->routed;
    have_previous = true;
    host_flags = Flags{0, 0, 0};
    APGX_HIP(hipMemcpyAsync(b.flags, &host_flags, sizeof(Flags), hipMemcpyHostToDevice, stream), "flags clear");
    hipLaunchKernelGGL(HistoryKernel, dim3(grid), dim3(kBlock), 0, stream, static_cast<const uint8_t*>(b.cap),
                       static_cast<const uint32_t*>(b.usage), b.history, cells, pol.historical_cost_increment,
                       max_cong, b.flags);
    APGX_HIP(hipGetLastError(), "history launch");
    APGX_HIP(hipMemcpyAsync(&host_flags, b.flags, sizeof(Flags), hipMemcpyDeviceToHost, stream), "flags readback");
    APGX_HIP(hipStreamSynchronize(stream), "hipStreamSynchronize");
    if (host_flags.bound)
      return Fail(APGX_E_RESOURCE_EXHAUSTED, "apgx_negotiate: historical congestion cost exceeds the value bound");
  }
  neg->history.resize(cells);
  neg->usage.resize(cells);
  APGX_HIP(hipMemcpy(neg->history.data(), b.history, 4 * cells, hipMemcpyDeviceToHost), "history readback");
  APGX_HIP(hipMemcpy(neg->usage.data(), b.usage, 4 * cells, hipMemcpyDeviceToHost), "usage readback");
  sum.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - wall_start).count();
  *out = owner.release();
  ClearError();
  return APGX_OK;
}

extern "C" int apgx_negotiation_get_summary(const apgx_negotiation* n, apgx_negotiation_summary* out,
                                            uint64_t* overuse_per_pass, uint32_t capacity) {
  using namespace apgx;
  if (n == nullptr || out == nullptr) return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_negotiation_get_summary: null argument");
  *out = n->summary;
  if (overuse_per_pass)
    for (uint32_t i = 0; i < std::min<size_t>(capacity, n->overuse_per_pass.size()); ++i)
      overuse_per_pass[i] = n->overuse_per_pass[i];
  ClearError();
  return APGX_OK;
}

extern "C" int apgx_negotiation_planes(const apgx_negotiation* n, uint32_t* history, uint32_t* usage) {
  using namespace apgx;
  if (n == nullptr) return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_negotiation_planes: null argument");
  if (history) std::memcpy(history, n->history.data(), 4 * n->cells);
  if (usage) std::memcpy(usage, n->usage.data(), 4 * n->cells);
  ClearError();
  return APGX_OK;
}

extern "C" int apgx_negotiation_routes(const apgx_negotiation* n, uint32_t* route_len, uint32_t* cells,
                                       uint64_t cell_capacity, uint64_t* cell_count) {
  using namespace apgx;
  if (n == nullptr) return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_negotiation_routes: null argument");
  uint64_t total = 0;
  for (const auto& r : n->routes) total += r.size();
  if (cell_count) *cell_count = total;
  if (cells && total > cell_capacity)
    return Fail(APGX_E_RESOURCE_EXHAUSTED, "apgx_negotiation_routes: cell capacity too small");
  uint64_t w = 0;
  for (uint32_t i = 0; i < n->net_count; ++i) {
    if (route_len) route_len[i] = static_cast<uint32_t>(n->routes[i].size());
    if (cells && !n->routes[i].empty()) std::memcpy(cells + w, n->routes[i].data(), 4 * n->routes[i].size());
    w += n->routes[i].size();
  }
  ClearError();
  return APGX_OK;
}

extern "C" void apgx_negotiation_release(apgx_negotiation* n) { delete n; }
