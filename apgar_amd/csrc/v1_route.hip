// v1_route.hip — single routes on an uploaded V1 view (ExecuteRoute /
// ReadbackRoute, gpu/cuda/planar_router.cu:2049-2181 and 2794-2853).
//
// The reference runs one kernel launch per round and copies a stats word back
// to the host after every round.  Here one workgroup runs the whole round loop
// of a route in a single launch: rounds end at workgroup barriers, the round
// statistics are LDS reductions, and cancellation is read once per round from
// a host-mapped word that the calling thread mirrors from the caller's flag.
// The round semantics are the reference's, so rounds, examined work and turn
// relaxations are its figures:
//
//  * Bucketed frontier (FrontierRelax / RunFrontier, :363-412, :1552-1640):
//    delta-stepping over heading states.  A state is active in round k when
//    round k-1 deferred it or lowered its label.  An active state whose label
//    lies in the bucket relaxes its legal edges, the others are deferred; the
//    next bucket starts at the smallest deferred or lowered label rounded down
//    to a multiple of delta = min(orthogonal, diagonal step).  Every candidate
//    made inside the bucket lies above it (each step costs at least delta), so
//    which states run in a round, and the labels they run with, do not depend
//    on the order of the atomics.
//  * Heading-aware sweep (ComputeDepartureLabels / SweepRuns / RunSweep,
//    :414-484, :1642-1722): departures, then segmented min-plus scans along
//    the view's maximal runs, until a round lowers no label.
//
// Then every state's lowest tight predecessor (SelectStablePredecessors,
// :486-524) in a grid-wide kernel.  Device bytes follow the reference's
// deterministic accounting exactly: result header, labels, predecessors, and
// the generator's scratch (two round marks per state and two words, or eight
// departures per node and two words).
#include <hip/hip_runtime.h>

#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#include "apgx.h"
#include "common.h"
#include "v1_common.h"
#include "v1_view.h"

namespace apgx {
namespace v1route {
namespace {

constexpr uint64_t kInf64 = ~0ull;
constexpr uint32_t kInvalid = APGX_V1_INVALID_INDEX;
constexpr int kRouteThreads = 1024;
constexpr uint32_t kStopBudget = 0, kStopConverged = 1, kStopCancelled = 2;

static_assert(sizeof(apgx_v1_route_header) == 64, "DeviceRouteResultHeaderV1 is 64 bytes");
static_assert(sizeof(apgx_v1_kernel_telemetry) == 56, "KernelTelemetry mirror");

struct RouteArgs {
  const apgx_v1_node* nodes;
  const apgx_v1_run* runs;
  const uint32_t* run_nodes;
  uint64_t node_count;
  uint64_t run_count;
  uint32_t orth, diag, bend;
  uint32_t start_node, goal_node, max_rounds;
  uint64_t* labels;      // node_count * 9
  uint32_t* preds;       // node_count * 9
  uint32_t* marks[2];    // frontier: round marks, alternating by round parity
  uint64_t* departures;  // sweep: node_count * 8
  // Two words: frontier {examined work, rounds | stop << 32}; sweep
  // {turn relaxations, rounds | stop << 32}.
  unsigned long long* stats;
  const uint32_t* cancel;  // host-mapped; nullptr when the route cannot be cancelled
};

__device__ __forceinline__ bool CheckedAdd(uint64_t a, uint64_t b, uint64_t* out) {
  const uint64_t max_finite = kInf64 - 1;
  if (a > max_finite || b > max_finite - a) return false;
  *out = a + b;
  return true;
}

// DeviceStepCost (:318-326): step of the direction, plus the bend when the
// state was entered with another heading (none at the start).
__device__ __forceinline__ uint64_t StepCost(const RouteArgs& a, uint32_t dir, uint32_t incoming) {
  uint64_t c = (dir & 1u) ? a.diag : a.orth;
  if (incoming != 8u && incoming != dir) c += a.bend;
  return c;
}

__device__ __forceinline__ bool Cancelled(const RouteArgs& a) {
  return a.cancel != nullptr && __hip_atomic_load(a.cancel, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM) != 0;
}

__device__ __forceinline__ uint64_t LoadLabel(const uint64_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint32_t LoadMark(const uint32_t* p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void StoreMark(uint32_t* p, uint32_t v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// InitializeResultHeader + InitializeSearch (:295-316, :349-361): labels
// unreachable except the start state, no predecessors, the start state active
// in round 0 (mark 1), the other marks clear.
__global__ void RouteInitKernel(RouteArgs a, apgx_v1_route_header header, apgx_v1_route_header* out) {
  const uint64_t S = a.node_count * 9;
  const uint64_t start_state = static_cast<uint64_t>(a.start_node) * 9 + 8;
  for (uint64_t s = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; s < S;
       s += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    a.labels[s] = s == start_state ? 0 : kInf64;
    a.preds[s] = kInvalid;
    if (a.marks[0] != nullptr) {
      a.marks[0][s] = s == start_state ? 1u : 0u;
      a.marks[1][s] = 0;
    }
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    *out = header;
    a.stats[0] = 0;
    a.stats[1] = 0;
  }
}

// The whole delta-stepping loop in one workgroup.  Round k processes the
// states whose mark is k + 1 (marks of round k - 1, or the start's initial
// mark) and marks the states it defers or lowers with k + 2 in the other
// array: exactly the reference's generation-stamped active arrays.
__global__ __launch_bounds__(kRouteThreads) void RouteFrontierKernel(RouteArgs a) {
  __shared__ unsigned long long s_min, s_work;
  __shared__ uint32_t s_cancel;
  const uint64_t S = a.node_count * 9;
  const uint64_t delta = a.orth < a.diag ? a.orth : a.diag;
  uint64_t bucket_upper = 0, examined = 0;
  uint32_t rounds = 0, stop = kStopBudget;
  for (uint32_t round = 0; round < a.max_rounds; ++round) {
    if (threadIdx.x == 0) {
      s_min = kInf64;
      s_work = 0;
      s_cancel = Cancelled(a) ? 1u : 0u;
    }
    __syncthreads();
    if (s_cancel) {
      stop = kStopCancelled;
      break;
    }
    const uint32_t* cur = a.marks[round & 1u];
    uint32_t* nxt = a.marks[(round + 1u) & 1u];
    const uint32_t here = round + 1u, there = round + 2u;
    uint64_t local_min = kInf64, local_work = 0;
    for (uint64_t s = threadIdx.x; s < S; s += kRouteThreads) {
      if (LoadMark(&cur[s]) != here) continue;
      const uint64_t label = LoadLabel(&a.labels[s]);
      if (label == kInf64) continue;
      if (label > bucket_upper) {
        StoreMark(&nxt[s], there);
        local_min = label < local_min ? label : local_min;
        continue;
      }
      const uint32_t node = static_cast<uint32_t>(s / 9), incoming = static_cast<uint32_t>(s % 9);
      const apgx_v1_node nd = a.nodes[node];
      for (uint32_t dir = 0; dir < 8; ++dir) {
        if ((nd.legal_edges & (1u << dir)) == 0) continue;
        ++local_work;
        uint64_t cand;
        if (!CheckedAdd(label, StepCost(a, dir, incoming), &cand)) continue;
        const uint64_t target = static_cast<uint64_t>(nd.neighbors[dir]) * 9 + dir;
        const unsigned long long prev =
            atomicMin(reinterpret_cast<unsigned long long*>(&a.labels[target]), static_cast<unsigned long long>(cand));
        if (cand < prev) {
          StoreMark(&nxt[target], there);
          local_min = cand < local_min ? cand : local_min;
        }
      }
    }
    if (local_min != kInf64) atomicMin(&s_min, static_cast<unsigned long long>(local_min));
    if (local_work) atomicAdd(&s_work, static_cast<unsigned long long>(local_work));
    __syncthreads();
    examined += s_work;
    rounds = round + 1;
    const uint64_t next_min = s_min;
    // The round's stats are the batch boundary: cancellation is resampled
    // before convergence or the budget is classified.
    if (threadIdx.x == 0) s_cancel = Cancelled(a) ? 1u : 0u;
    __syncthreads();
    if (s_cancel) {
      stop = kStopCancelled;
      break;
    }
    if (next_min == kInf64) {
      stop = kStopConverged;
      break;
    }
    const uint64_t base = (next_min / delta) * delta;
    bucket_upper = base > kInf64 - (delta - 1) ? kInf64 : base + delta - 1;
  }
  if (threadIdx.x == 0) {
    a.stats[0] = examined;
    a.stats[1] = static_cast<unsigned long long>(rounds) | (static_cast<unsigned long long>(stop) << 32);
  }
}

// The whole sweep loop in one workgroup: per round the departure of every
// (node, direction) from the node's labels (the start node departs at 0; a
// turn adds the bend), then a segmented min-plus scan along every maximal
// run.  Each (target, heading) label has one incoming run edge, so every
// label has a single writer.  A departure counts as a turn relaxation when
// its final minimum came from another heading.
__global__ __launch_bounds__(kRouteThreads) void RouteSweepKernel(RouteArgs a) {
  uint64_t turns = 0;
  uint32_t rounds = 0, stop = kStopBudget;
  __shared__ uint32_t s_cancel;
  const uint64_t items = a.node_count * 8;
  for (uint32_t round = 0; round < a.max_rounds; ++round) {
    if (threadIdx.x == 0) s_cancel = Cancelled(a) ? 1u : 0u;
    __syncthreads();
    if (s_cancel) {
      stop = kStopCancelled;
      break;
    }
    for (uint64_t item = threadIdx.x; item < items; item += kRouteThreads) {
      const uint32_t node = static_cast<uint32_t>(item >> 3), dir = static_cast<uint32_t>(item & 7u);
      uint64_t best = node == a.start_node ? 0 : a.labels[node * 9ull + dir];
      bool used_turn = false;
      for (uint32_t inc = 0; inc < 8; ++inc) {
        const uint64_t label = a.labels[node * 9ull + inc];
        if (label == kInf64) continue;
        uint64_t cand;
        if (CheckedAdd(label, inc == dir ? 0 : a.bend, &cand) && cand < best) {
          best = cand;
          used_turn = inc != dir;
        }
      }
      a.departures[item] = best;
      if (used_turn) ++turns;
    }
    __syncthreads();
    bool changed = false;
    for (uint64_t r = threadIdx.x; r < a.run_count; r += kRouteThreads) {
      const apgx_v1_run run = a.runs[r];
      const uint32_t dir = run.direction;
      const uint64_t step = (dir & 1u) ? a.diag : a.orth;
      uint64_t carry = a.departures[static_cast<uint64_t>(a.run_nodes[run.node_offset]) * 8 + dir];
      for (uint32_t pos = 0; pos + 1 < run.node_count; ++pos) {
        const uint32_t source = a.run_nodes[run.node_offset + pos];
        const uint32_t target = a.run_nodes[run.node_offset + pos + 1];
        const uint64_t dep_s = a.departures[static_cast<uint64_t>(source) * 8 + dir];
        carry = dep_s < carry ? dep_s : carry;
        const uint64_t dep_t = a.departures[static_cast<uint64_t>(target) * 8 + dir];
        uint64_t cand;
        if (CheckedAdd(carry, step, &cand)) {
          uint64_t& slot = a.labels[static_cast<uint64_t>(target) * 9 + dir];
          if (cand < slot) {
            slot = cand;
            changed = true;
          }
          carry = cand < dep_t ? cand : dep_t;
        } else {
          carry = dep_t;
        }
      }
    }
    changed = __syncthreads_or(changed);
    rounds = round + 1;
    if (threadIdx.x == 0) s_cancel = Cancelled(a) ? 1u : 0u;
    __syncthreads();
    if (s_cancel) {
      stop = kStopCancelled;
      break;
    }
    if (!changed) {
      stop = kStopConverged;
      break;
    }
  }
  if (turns) atomicAdd(&a.stats[0], static_cast<unsigned long long>(turns));
  if (threadIdx.x == 0)
    a.stats[1] = static_cast<unsigned long long>(rounds) | (static_cast<unsigned long long>(stop) << 32);
}

// SelectStablePredecessors (:486-524): the lowest-index predecessor state p
// of the node behind the incoming heading (heading 8 only at the start node)
// with label[p] + step == label[s].
__global__ void RoutePredecessorsKernel(RouteArgs a) {
  const uint64_t S = a.node_count * 9;
  const uint64_t start_state = static_cast<uint64_t>(a.start_node) * 9 + 8;
  for (uint64_t s = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; s < S;
       s += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint32_t node = static_cast<uint32_t>(s / 9), inc = static_cast<uint32_t>(s % 9);
    const uint64_t label = a.labels[s];
    if (s == start_state || inc == 8u || label == kInf64) {
      a.preds[s] = kInvalid;
      continue;
    }
    const uint32_t pnode = a.nodes[node].neighbors[(inc + 4u) % 8u];
    if (pnode == kInvalid) {
      a.preds[s] = kInvalid;
      continue;
    }
    uint32_t best = kInvalid;
    for (uint32_t ph = 0; ph < 9; ++ph) {
      if (ph == 8u && pnode != a.start_node) continue;
      const uint32_t ps = pnode * 9u + ph;
      uint64_t cand;
      if (CheckedAdd(a.labels[ps], StepCost(a, inc, ph), &cand) && cand == label && ps < best) best = ps;
    }
    a.preds[s] = best;
  }
}

}  // namespace

// Deterministic device bytes of a route (ValidateKernelTelemetry's formula).
uint64_t RouteBytes(bool frontier, uint64_t node_count) {
  const uint64_t S = node_count * 9;
  return sizeof(apgx_v1_route_header) + S * (8 + 4) + (frontier ? S * 2 * 4 + 16 : node_count * 8 * 8 + 16);
}

}  // namespace v1route
}  // namespace apgx

struct apgx_v1_route_pending {
  apgx_v1_view* view = nullptr;
  std::unique_lock<std::mutex> lease;
  void* d_block = nullptr;
  apgx_v1_route_header* d_header = nullptr;
  uint64_t* d_labels = nullptr;
  uint32_t* d_preds = nullptr;
  uint64_t state_count = 0;
  uint32_t completion = APGX_V1_DISCONNECTED;
  uint32_t fault = APGX_V1_ROUTE_FAULT_NONE;
  apgx_v1_kernel_telemetry telemetry{};
  ~apgx_v1_route_pending() {
    if (d_block) (void)hipFree(d_block);
  }
};

extern "C" int apgx_v1_expected_route_bytes(uint32_t generator, uint64_t node_count, uint64_t* device_bytes) {
  using namespace apgx;
  if (device_bytes == nullptr) return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_v1_expected_route_bytes: null argument");
  v1::GeneratorShape shape;
  if (!v1::DescribeGenerator(generator, &shape))
    return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_v1_expected_route_bytes: unknown generator");
  *device_bytes = v1route::RouteBytes(generator == APGX_V1_GENERATOR_BUCKETED_FRONTIER, node_count);
  ClearError();
  return APGX_OK;
}

extern "C" int apgx_v1_route_execute(apgx_v1_view* view, const apgx_v1_route_request* req,
                                     const volatile uint32_t* cancel, apgx_v1_route_pending** out) {
  using namespace apgx;
  using namespace apgx::v1route;
  if (view == nullptr || req == nullptr || out == nullptr)
    return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_v1_route_execute: null argument");
  *out = nullptr;
  int device = -1;
  APGX_HIP(hipGetDevice(&device), "hipGetDevice(execute)");
  if (device != view->device)
    return Fail(APGX_E_BACKEND_FAILURE, "route execute current device differs from upload device");
  std::unique_lock<std::mutex> lease(view->route_lease);
  const uint64_t N = view->node_count, S = N * 9;
  if (req->start_node >= N || req->goal_node >= N || req->maximum_rounds == 0)
    return Fail(APGX_E_INTERNAL_INVARIANT, "route execute received invalid endpoint or budget bounds");
  if (req->start_node == req->goal_node)
    return Fail(APGX_E_INTERNAL_INVARIANT, "route execute requires distinct start and goal nodes");
  v1::GeneratorShape shape;
  if (!v1::DescribeGenerator(req->generator, &shape))
    return Fail(APGX_E_INTERNAL_INVARIANT, "route execute received an unknown generator");
  const bool frontier = req->generator == APGX_V1_GENERATOR_BUCKETED_FRONTIER;
  const apgx_v1_header& h = view->header;
  if (frontier && std::min(h.orthogonal_step, h.diagonal_step) == 0)
    return Fail(APGX_E_UNSUPPORTED, "route frontier needs positive step costs (its bucket width)");
  const uint64_t batch_bytes = RouteBytes(frontier, N);
  if (view->persistent_bytes > req->maximum_device_bytes ||
      batch_bytes > req->maximum_device_bytes - view->persistent_bytes)
    return Fail(APGX_E_RESOURCE_EXHAUSTED, "route exceeds the deterministic device-memory budget");

  auto p = std::make_unique<apgx_v1_route_pending>();
  p->view = view;
  p->state_count = S;
  // One block: header | labels | predecessors | scratch, 256-byte aligned sections.
  auto align = [](uint64_t v) { return (v + 255) & ~255ull; };
  const uint64_t off_labels = 256, off_preds = align(off_labels + S * 8), off_scratch = align(off_preds + S * 4);
  const uint64_t scratch = frontier ? align(S * 4) * 2 + 16 : N * 64 + 16;
  APGX_HIP(hipMalloc(&p->d_block, off_scratch + scratch + 256), "route workspace allocation");
  char* base = static_cast<char*>(p->d_block);
  p->d_header = reinterpret_cast<apgx_v1_route_header*>(base);
  p->d_labels = reinterpret_cast<uint64_t*>(base + off_labels);
  p->d_preds = reinterpret_cast<uint32_t*>(base + off_preds);

  RouteArgs a{};
  a.nodes = view->d_nodes;
  a.runs = view->d_runs;
  a.run_nodes = view->d_run_nodes;
  a.node_count = N;
  a.run_count = view->run_count;
  a.orth = h.orthogonal_step;
  a.diag = h.diagonal_step;
  a.bend = h.bend;
  a.start_node = req->start_node;
  a.goal_node = req->goal_node;
  a.max_rounds = req->maximum_rounds;
  a.labels = p->d_labels;
  a.preds = p->d_preds;
  if (frontier) {
    a.marks[0] = reinterpret_cast<uint32_t*>(base + off_scratch);
    a.marks[1] = reinterpret_cast<uint32_t*>(base + off_scratch + align(S * 4));
    a.stats = reinterpret_cast<unsigned long long*>(base + off_scratch + 2 * align(S * 4));
  } else {
    a.departures = reinterpret_cast<uint64_t*>(base + off_scratch);
    a.stats = reinterpret_cast<unsigned long long*>(base + off_scratch + N * 64);
  }
  if (cancel != nullptr) {
    if (view->cancel_word == nullptr)
      APGX_HIP(hipHostMalloc(reinterpret_cast<void**>(&view->cancel_word), 64,
                             hipHostMallocMapped | hipHostMallocCoherent),
               "cancellation word allocation");
    uint32_t* dev_word = nullptr;
    APGX_HIP(hipHostGetDevicePointer(reinterpret_cast<void**>(&dev_word), view->cancel_word, 0),
             "cancellation word mapping");
    __atomic_store_n(view->cancel_word, *cancel ? 1u : 0u, __ATOMIC_SEQ_CST);
    a.cancel = dev_word;
  }
  apgx_v1_route_header header{};
  header.schema_version = APGX_V1_ROUTE_RESULT_SCHEMA_VERSION;
  header.compiler_version = h.compiler_version;
  header.start_node = req->start_node;
  header.goal_node = req->goal_node;
  header.source_board_content_hash = h.source_board_content_hash;
  header.compiler_profile_fingerprint = h.compiler_profile_fingerprint;
  header.routing_profile_fingerprint = req->routing_profile_fingerprint;
  header.rule_bucket_identity = h.rule_bucket_identity;
  header.device_view_fingerprint = h.device_view_fingerprint;
  header.generator = static_cast<uint8_t>(req->generator);

  hipEvent_t ev[2] = {nullptr, nullptr};
  APGX_HIP(hipEventCreate(&ev[0]), "hipEventCreate(start)");
  hipError_t e = hipEventCreate(&ev[1]);
  const uint32_t grid = static_cast<uint32_t>(std::min<uint64_t>((S + 255) / 256, 1024));
  if (e == hipSuccess) e = hipEventRecord(ev[0], nullptr);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(RouteInitKernel, dim3(grid), dim3(256), 0, nullptr, a, header, p->d_header);
    e = hipGetLastError();
  }
  if (e == hipSuccess) {
    if (frontier)
      hipLaunchKernelGGL(RouteFrontierKernel, dim3(1), dim3(kRouteThreads), 0, nullptr, a);
    else
      hipLaunchKernelGGL(RouteSweepKernel, dim3(1), dim3(kRouteThreads), 0, nullptr, a);
    e = hipGetLastError();
  }
  if (e == hipSuccess) {
    hipLaunchKernelGGL(RoutePredecessorsKernel, dim3(grid), dim3(256), 0, nullptr, a);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipEventRecord(ev[1], nullptr);
  // Wait; a cancellable route mirrors the caller's flag into the mapped word
  // while it runs.
  if (e == hipSuccess && a.cancel != nullptr) {
    for (;;) {
      const hipError_t q = hipEventQuery(ev[1]);
      if (q == hipSuccess) break;
      if (q != hipErrorNotReady) {
        e = q;
        break;
      }
      if (*cancel) __atomic_store_n(view->cancel_word, 1u, __ATOMIC_SEQ_CST);
      std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
  }
  if (e == hipSuccess) e = hipEventSynchronize(ev[1]);
  float ms = 0.f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, ev[0], ev[1]);
  unsigned long long stats[2] = {0, 0};
  uint64_t goal_labels[8];
  if (e == hipSuccess) e = hipMemcpy(stats, a.stats, sizeof(stats), hipMemcpyDeviceToHost);
  if (e == hipSuccess)
    e = hipMemcpy(goal_labels, p->d_labels + static_cast<uint64_t>(req->goal_node) * 9, sizeof(goal_labels),
                  hipMemcpyDeviceToHost);
  (void)hipEventDestroy(ev[0]);
  if (ev[1]) (void)hipEventDestroy(ev[1]);
  if (e != hipSuccess) return HipFail(e, "route execution");

  // CompleteBoundedSearch / DetermineCompletion / FinishExecution (:1455-1550).
  const uint32_t rounds = static_cast<uint32_t>(stats[1] & 0xFFFFFFFFull);
  const uint32_t stop = static_cast<uint32_t>(stats[1] >> 32);
  if (stop == kStopCancelled) {
    p->completion = APGX_V1_CANCELLED;
  } else if (stop == kStopConverged) {
    bool any = false;
    for (uint64_t l : goal_labels) any |= l != kInf64;
    p->completion = any ? APGX_V1_REACHED : APGX_V1_DISCONNECTED;
  } else {
    p->completion = APGX_V1_BUDGET_EXHAUSTED;
  }
  if (cancel != nullptr && *cancel) p->completion = APGX_V1_CANCELLED;
  apgx_v1_kernel_telemetry& t = p->telemetry;
  t.persistent_device_bytes = view->persistent_bytes;
  t.batch_device_bytes = batch_bytes;
  t.peak_device_bytes = view->persistent_bytes + batch_bytes;
  t.rounds = rounds;
  t.examined_work = frontier ? stats[0] : static_cast<uint64_t>(rounds) * (view->run_count ? view->legal_edges : 0);
  t.heading_turn_relaxations = frontier ? 0 : stats[0];
  t.kernel_milliseconds = ms;
  p->lease = std::move(lease);
  *out = p.release();
  ClearError();
  return APGX_OK;
}

extern "C" int apgx_v1_route_readback(apgx_v1_route_pending* p, apgx_v1_route_result* result, uint64_t* labels,
                                      uint32_t* predecessors, uint64_t capacity) {
  using namespace apgx;
  if (p == nullptr || result == nullptr)
    return Fail(APGX_E_INTERNAL_INVARIANT, "route readback received a null job");
  int device = -1;
  APGX_HIP(hipGetDevice(&device), "hipGetDevice(readback)");
  if (device != p->view->device)
    return Fail(APGX_E_BACKEND_FAILURE, "route readback current device differs from execution device");
  if ((labels || predecessors) && capacity < p->state_count)
    return Fail(APGX_E_RESOURCE_EXHAUSTED, "route readback arrays are smaller than the represented states");
  std::memset(result, 0, sizeof(*result));
  APGX_HIP(hipMemcpy(&result->header, p->d_header, sizeof(result->header), hipMemcpyDeviceToHost),
           "hipMemcpy(route result header)");
  result->completion = p->completion;
  result->goal_state = APGX_V1_INVALID_INDEX;
  result->state_count = p->state_count;
  result->telemetry = p->telemetry;
  std::vector<uint64_t> lab;
  if (labels == nullptr && p->completion == APGX_V1_REACHED) {
    lab.resize(p->state_count);
    labels = lab.data();
  }
  if (labels)
    APGX_HIP(hipMemcpy(labels, p->d_labels, 8 * p->state_count, hipMemcpyDeviceToHost), "hipMemcpy(route labels)");
  if (predecessors)
    APGX_HIP(hipMemcpy(predecessors, p->d_preds, 4 * p->state_count, hipMemcpyDeviceToHost),
             "hipMemcpy(route predecessors)");
  if (p->completion == APGX_V1_REACHED) {
    // The goal state: least goal label, lowest state index on a tie (:2838-2848).
    uint64_t best = ~0ull;
    for (uint32_t heading = 0; heading < 8; ++heading) {
      const uint32_t state = result->header.goal_node * 9u + heading;
      if (state < p->state_count && (labels[state] < best || (labels[state] == best && state < result->goal_state))) {
        best = labels[state];
        result->goal_state = state;
      }
    }
  }
  // Replay/test decorator: the goal names itself as its predecessor.
  if (p->fault == APGX_V1_ROUTE_FAULT_GOAL_PREDECESSOR_SELF_CYCLE && predecessors &&
      result->completion == APGX_V1_REACHED && result->goal_state < p->state_count)
    predecessors[result->goal_state] = result->goal_state;
  ClearError();
  return APGX_OK;
}

extern "C" int apgx_v1_route_pending_inject_fault(apgx_v1_route_pending* p, uint32_t fault) {
  using namespace apgx;
  if (p == nullptr) return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_v1_route_pending_inject_fault: null job");
  if (fault > APGX_V1_ROUTE_FAULT_GOAL_PREDECESSOR_SELF_CYCLE)
    return Fail(APGX_E_UNSUPPORTED, "apgx_v1_route_pending_inject_fault: unknown fault");
  p->fault = fault;
  ClearError();
  return APGX_OK;
}

extern "C" void apgx_v1_route_pending_release(apgx_v1_route_pending* p) { delete p; }
