// v1_sweep.hip — heading-aware sweep and bucketed frontier over reference V1
// device views.
//
// Semantics follow the reference batch sweep exactly (departures,
// segmented run scans, round convergence, goal tie-break and lowest
// predecessor paths: gpu/cuda/planar_router.cu:847-1222), so labels, goal
// states, paths and even round counts match.  The schedule is MI355X-shaped:
// one workgroup owns one query and runs up to a chunk of rounds inside one
// launch (departure pass, barrier, run pass, barrier-or for convergence), so a
// chunk costs one launch instead of four launches per round, and the host
// samples cancellation and status once per chunk.
//
// Readback follows the reference's per-generator contract: the sweep returns
// compact canonical paths; the bucketed frontier returns its full query-major
// workspace (labels, lowest tight predecessors and the two ownership stamps
// of every state), from which the host reconstructs and re-validates routes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <vector>

#include "apgx.h"
#include "common.h"
#include "v1_common.h"
#include "v1_view.h"

namespace apgx {
namespace v1 {
namespace {

constexpr uint64_t kInf64 = ~0ull;
constexpr uint32_t kInvalid = APGX_V1_INVALID_INDEX;
constexpr uint32_t kRunning = 0xFFu;
constexpr int kThreads = 256;

struct V1Args {
  const apgx_v1_node* nodes;
  const apgx_v1_run* runs;
  const uint32_t* run_nodes;
  uint64_t node_count;
  uint64_t run_count;
  uint64_t legal_edges;
  const apgx_v1_query* queries;
  const apgx_v1_policy_edge* policy;
  uint32_t query_count;
  uint64_t* labels;      // query-major, 9 * node_count per query
  uint64_t* departures;  // sweep: query-major, 8 * node_count per query (path scratch afterwards)
  uint32_t* predecessors;        // frontier: lowest tight predecessor of every state
  uint64_t* state_owners;        // frontier: ownership stamp of every label slot
  uint64_t* predecessor_owners;  // frontier: ownership stamp of every predecessor slot
  uint8_t* closed;               // frontier: closed flag of every state
  apgx_v1_result* results;
  apgx_v1_compact_path* paths;
  uint32_t* packed;
  uint32_t orth, diag, bend;
};

__device__ __forceinline__ bool CheckedAdd(uint64_t a, uint64_t b, uint64_t* out) {
  const uint64_t max_finite = kInf64 - 1;
  if (a > max_finite || b > max_finite - a) return false;
  *out = a + b;
  return true;
}

__device__ uint64_t PolicyAdjustment(const V1Args& a, const apgx_v1_query& q, uint32_t edge) {
  uint32_t first = q.policy_offset, last = q.policy_offset + q.policy_count;
  while (first < last) {
    const uint32_t mid = first + (last - first) / 2;
    if (a.policy[mid].physical_edge_index < edge) first = mid + 1; else last = mid;
  }
  if (first < q.policy_offset + q.policy_count && a.policy[first].physical_edge_index == edge)
    return a.policy[first].adjustment;
  return 0;
}

// Transition of the directed edge source -> neighbour(dir) entered with
// `incoming` heading (DeviceBatchTransitionCost semantics): step + surcharge,
// bend + bend surcharge on a heading change, policy adjustment of the canonical
// physical edge; banned or overflowing transitions do not exist.
__device__ bool Transition(const V1Args& a, const apgx_v1_query& q, uint32_t source, uint32_t dir,
                           uint32_t incoming, uint64_t* cost) {
  uint32_t edge;
  if (dir < 4) {
    edge = source * 8u + dir;
  } else {
    const uint32_t canonical = a.nodes[source].neighbors[dir];
    if (canonical == kInvalid) return false;
    edge = canonical * 8u + dir - 4u;
  }
  const uint64_t adj = PolicyAdjustment(a, q, edge);
  if (adj == APGX_V1_BANNED_ADJUSTMENT) return false;
  uint64_t v;
  if (!CheckedAdd((dir & 1u) ? a.diag : a.orth,
                  (dir & 1u) ? q.diagonal_step_surcharge : q.orthogonal_step_surcharge, &v))
    return false;
  if (incoming != 8u && incoming != dir) {
    uint64_t b;
    if (!CheckedAdd(a.bend, q.bend_surcharge, &b) || !CheckedAdd(v, b, &v)) return false;
  }
  return CheckedAdd(v, adj, cost);
}

__global__ void V1InitKernel(V1Args a, uint64_t total_states) {
  const uint64_t slot = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (slot >= total_states) return;
  const uint64_t states = a.node_count * 9;
  const uint32_t qi = static_cast<uint32_t>(slot / states);
  const uint64_t local = slot - qi * states;
  a.labels[slot] = local == static_cast<uint64_t>(a.queries[qi].start_node) * 9 + 8 ? 0 : kInf64;
  if (a.state_owners != nullptr) {  // frontier workspace (InitializeCandidateBatch)
    const uint64_t owner = a.queries[qi].workspace_owner;
    a.state_owners[slot] = owner;
    a.predecessor_owners[slot] = owner;
    a.predecessors[slot] = kInvalid;
    a.closed[slot] = 0;
  }
}

// Completion of the queries still running when the round loop stops: budget
// exhausted, or cancelled at a chunk boundary (FinalizeCandidateBatchStatus).
__global__ void V1FinalizeKernel(V1Args a, uint32_t completion) {
  const uint32_t qi = blockIdx.x * blockDim.x + threadIdx.x;
  if (qi >= a.query_count) return;
  apgx_v1_result& res = a.results[qi];
  if (res.reserved[0] != 0) return;
  res.completion = static_cast<uint8_t>(completion);
  res.reserved[0] = 1;
}

// Up to `chunk` rounds for one query per workgroup.
__global__ __launch_bounds__(kThreads) void V1SweepKernel(V1Args a, uint32_t chunk,
                                                         uint32_t round_budget) {
  const uint32_t qi = blockIdx.x;
  apgx_v1_result& res = a.results[qi];
  __shared__ uint32_t s_status;
  if (threadIdx.x == 0) s_status = res.reserved[0] == 0 ? kRunning : 0;
  __syncthreads();
  if (s_status != kRunning) return;
  const apgx_v1_query q = a.queries[qi];
  uint64_t* lab = a.labels + static_cast<uint64_t>(qi) * a.node_count * 9;
  uint64_t* dep = a.departures + static_cast<uint64_t>(qi) * a.node_count * 8;
  uint64_t turns = 0;
  for (uint32_t k = 0; k < chunk; ++k) {
    if (res.rounds >= round_budget) break;
    // Departures from this round's labels (DeviceSweepDeparture).
    for (uint64_t item = threadIdx.x; item < a.node_count * 8; item += kThreads) {
      const uint32_t node = static_cast<uint32_t>(item >> 3);
      const uint32_t d = static_cast<uint32_t>(item & 7u);
      uint64_t best = node == q.start_node ? 0 : lab[node * 9ull + d];
      for (uint32_t inc = 0; inc < 8; ++inc) {
        uint64_t cand = lab[node * 9ull + inc];
        if (cand == kInf64) continue;
        if (inc != d) {
          uint64_t b;
          if (!CheckedAdd(a.bend, q.bend_surcharge, &b) || !CheckedAdd(cand, b, &cand)) continue;
        }
        if (cand < best) {
          best = cand;
          if (inc != d) ++turns;
        }
      }
      dep[item] = best;
    }
    __syncthreads();
    // Segmented min-plus scans along every maximal run (SweepCandidateRuns).
    // Each (target, heading) label has exactly one incoming run edge, so one
    // thread owns each write.
    bool changed = false;
    for (uint64_t r = threadIdx.x; r < a.run_count; r += kThreads) {
      const apgx_v1_run run = a.runs[r];
      const uint32_t dir = run.direction;
      uint64_t carry = kInf64;
      for (uint32_t pos = 0; pos + 1 < run.node_count; ++pos) {
        const uint32_t source = a.run_nodes[run.node_offset + pos];
        const uint32_t target = a.run_nodes[run.node_offset + pos + 1];
        carry = min(carry, dep[source * 8ull + dir]);
        uint64_t w;
        if (!Transition(a, q, source, dir, dir, &w)) {
          carry = kInf64;
          continue;
        }
        uint64_t cand;
        if (CheckedAdd(carry, w, &cand)) {
          uint64_t& slot = lab[target * 9ull + dir];
          if (cand < slot) {
            slot = cand;
            changed = true;
          }
          carry = min(cand, dep[target * 8ull + dir]);
        } else {
          carry = dep[target * 8ull + dir];
        }
      }
    }
    changed = __syncthreads_or(changed);
    if (threadIdx.x == 0) {
      res.rounds += 1;
      res.examined_work += a.legal_edges;
    }
    if (!changed) {
      if (threadIdx.x == 0) {
        uint64_t best = kInf64;
        uint32_t best_state = kInvalid;
        for (uint32_t h = 0; h < 8; ++h) {
          const uint32_t st = q.goal_node * 9u + h;
          if (lab[st] < best || (lab[st] == best && st < best_state)) {
            best = lab[st];
            best_state = st;
          }
        }
        if (best == kInf64) {
          res.completion = APGX_V1_DISCONNECTED;
        } else {
          res.goal_state = best_state;
          res.completion = APGX_V1_REACHED;
        }
        res.reserved[0] = 1;  // finished marker (cleared before readback)
      }
      break;
    }
    __syncthreads();
  }
  if (turns) atomicAdd(reinterpret_cast<unsigned long long*>(&res.heading_turn_relaxations),
                       static_cast<unsigned long long>(turns));
}

// Bucketed-frontier generator (CandidateFrontierChunk,
// gpu/cuda/planar_router.cu:715-845): each round one workgroup takes the open
// state with the least (f, g, state) -- f = g + the admissible heuristic
// max(|dx|, |dy|) * min(orthogonal, diagonal step) of
// DeviceAdmissibleHeuristic (:590-608), winner order DeviceFrontierWinnerLess
// (:610-619) -- closes it, finishes at the goal node, and otherwise relaxes
// its legal edges (each counted as examined work).  The argmin is a
// workgroup reduction; one thread relaxes the <= 8 edges.  Every close and
// every label write first checks the slot's ownership stamp (the query's
// workspace_owner, written at init); a foreign stamp skips the write and
// zeroes the result's workspace_owner, so the host rejects the query instead
// of trusting contaminated state (DeviceClaimCandidateState, :620-637).
__device__ __forceinline__ uint64_t FrontierHeuristic(const V1Args& a, const apgx_v1_query& q, uint32_t node,
                                                     int64_t gx, int64_t gy) {
  const int64_t x = a.nodes[node].lattice_x, y = a.nodes[node].lattice_y;
  const uint64_t dx = x >= gx ? static_cast<uint64_t>(x) - static_cast<uint64_t>(gx)
                              : static_cast<uint64_t>(gx) - static_cast<uint64_t>(x);
  const uint64_t dy = y >= gy ? static_cast<uint64_t>(y) - static_cast<uint64_t>(gy)
                              : static_cast<uint64_t>(gy) - static_cast<uint64_t>(y);
  const uint64_t steps = dx > dy ? dx : dy;
  const uint64_t orth = static_cast<uint64_t>(a.orth) + q.orthogonal_step_surcharge;
  const uint64_t diag = static_cast<uint64_t>(a.diag) + q.diagonal_step_surcharge;
  const uint64_t step = orth < diag ? orth : diag;
  if (steps != 0 && step > (kInf64 - 1) / steps) return kInf64 - 1;
  return steps * step;
}

__device__ __forceinline__ bool WinnerLess(uint64_t f, uint64_t g, uint32_t s, uint64_t bf, uint64_t bg,
                                           uint32_t bs) {
  return f < bf || (f == bf && (g < bg || (g == bg && s < bs)));
}

__global__ __launch_bounds__(kThreads) void V1FrontierKernel(V1Args a, uint32_t chunk, uint32_t round_budget) {
  const uint32_t qi = blockIdx.x;
  apgx_v1_result& res = a.results[qi];
  const apgx_v1_query q = a.queries[qi];
  const uint64_t states = a.node_count * 9;
  uint64_t* lab = a.labels + static_cast<uint64_t>(qi) * states;
  uint8_t* closed = a.closed + static_cast<uint64_t>(qi) * states;
  const uint64_t* own = a.state_owners + static_cast<uint64_t>(qi) * states;
  const int64_t gx = a.nodes[q.goal_node].lattice_x, gy = a.nodes[q.goal_node].lattice_y;
  __shared__ uint64_t wf[kThreads], wg[kThreads];
  __shared__ uint32_t ws[kThreads];
  __shared__ uint32_t s_run;
  for (uint32_t k = 0; k < chunk; ++k) {
    if (threadIdx.x == 0) s_run = (res.reserved[0] == 0 && res.rounds < round_budget) ? 1u : 0u;
    __syncthreads();
    if (s_run == 0) break;
    uint64_t bf = kInf64, bg = kInf64;
    uint32_t bs = kInvalid;
    for (uint64_t st = threadIdx.x; st < states; st += kThreads) {
      if (closed[st]) continue;
      const uint64_t g = lab[st];
      if (g == kInf64) continue;
      uint64_t f;
      if (!CheckedAdd(g, FrontierHeuristic(a, q, static_cast<uint32_t>(st / 9), gx, gy), &f)) f = kInf64 - 1;
      if (WinnerLess(f, g, static_cast<uint32_t>(st), bf, bg, bs)) {
        bf = f;
        bg = g;
        bs = static_cast<uint32_t>(st);
      }
    }
    wf[threadIdx.x] = bf;
    wg[threadIdx.x] = bg;
    ws[threadIdx.x] = bs;
    __syncthreads();
    for (uint32_t stride = kThreads / 2; stride != 0; stride /= 2) {
      if (threadIdx.x < stride &&
          WinnerLess(wf[threadIdx.x + stride], wg[threadIdx.x + stride], ws[threadIdx.x + stride], wf[threadIdx.x],
                     wg[threadIdx.x], ws[threadIdx.x])) {
        wf[threadIdx.x] = wf[threadIdx.x + stride];
        wg[threadIdx.x] = wg[threadIdx.x + stride];
        ws[threadIdx.x] = ws[threadIdx.x + stride];
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) {
      res.rounds += 1;
      const uint32_t sel = ws[0];
      if (sel == kInvalid) {
        res.completion = APGX_V1_DISCONNECTED;
        res.reserved[0] = 1;
      } else if (own[sel] != q.workspace_owner) {
        res.workspace_owner = 0;
        res.completion = APGX_V1_DISCONNECTED;
        res.reserved[0] = 1;
      } else {
        closed[sel] = 1;
        const uint32_t node = sel / 9u, incoming = sel % 9u;
        if (node == q.goal_node) {
          res.goal_state = sel;
          res.completion = APGX_V1_REACHED;
          res.reserved[0] = 1;
        } else {
          const apgx_v1_node nd = a.nodes[node];
          for (uint32_t dir = 0; dir < 8; ++dir) {
            if ((nd.legal_edges & (1u << dir)) == 0) continue;
            res.examined_work += 1;
            uint64_t w, cand;
            if (!Transition(a, q, node, dir, incoming, &w)) continue;
            if (!CheckedAdd(wg[0], w, &cand)) continue;
            const uint64_t t = nd.neighbors[dir] * 9ull + dir;
            if (cand < lab[t]) {
              if (own[t] == q.workspace_owner) lab[t] = cand; else res.workspace_owner = 0;
            }
          }
        }
      }
    }
    __syncthreads();
  }
}

// Frontier readback: the lowest tight predecessor of every finite state from
// the final labels, one thread per (query, state) slot
// (SelectCandidateBatchPredecessors, gpu/cuda/planar_router.cu:1021-1082).
__global__ void V1SelectPredecessorsKernel(V1Args a, uint64_t total_states) {
  const uint64_t slot = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (slot >= total_states) return;
  const uint64_t states = a.node_count * 9;
  const uint32_t qi = static_cast<uint32_t>(slot / states);
  const uint64_t base = static_cast<uint64_t>(qi) * states;
  const uint32_t state = static_cast<uint32_t>(slot - base);
  const apgx_v1_query q = a.queries[qi];
  apgx_v1_result& res = a.results[qi];
  if (a.state_owners[slot] != q.workspace_owner || a.predecessor_owners[slot] != q.workspace_owner) {
    res.workspace_owner = 0;
    return;
  }
  const uint32_t node = state / 9u, inc = state % 9u;
  const uint32_t start_state = q.start_node * 9u + 8u;
  if (state == start_state || inc == 8u || a.labels[slot] == kInf64) {
    a.predecessors[slot] = kInvalid;
    return;
  }
  const uint32_t pred = a.nodes[node].neighbors[(inc + 4u) % 8u];
  if (pred == kInvalid) {
    a.predecessors[slot] = kInvalid;
    return;
  }
  uint32_t best = kInvalid;
  for (uint32_t ph = 0; ph < 9; ++ph) {
    if (ph == 8u && pred != q.start_node) continue;
    const uint32_t ps = pred * 9u + ph;
    if (a.state_owners[base + ps] != q.workspace_owner) {
      a.predecessors[slot] = kInvalid;
      res.workspace_owner = 0;
      return;
    }
    uint64_t w, c;
    if (!Transition(a, q, pred, inc, ph, &w)) continue;
    if (CheckedAdd(a.labels[base + ps], w, &c) && c == a.labels[slot] && ps < best) best = ps;
  }
  a.predecessors[slot] = best;
}

// Lowest predecessor state walk from the goal state (BuildQueryMajorCandidateSweepPaths).
__global__ void V1PathKernel(V1Args a) {
  const uint32_t qi = blockIdx.x * blockDim.x + threadIdx.x;
  if (qi >= a.query_count) return;
  const apgx_v1_query q = a.queries[qi];
  apgx_v1_result& res = a.results[qi];
  apgx_v1_compact_path& cp = a.paths[qi];
  cp.schema_version = 1;
  cp.state_count = 0;
  cp.query_id = q.query_id;
  cp.workspace_owner = q.workspace_owner;
  cp.state_offset = 0;
  if (res.completion != APGX_V1_REACHED) return;
  const uint64_t states = a.node_count * 9;
  const uint64_t* lab = a.labels + static_cast<uint64_t>(qi) * states;
  uint32_t* out = reinterpret_cast<uint32_t*>(a.departures + static_cast<uint64_t>(qi) * a.node_count * 8);
  const uint32_t start_state = q.start_node * 9u + 8u;
  uint32_t cursor = res.goal_state;
  for (uint64_t n = 0; n < states; ++n) {
    if (cursor >= states) break;
    out[n] = cursor;
    if (cursor == start_state) {
      cp.state_count = static_cast<uint32_t>(n + 1);
      return;
    }
    const uint32_t node = cursor / 9u, inc = cursor % 9u;
    if (inc == 8u) break;
    const uint32_t pred = a.nodes[node].neighbors[(inc + 4u) % 8u];
    if (pred == kInvalid) break;
    uint32_t best = kInvalid;
    for (uint32_t ph = 0; ph < 9; ++ph) {
      if (ph == 8u && pred != q.start_node) continue;
      const uint32_t ps = pred * 9u + ph;
      uint64_t w, c;
      if (!Transition(a, q, pred, inc, ph, &w)) continue;
      if (CheckedAdd(lab[ps], w, &c) && c == lab[cursor] && ps < best) best = ps;
    }
    if (best == kInvalid) break;
    cursor = best;
  }
  cp.state_count = 0;
  res.workspace_owner = 0;  // an incomplete chain invalidates the query's association
}

__global__ void V1ScanKernel(V1Args a) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  uint64_t offset = 0;
  for (uint32_t i = 0; i < a.query_count; ++i) {
    a.paths[i].state_offset = offset;
    offset += a.paths[i].state_count;
  }
}

__global__ void V1ScatterKernel(V1Args a) {
  const uint32_t qi = blockIdx.x;
  const apgx_v1_compact_path cp = a.paths[qi];
  const uint32_t* src = reinterpret_cast<const uint32_t*>(a.departures + static_cast<uint64_t>(qi) * a.node_count * 8);
  for (uint32_t i = threadIdx.x; i < cp.state_count; i += blockDim.x) a.packed[cp.state_offset + i] = src[i];
}

}  // namespace
}  // namespace v1
}  // namespace apgx


struct apgx_v1_pending {
  uint32_t query_count = 0;
  uint32_t readback_kind = APGX_V1_READBACK_COMPACT_PATHS;
  uint64_t states_per_query = 0;
  std::vector<apgx_v1_result> results;
  std::vector<apgx_v1_query_envelope> envelopes;  // workspace slice + query telemetry beside each header
  std::vector<apgx_v1_compact_path> paths;
  std::vector<uint32_t> states;
  // Full-workspace readback (bucketed frontier), query-major.
  std::vector<uint64_t> labels, state_owners, predecessor_owners;
  std::vector<uint32_t> predecessors;
  apgx_v1_batch_telemetry telemetry{};
};

namespace apgx {
namespace v1 {
namespace {

int ValidateView(const apgx_v1_header* h, const apgx_v1_layer_range* layers, uint64_t nl,
                 const apgx_v1_node* nodes, uint64_t nn, const apgx_v1_run* runs, uint64_t nr,
                 const uint32_t* run_nodes, uint64_t nrn, uint64_t* legal_edges) {
  auto bad = [](const char* why) { return Fail(APGX_E_INTERNAL_INVARIANT, std::string("V1 view rejected: ") + why); };
  if (h->schema_version != APGX_V1_SCHEMA_VERSION) return Fail(APGX_E_UNSUPPORTED, "V1 view schema version is not supported");
  if (nn == 0 || nn > kMaximumRepresentedNodes || h->represented_nodes != nn || h->represented_states != nn * 9)
    return bad("node counts disagree with the header");
  if ((h->heading_mask & ~7u) != 0 || h->heading_mask == 0) return bad("heading mask");
  uint64_t covered = 0;
  for (uint64_t i = 0; i < nl; ++i) {
    const apgx_v1_layer_range& r = layers[i];
    if (r.node_offset != covered || r.node_count == 0 || covered + r.node_count > nn) return bad("layer ranges");
    if (i > 0 && layers[i - 1].layer >= r.layer) return bad("layer order");
    int64_t mnx = nodes[r.node_offset].lattice_x, mxx = mnx, mny = nodes[r.node_offset].lattice_y, mxy = mny;
    for (uint64_t k = r.node_offset; k < r.node_offset + r.node_count; ++k) {
      if (nodes[k].layer != r.layer) return bad("layer membership");
      mnx = std::min(mnx, nodes[k].lattice_x); mxx = std::max(mxx, nodes[k].lattice_x);
      mny = std::min(mny, nodes[k].lattice_y); mxy = std::max(mxy, nodes[k].lattice_y);
    }
    if (mnx != r.minimum_lattice_x || mxx != r.maximum_lattice_x || mny != r.minimum_lattice_y || mxy != r.maximum_lattice_y)
      return bad("layer bounds");
    covered += r.node_count;
  }
  if (covered != nn) return bad("layer ranges do not cover every node");
  uint64_t legal = 0;
  for (uint64_t i = 0; i < nn; ++i) {
    const apgx_v1_node& d = nodes[i];
    for (int dir = 0; dir < 8; ++dir) {
      const bool l = (d.legal_edges >> dir) & 1u;
      const uint32_t nb = d.neighbors[dir];
      if (l != (nb != APGX_V1_INVALID_INDEX)) return bad("legal mask and neighbour disagree");
      if (!l) continue;
      if ((h->heading_mask & HeadingFor(dir)) == 0) return bad("legal edge outside the heading mask");
      if (nb >= nn) return bad("neighbour index out of range");
      const apgx_v1_node& t = nodes[nb];
      if (t.layer != d.layer || t.lattice_x != d.lattice_x + kDx[dir] || t.lattice_y != d.lattice_y + kDy[dir])
        return bad("neighbour is not the adjacent lattice node");
      if (t.neighbors[(dir + 4) % 8] != i) return bad("edge lacks its reverse");
      ++legal;
    }
  }
  uint64_t run_edges = 0, prev_dir = 0, prev_start = 0;
  for (uint64_t r = 0; r < nr; ++r) {
    const apgx_v1_run& run = runs[r];
    if (run.direction >= 8 || run.node_count < 2 || static_cast<uint64_t>(run.node_offset) + run.node_count > nrn)
      return bad("run bounds");
    const uint32_t first = run_nodes[run.node_offset];
    if (first >= nn) return bad("run node out of range");
    if (r > 0 && (run.direction < prev_dir || (run.direction == prev_dir && first <= prev_start)))
      return bad("runs are not in canonical (direction, start) order");
    prev_dir = run.direction;
    prev_start = first;
    const uint32_t pred = nodes[first].neighbors[(run.direction + 4) % 8];
    if (pred != APGX_V1_INVALID_INDEX && ((nodes[pred].legal_edges >> run.direction) & 1u))
      return bad("run is not maximal at its start");
    for (uint32_t k = 0; k + 1 < run.node_count; ++k) {
      const uint32_t s = run_nodes[run.node_offset + k], t = run_nodes[run.node_offset + k + 1];
      if (s >= nn || t >= nn || nodes[s].neighbors[run.direction] != t) return bad("run step is not a legal edge");
    }
    const uint32_t last = run_nodes[run.node_offset + run.node_count - 1];
    if ((nodes[last].legal_edges >> run.direction) & 1u) return bad("run is not maximal at its end");
    run_edges += run.node_count - 1;
  }
  if (run_edges != legal) return bad("runs do not cover every legal directed edge exactly once");
  if (h->estimated_persistent_device_bytes != PersistentBytes(nl, nn, nr, nrn)) return bad("persistent byte formula");
  if (Fingerprint(*h, layers, nl, nodes, nn, runs, nr, run_nodes, nrn) != h->device_view_fingerprint)
    return bad("device view fingerprint mismatch");
  *legal_edges = legal;
  return APGX_OK;
}

}  // namespace
}  // namespace v1
}  // namespace apgx

extern "C" int apgx_v1_view_upload(const apgx_v1_header* header, const apgx_v1_layer_range* layers,
                                   uint64_t nl, const apgx_v1_node* nodes, uint64_t nn,
                                   const apgx_v1_run* runs, uint64_t nr, const uint32_t* run_nodes,
                                   uint64_t nrn, apgx_v1_view** out) {
  using namespace apgx;
  if (out == nullptr || header == nullptr || (nl && !layers) || (nn && !nodes) || (nr && !runs) || (nrn && !run_nodes))
    return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_v1_view_upload: null argument");
  *out = nullptr;
  uint64_t legal = 0;
  if (int s = v1::ValidateView(header, layers, nl, nodes, nn, runs, nr, run_nodes, nrn, &legal)) return s;
  auto* v = new (std::nothrow) apgx_v1_view();
  if (v == nullptr) return Fail(APGX_E_RESOURCE_EXHAUSTED, "apgx_v1_view_upload: host allocation");
  v->header = *header;
  v->layer_count = nl; v->node_count = nn; v->run_count = nr; v->run_node_count = nrn; v->legal_edges = legal;
  v->persistent_bytes = header->estimated_persistent_device_bytes;
  hipError_t e = hipGetDevice(&v->device);
  if (e == hipSuccess) e = hipMalloc(&v->d_nodes, nn * sizeof(apgx_v1_node));
  if (e == hipSuccess) e = hipMalloc(&v->d_runs, std::max<uint64_t>(nr, 1) * sizeof(apgx_v1_run));
  if (e == hipSuccess) e = hipMalloc(&v->d_run_nodes, std::max<uint64_t>(nrn, 1) * 4);
  if (e == hipSuccess) e = hipMemcpy(v->d_nodes, nodes, nn * sizeof(apgx_v1_node), hipMemcpyHostToDevice);
  if (e == hipSuccess && nr) e = hipMemcpy(v->d_runs, runs, nr * sizeof(apgx_v1_run), hipMemcpyHostToDevice);
  if (e == hipSuccess && nrn) e = hipMemcpy(v->d_run_nodes, run_nodes, nrn * 4, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    apgx_v1_view_release(v);
    return HipFail(e, "V1 view upload");
  }
  *out = v;
  ClearError();
  return APGX_OK;
}

extern "C" void apgx_v1_view_release(apgx_v1_view* v) {
  if (v == nullptr) return;
  if (v->cancel_word) (void)hipHostFree(v->cancel_word);
  if (v->d_nodes) (void)hipFree(v->d_nodes);
  if (v->d_runs) (void)hipFree(v->d_runs);
  if (v->d_run_nodes) (void)hipFree(v->d_run_nodes);
  delete v;
}

extern "C" int apgx_v1_batch_execute(apgx_v1_view* view, const apgx_v1_batch_request* req,
                                     const apgx_v1_query* queries, uint32_t qn,
                                     const apgx_v1_policy_edge* policy, uint64_t pn,
                                     const volatile uint32_t* cancel, apgx_v1_pending** out) {
  using namespace apgx;
  using namespace apgx::v1;
  const auto wall0 = std::chrono::steady_clock::now();
  if (view == nullptr || req == nullptr || out == nullptr || (qn && !queries) || (pn && !policy))
    return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_v1_batch_execute: null argument");
  *out = nullptr;
  int device = 0;
  APGX_HIP(hipGetDevice(&device), "hipGetDevice");
  if (device != view->device)
    return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_v1_batch_execute: current device differs from the upload device");
  if (req->schema_version != APGX_V1_SCHEMA_VERSION)
    return Fail(APGX_E_INTERNAL_INVARIANT, "candidate batch schema version mismatch");
  if (req->generator != APGX_V1_GENERATOR_HEADING_AWARE_SWEEP && req->generator != APGX_V1_GENERATOR_BUCKETED_FRONTIER)
    return Fail(APGX_E_UNSUPPORTED, "unknown candidate generator");
  const bool frontier = req->generator == APGX_V1_GENERATOR_BUCKETED_FRONTIER;
  if (req->maximum_rounds == 0) return Fail(APGX_E_INTERNAL_INVARIANT, "candidate batch round budget must be positive");
  const uint64_t nn = view->node_count, S = nn * 9;
  if (S > req->maximum_workspace_states_per_query)
    return Fail(APGX_E_RESOURCE_EXHAUSTED, "query workspace exceeds maximum_workspace_states_per_query");
  if (frontier && S > req->maximum_frontier_states_per_query)
    return Fail(APGX_E_RESOURCE_EXHAUSTED, "query frontier exceeds maximum_frontier_states_per_query");
  // Request shape (ValidateCandidateBatchExecution, gpu/cuda/planar_router.cu:1731-1909).
  for (uint32_t i = 0; i < qn; ++i) {
    const apgx_v1_query& q = queries[i];
    if (q.schema_version != APGX_V1_SCHEMA_VERSION || q.batch_id != req->batch_id || q.query_id == 0 ||
        (i > 0 && queries[i - 1].query_id >= q.query_id) || q.workspace_owner != q.query_id ||
        q.start_node >= nn || q.goal_node >= nn || q.start_node == q.goal_node ||
        q.workspace_offset != static_cast<uint64_t>(i) * S || q.workspace_state_count != S ||
        q.frontier_state_capacity != (frontier ? S : 0) || q.generator != req->generator ||
        static_cast<uint64_t>(q.policy_offset) + q.policy_count > pn)
      return Fail(APGX_E_INTERNAL_INVARIANT, "malformed candidate batch query record");
    for (uint32_t k = 1; k < q.policy_count; ++k)
      if (policy[q.policy_offset + k - 1].physical_edge_index >= policy[q.policy_offset + k].physical_edge_index)
        return Fail(APGX_E_INTERNAL_INVARIANT, "query policy slice is not strictly sorted by physical edge");
  }
  const uint64_t total_states = static_cast<uint64_t>(qn) * S;
  const BatchLayout lay = LayoutBatch(frontier, qn, pn, nn);
  const uint64_t batch_bytes = lay.total;
  if (batch_bytes + view->persistent_bytes > req->maximum_device_bytes)
    return Fail(APGX_E_RESOURCE_EXHAUSTED, "candidate batch exceeds maximum_device_bytes");
  // Host budget before execution, with every compact path at its largest
  // (EstimateCandidateBatchHostBytesV1, src/gpu/planar_router.cc:738-794).
  if (HostBytes(frontier, qn, nn, total_states) > req->maximum_host_bytes)
    return Fail(APGX_E_RESOURCE_EXHAUSTED, "candidate batch exceeds maximum_host_bytes");
  auto* pend = new (std::nothrow) apgx_v1_pending();
  if (pend == nullptr) return Fail(APGX_E_RESOURCE_EXHAUSTED, "apgx_v1_batch_execute: host allocation");
  pend->query_count = qn;
  pend->states_per_query = S;
  pend->readback_kind = frontier ? APGX_V1_READBACK_FULL_WORKSPACE : APGX_V1_READBACK_COMPACT_PATHS;
  pend->results.resize(qn);
  pend->paths.resize(qn);
  for (uint32_t i = 0; i < qn; ++i) {
    apgx_v1_result& r = pend->results[i];
    std::memset(&r, 0, sizeof(r));
    const apgx_v1_query& q = queries[i];
    r.schema_version = APGX_V1_SCHEMA_VERSION;
    r.compiler_version = view->header.compiler_version;
    r.input_ordinal = q.input_ordinal;
    r.start_node = q.start_node;
    r.goal_node = q.goal_node;
    r.goal_state = APGX_V1_INVALID_INDEX;
    r.generator = q.generator;
    r.completion = APGX_V1_BUDGET_EXHAUSTED;
    r.batch_id = q.batch_id;
    r.query_id = q.query_id;
    r.workspace_owner = q.workspace_owner;
    r.policy_identity = q.policy_identity;
    r.routing_profile_fingerprint = q.routing_profile_fingerprint;
    r.source_board_content_hash = view->header.source_board_content_hash;
    r.compiler_profile_fingerprint = view->header.compiler_profile_fingerprint;
    r.rule_bucket_identity = view->header.rule_bucket_identity;
    r.device_view_fingerprint = view->header.device_view_fingerprint;
  }
  apgx_v1_batch_telemetry& t = pend->telemetry;
  t.persistent_device_bytes = view->persistent_bytes;
  t.batch_device_bytes = batch_bytes;
  t.workspace_capacity_device_bytes = batch_bytes;
  t.peak_device_bytes = batch_bytes + view->persistent_bytes;
  GeneratorShape shape{};
  DescribeGenerator(req->generator, &shape);
  const uint32_t chunk_rounds = shape.chunk_rounds;
  t.chunk_rounds = chunk_rounds;
  if (qn == 0) {
    *out = pend;
    ClearError();
    return APGX_OK;
  }
  V1Args a{};
  a.nodes = view->d_nodes; a.runs = view->d_runs; a.run_nodes = view->d_run_nodes;
  a.node_count = nn; a.run_count = view->run_count; a.legal_edges = view->legal_edges;
  a.query_count = qn;
  a.orth = view->header.orthogonal_step; a.diag = view->header.diagonal_step; a.bend = view->header.bend;
  void* block = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int status = APGX_OK;
  auto check = [&](hipError_t err, const char* op) {
    if (status == APGX_OK && err != hipSuccess) status = HipFail(err, op);
    return status == APGX_OK;
  };
  check(hipMalloc(&block, lay.total), "candidate batch workspace allocation");
  if (status != APGX_OK) { delete pend; return status; }
  char* base = static_cast<char*>(block);
  a.queries = reinterpret_cast<apgx_v1_query*>(base + lay.queries);
  a.policy = reinterpret_cast<apgx_v1_policy_edge*>(base + lay.policy);
  a.results = reinterpret_cast<apgx_v1_result*>(base + lay.results);
  a.labels = reinterpret_cast<uint64_t*>(base + lay.labels);
  if (frontier) {
    a.predecessors = reinterpret_cast<uint32_t*>(base + lay.predecessors);
    a.state_owners = reinterpret_cast<uint64_t*>(base + lay.state_owners);
    a.predecessor_owners = reinterpret_cast<uint64_t*>(base + lay.predecessor_owners);
    a.closed = reinterpret_cast<uint8_t*>(base + lay.closed);
  } else {
    a.paths = reinterpret_cast<apgx_v1_compact_path*>(base + lay.paths);
    a.packed = reinterpret_cast<uint32_t*>(base + lay.packed);
    a.departures = reinterpret_cast<uint64_t*>(base + lay.departures);
  }
  check(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking), "hipStreamCreate");
  check(hipEventCreate(&e0), "hipEventCreate");
  check(hipEventCreate(&e1), "hipEventCreate");
  check(hipMemcpyAsync(const_cast<apgx_v1_query*>(a.queries), queries, qn * sizeof(apgx_v1_query), hipMemcpyHostToDevice, stream), "query upload");
  if (pn) check(hipMemcpyAsync(const_cast<apgx_v1_policy_edge*>(a.policy), policy, pn * 16, hipMemcpyHostToDevice, stream), "policy upload");
  check(hipMemcpyAsync(a.results, pend->results.data(), qn * sizeof(apgx_v1_result), hipMemcpyHostToDevice, stream), "result init");
  uint64_t launches = 0, syncs = 0;
  uint32_t dispatched = 0, finalization = 0;
  bool cancelled = false, running = true;
  if (status == APGX_OK) {
    check(hipEventRecord(e0, stream), "hipEventRecord");
    const uint32_t state_blocks = static_cast<uint32_t>((total_states + 255) / 256);
    hipLaunchKernelGGL(V1InitKernel, dim3(state_blocks), dim3(256), 0, stream, a, total_states);
    check(hipGetLastError(), "V1 init launch");
    ++launches;
    std::vector<apgx_v1_result> snapshot(qn);
    while (status == APGX_OK && dispatched < req->maximum_rounds) {
      if (cancel != nullptr && *cancel != 0) { cancelled = true; break; }
      const uint32_t chunk = std::min(chunk_rounds, req->maximum_rounds - dispatched);
      if (frontier)
        hipLaunchKernelGGL(V1FrontierKernel, dim3(qn), dim3(kThreads), 0, stream, a, chunk, req->maximum_rounds);
      else
        hipLaunchKernelGGL(V1SweepKernel, dim3(qn), dim3(kThreads), 0, stream, a, chunk, req->maximum_rounds);
      if (!check(hipGetLastError(), "V1 generator launch")) break;
      ++launches;
      dispatched += chunk;
      check(hipMemcpyAsync(snapshot.data(), a.results, qn * sizeof(apgx_v1_result), hipMemcpyDeviceToHost, stream), "status readback");
      check(hipStreamSynchronize(stream), "hipStreamSynchronize");
      ++syncs;
      running = false;
      for (const apgx_v1_result& r : snapshot) running |= r.reserved[0] == 0;
      if (!running) break;
    }
    if (status == APGX_OK && running) {
      const uint32_t completion = cancelled ? APGX_V1_CANCELLED : APGX_V1_BUDGET_EXHAUSTED;
      hipLaunchKernelGGL(V1FinalizeKernel, dim3((qn + 63) / 64), dim3(64), 0, stream, a, completion);
      check(hipGetLastError(), "V1 finalization launch");
      ++launches;
      finalization = 1;
    }
    if (status == APGX_OK) {
      if (frontier) {
        hipLaunchKernelGGL(V1SelectPredecessorsKernel, dim3(state_blocks), dim3(256), 0, stream, a, total_states);
        check(hipGetLastError(), "V1 predecessor launch");
        launches += 1;
      } else {
        hipLaunchKernelGGL(V1PathKernel, dim3((qn + 63) / 64), dim3(64), 0, stream, a);
        hipLaunchKernelGGL(V1ScanKernel, dim3(1), dim3(1), 0, stream, a);
        hipLaunchKernelGGL(V1ScatterKernel, dim3(qn), dim3(256), 0, stream, a);
        check(hipGetLastError(), "V1 path launch");
        launches += 3;
      }
      check(hipEventRecord(e1, stream), "hipEventRecord");
      check(hipStreamSynchronize(stream), "hipStreamSynchronize");
    }
  }
  float ms = 0.f;
  if (status == APGX_OK) check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
  if (status == APGX_OK)
    check(hipMemcpy(pend->results.data(), a.results, qn * sizeof(apgx_v1_result), hipMemcpyDeviceToHost), "result readback");
  uint64_t used = 0;
  if (status == APGX_OK && frontier) {
    pend->labels.resize(total_states);
    pend->predecessors.resize(total_states);
    pend->state_owners.resize(total_states);
    pend->predecessor_owners.resize(total_states);
    check(hipMemcpy(pend->labels.data(), a.labels, total_states * 8, hipMemcpyDeviceToHost), "label readback");
    check(hipMemcpy(pend->predecessors.data(), a.predecessors, total_states * 4, hipMemcpyDeviceToHost), "predecessor readback");
    check(hipMemcpy(pend->state_owners.data(), a.state_owners, total_states * 8, hipMemcpyDeviceToHost), "owner readback");
    check(hipMemcpy(pend->predecessor_owners.data(), a.predecessor_owners, total_states * 8, hipMemcpyDeviceToHost), "owner readback");
  } else if (status == APGX_OK) {
    check(hipMemcpy(pend->paths.data(), a.paths, qn * sizeof(apgx_v1_compact_path), hipMemcpyDeviceToHost), "path header readback");
    for (const auto& p : pend->paths) used += p.state_count;
    pend->states.resize(used);
    if (used && status == APGX_OK) check(hipMemcpy(pend->states.data(), a.packed, used * 4, hipMemcpyDeviceToHost), "path readback");
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (stream) (void)hipStreamDestroy(stream);
  (void)hipFree(block);
  if (status != APGX_OK) { delete pend; return status; }
  pend->envelopes.resize(qn);
  for (uint32_t i = 0; i < qn; ++i) {
    apgx_v1_result& r = pend->results[i];
    r.reserved[0] = 0;  // device status word
    pend->envelopes[i] = apgx_v1_query_envelope{static_cast<uint64_t>(i) * S, S, r.examined_work,
                                                r.heading_turn_relaxations, r.rounds, 0};
  }
  t.kernel_launch_count = launches;
  t.blocking_status_readback_count = syncs;
  t.dispatched_rounds = dispatched;
  t.finalization_launch_count = finalization;
  t.kernel_milliseconds = ms;
  t.device_to_host_readback_bytes = ReadbackBytes(frontier, qn, nn, used);
  t.batch_host_bytes = HostBytes(frontier, qn, nn, used);
  (void)wall0;
  *out = pend;
  ClearError();
  return APGX_OK;
}

extern "C" int apgx_v1_batch_readback(apgx_v1_pending* p, apgx_v1_result* results,
                                      apgx_v1_compact_path* paths, uint32_t* states,
                                      uint64_t cap, uint64_t* used, apgx_v1_batch_telemetry* tel) {
  using namespace apgx;
  if (p == nullptr) return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_v1_batch_readback: null pending");
  if (used) *used = p->states.size();
  if (p->states.size() > cap && states != nullptr)
    return Fail(APGX_E_RESOURCE_EXHAUSTED, "apgx_v1_batch_readback: state capacity too small");
  if (results) std::memcpy(results, p->results.data(), p->results.size() * sizeof(apgx_v1_result));
  if (paths) std::memcpy(paths, p->paths.data(), p->paths.size() * sizeof(apgx_v1_compact_path));
  if (states && !p->states.empty()) std::memcpy(states, p->states.data(), p->states.size() * 4);
  if (tel) *tel = p->telemetry;
  ClearError();
  return APGX_OK;
}

extern "C" int apgx_v1_batch_readback_workspace(apgx_v1_pending* p, uint32_t* readback_kind,
                                                apgx_v1_query_envelope* per_query, uint64_t* labels,
                                                uint32_t* predecessors, uint64_t* state_owners,
                                                uint64_t* predecessor_owners, uint64_t capacity,
                                                uint64_t* total) {
  using namespace apgx;
  if (p == nullptr) return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_v1_batch_readback_workspace: null pending");
  const uint64_t n = p->labels.size();
  if (readback_kind) *readback_kind = p->readback_kind;
  if (total) *total = n;
  if (per_query && !p->envelopes.empty())
    std::memcpy(per_query, p->envelopes.data(), p->envelopes.size() * sizeof(apgx_v1_query_envelope));
  const bool want = labels || predecessors || state_owners || predecessor_owners;
  if (want && n > capacity)
    return Fail(APGX_E_RESOURCE_EXHAUSTED, "apgx_v1_batch_readback_workspace: state capacity too small");
  if (n) {
    if (labels) std::memcpy(labels, p->labels.data(), n * 8);
    if (predecessors) std::memcpy(predecessors, p->predecessors.data(), n * 4);
    if (state_owners) std::memcpy(state_owners, p->state_owners.data(), n * 8);
    if (predecessor_owners) std::memcpy(predecessor_owners, p->predecessor_owners.data(), n * 8);
  }
  ClearError();
  return APGX_OK;
}

// Replay/test-only readback decorator (FaultInjectingCandidateBatchBackend,
// src/gpu/fault_injecting_backend.cc:113-224): corrupts the first query of an
// executed batch's host copy, after the normal readback, so the trust boundary
// can be shown to catch it.  The execution path itself has no fault controls.
extern "C" int apgx_v1_pending_inject_fault(apgx_v1_pending* p, uint32_t fault) {
  using namespace apgx;
  if (p == nullptr) return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_v1_pending_inject_fault: null pending");
  if (fault > APGX_V1_FAULT_COMPACT_PATH_EDGE) return Fail(APGX_E_UNSUPPORTED, "unknown readback fault");
  if (fault == APGX_V1_FAULT_NONE || p->query_count == 0) {
    ClearError();
    return APGX_OK;
  }
  apgx_v1_result& r = p->results[0];
  apgx_v1_compact_path& cp = p->paths[0];
  const bool full = p->readback_kind == APGX_V1_READBACK_FULL_WORKSPACE;
  const uint64_t S = p->states_per_query;
  constexpr uint64_t kMax64 = ~0ull;
  constexpr uint32_t kInvalidState = APGX_V1_INVALID_INDEX;
  // Compact path slice of query 0 (empty when out of range, as the decorator's).
  uint32_t* cs = nullptr;
  uint64_t cn = 0;
  if (!full && cp.state_offset <= p->states.size() && cp.state_count <= p->states.size() - cp.state_offset) {
    cs = p->states.data() + cp.state_offset;
    cn = cp.state_count;
  }
  auto contains = [&](uint32_t v) { return std::find(cs, cs + cn, v) != cs + cn; };
  switch (fault) {
    case APGX_V1_FAULT_WORKSPACE_BOUNDS: {
      uint64_t& count = p->envelopes[0].workspace_state_count;
      count = count == 0 ? 1 : count - 1;
      break;
    }
    case APGX_V1_FAULT_WORKSPACE_OWNER:
      if (full && S) p->state_owners[0] = r.workspace_owner == kMax64 ? 0 : r.workspace_owner + 1;
      break;
    case APGX_V1_FAULT_QUERY_TELEMETRY: {
      const uint32_t rounds = p->envelopes[0].rounds;
      r.rounds = rounds == 0xFFFFFFFFu ? rounds - 1 : rounds + 1;
      break;
    }
    case APGX_V1_FAULT_MEMORY_ACCOUNTING:
      p->telemetry.batch_host_bytes += p->telemetry.batch_host_bytes == kMax64 ? ~0ull : 1;
      break;
    case APGX_V1_FAULT_BATCH_TELEMETRY:
      p->telemetry.kernel_launch_count += p->telemetry.kernel_launch_count == kMax64 ? ~0ull : 1;
      break;
    case APGX_V1_FAULT_QUERY_IDENTITY:
      r.query_id = r.query_id == kMax64 ? 0 : r.query_id + 1;
      break;
    case APGX_V1_FAULT_FALSE_DISCONNECTED:
      r.completion = APGX_V1_DISCONNECTED;
      r.goal_state = kInvalidState;
      if (full) {
        std::fill(p->labels.begin(), p->labels.begin() + S, kMax64);
        std::fill(p->predecessors.begin(), p->predecessors.begin() + S, kInvalidState);
        const uint64_t start = static_cast<uint64_t>(r.start_node) * 9 + 8;
        if (start < S) p->labels[start] = 0;
      }
      break;
    case APGX_V1_FAULT_COMPACT_PATH_BOUNDS:
      if (!full) cp.schema_version = cp.schema_version == 0xFFFFFFFFu ? cp.schema_version - 1 : cp.schema_version + 1;
      break;
    case APGX_V1_FAULT_COMPACT_PATH_ENDPOINT:
      if (cn) cs[0] = kInvalidState;
      break;
    case APGX_V1_FAULT_COMPACT_PATH_CYCLE:
      if (cn >= 4) cs[1] = cs[2];
      break;
    case APGX_V1_FAULT_COMPACT_PATH_HEADING:
      for (uint64_t i = 1; i + 1 < cn; ++i) {
        const uint32_t node = cs[i] / 9u, rep = node * 9u + 8u;
        if (node != r.start_node && !contains(rep)) {
          cs[i] = rep;
          break;
        }
      }
      break;
    case APGX_V1_FAULT_COMPACT_PATH_EDGE: {
      bool done = false;
      for (uint64_t i = 1; i + 1 < cn && !done; ++i) {
        const uint32_t node = cs[i] / 9u, heading = cs[i] % 9u;
        for (uint32_t h = 0; h < 8; ++h) {
          const uint32_t rep = node * 9u + h;
          if (h != heading && !contains(rep)) {
            cs[i] = rep;
            done = true;
            break;
          }
        }
      }
      break;
    }
    default:
      break;
  }
  ClearError();
  return APGX_OK;
}

extern "C" void apgx_v1_pending_release(apgx_v1_pending* p) { delete p; }
