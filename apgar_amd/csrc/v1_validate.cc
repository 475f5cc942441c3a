// v1_validate.cc — host trust boundary for V1 candidate batches.
//
// Re-checks an untrusted readback (compact paths from the sweep, the full
// workspace from the frontier) against the view, the request and the queries
// before any result is used, as the reference's host router does
// (src/gpu/planar_router.cc:2195-2380, with the checks it calls).  A failure
// carries the reference's failure class and, where the reference defines one,
// its invariant identifier, so callers and replays can assert on them.
// Nothing here produces a route: the validator only accepts or rejects what
// the device returned.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <queue>
#include <string>
#include <system_error>
#include <thread>
#include <vector>

#include "apgx.h"
#include "common.h"
#include "v1_common.h"

namespace apgx {
namespace v1check {
namespace {

using namespace ::apgx::v1;

constexpr uint64_t kInf = ~0ull;
constexpr uint32_t kNone = APGX_V1_INVALID_INDEX;

bool CheckedAdd(uint64_t a, uint64_t b, uint64_t* out) {
  if (a > kInf - 1 || b > kInf - 1 - a) return false;
  *out = a + b;
  return true;
}

struct Context {
  const apgx_v1_header* header;
  const apgx_v1_node* nodes;
  uint64_t node_count;
  const apgx_v1_policy_edge* policy;
};

uint64_t Adjustment(const Context& c, const apgx_v1_query& q, uint32_t edge) {
  const apgx_v1_policy_edge* first = c.policy + q.policy_offset;
  const apgx_v1_policy_edge* last = first + q.policy_count;
  const apgx_v1_policy_edge* it = std::lower_bound(
      first, last, edge, [](const apgx_v1_policy_edge& e, uint32_t v) { return e.physical_edge_index < v; });
  return (it != last && it->physical_edge_index == edge) ? it->adjustment : 0;
}

// Canonical physical edge of a move from `source` in `dir` (the lower endpoint
// in E, NE, N, NW order), as CanonicalPhysicalEdgeResource numbers it.
bool CanonicalEdge(const Context& c, uint32_t source, uint32_t dir, uint32_t* edge) {
  if (dir < 4) {
    *edge = source * 8u + dir;
    return true;
  }
  const uint32_t canonical = c.nodes[source].neighbors[dir];
  if (canonical == kNone) return false;
  *edge = canonical * 8u + dir - 4u;
  return true;
}

// Step cost under the query's policy; false when banned, unaddressable or
// overflowing (StepCostUnderPolicy, src/routing/candidate_policy.cc:419-438).
bool Transition(const Context& c, const apgx_v1_query& q, uint32_t source, uint32_t dir, uint32_t incoming,
                uint64_t* cost, bool* banned) {
  *banned = false;
  uint32_t edge;
  if (!CanonicalEdge(c, source, dir, &edge)) return false;
  const uint64_t adj = Adjustment(c, q, edge);
  if (adj == APGX_V1_BANNED_ADJUSTMENT) {
    *banned = true;
    return false;
  }
  uint64_t v;
  if (!CheckedAdd((dir & 1u) ? c.header->diagonal_step : c.header->orthogonal_step,
                  (dir & 1u) ? q.diagonal_step_surcharge : q.orthogonal_step_surcharge, &v))
    return false;
  if (incoming != 8u && incoming != dir) {
    uint64_t b;
    if (!CheckedAdd(c.header->bend, q.bend_surcharge, &b) || !CheckedAdd(v, b, &v)) return false;
  }
  return CheckedAdd(v, adj, cost);
}

// Exact CPU search over heading states: is the goal node reachable at a
// finite cost?  (The reference confirms every "disconnected" with its CPU
// oracle before trusting it.)
bool GoalReachable(const Context& c, const apgx_v1_query& q) {
  const uint64_t S = c.node_count * 9;
  std::vector<uint64_t> dist(S, kInf);
  using Item = std::pair<uint64_t, uint32_t>;
  std::priority_queue<Item, std::vector<Item>, std::greater<Item>> heap;
  const uint32_t start = q.start_node * 9u + 8u;
  dist[start] = 0;
  heap.push({0, start});
  while (!heap.empty()) {
    const auto [d, s] = heap.top();
    heap.pop();
    if (d != dist[s]) continue;
    const uint32_t node = s / 9u, inc = s % 9u;
    if (node == q.goal_node) return true;
    const apgx_v1_node& nd = c.nodes[node];
    for (uint32_t dir = 0; dir < 8; ++dir) {
      if ((nd.legal_edges & (1u << dir)) == 0 || nd.neighbors[dir] == kNone) continue;
      uint64_t w, nc;
      bool banned;
      if (!Transition(c, q, node, dir, inc, &w, &banned) || !CheckedAdd(d, w, &nc)) continue;
      const uint32_t t = nd.neighbors[dir] * 9u + dir;
      if (nc < dist[t]) {
        dist[t] = nc;
        heap.push({nc, t});
      }
    }
  }
  return false;
}


// One adjacent, legal, same-layer compiled edge from `pnode` to `node` in
// direction `dir`.
bool AdjacentLegal(const Context& c, uint32_t pnode, uint32_t node, uint32_t dir) {
  static const int kDx8[8] = {1, 1, 0, -1, -1, -1, 0, 1}, kDy8[8] = {0, 1, 1, 1, 0, -1, -1, -1};
  const apgx_v1_node& pn = c.nodes[pnode];
  const int64_t dx = c.nodes[node].lattice_x - pn.lattice_x, dy = c.nodes[node].lattice_y - pn.lattice_y;
  return pn.neighbors[dir] == node && (pn.legal_edges & (1u << dir)) != 0 && dx == kDx8[dir] && dy == kDy8[dir] &&
         pn.layer == c.nodes[node].layer;
}

// ValidateAllPredecessors (src/gpu/planar_router.cc:491-612) over one query's
// labels and predecessors: exact (label = predecessor label + step) for
// converged results, a lower bound (and missing predecessors allowed) for
// unfinished ones.
bool AllPredecessorsValid(const Context& c, const apgx_v1_query& q, const uint64_t* lab, const uint32_t* pre,
                          bool exact, std::string* why, const char** id) {
  const uint64_t S = c.node_count * 9;
  const uint32_t start_state = q.start_node * 9u + 8u;
  if (lab[start_state] != 0 || pre[start_state] != kNone) {
    *why = "GPU start state does not have the canonical zero label and no predecessor";
    return false;
  }
  for (uint64_t s = 0; s < S; ++s) {
    const uint32_t state = static_cast<uint32_t>(s);
    if (state == start_state) continue;
    const uint64_t label = lab[state];
    const uint32_t p = pre[state];
    const uint32_t inc = state % 9u;
    if (inc == 8u) {
      if (label != kInf || p != kNone) {
        *why = "GPU created a no-incoming-heading state away from the start";
        return false;
      }
      continue;
    }
    if (label == kInf) {
      if (p != kNone) {
        *why = "GPU unreachable state carries a predecessor";
        return false;
      }
      continue;
    }
    if (p == kNone && !exact) continue;
    if (p >= S) {
      *why = "GPU predecessor index is out of bounds";
      return false;
    }
    if (p == state) {
      *why = "GPU predecessor is self-referential";
      *id = "gpu.predecessor.self_reference.v1";
      return false;
    }
    const uint32_t node = state / 9u, pnode = p / 9u, pinc = p % 9u;
    if (!AdjacentLegal(c, pnode, node, inc)) {
      *why = "GPU predecessor does not follow one adjacent compiled legal edge";
      return false;
    }
    if (pinc == 8u && p != start_state) {
      *why = "GPU predecessor uses the start-only heading sentinel elsewhere";
      return false;
    }
    if (lab[p] == kInf) {
      *why = "GPU finite state follows an unreachable predecessor";
      return false;
    }
    uint64_t w, expected;
    bool banned;
    if (!Transition(c, q, pnode, inc, pinc, &w, &banned)) {
      if (banned) {
        *why = "GPU predecessor traverses a banned or unaddressable resource";
        *id = "gpu.policy.banned_predecessor.v1";
        return false;
      }
      *why = exact ? "GPU predecessor label does not reconstruct the exact scalar cost"
                   : "GPU partial predecessor label violates the scalar-cost lower bound";
      return false;
    }
    if (!CheckedAdd(lab[p], w, &expected) || (exact ? expected != label : expected > label)) {
      *why = exact ? "GPU predecessor label does not reconstruct the exact scalar cost"
                   : "GPU partial predecessor label violates the scalar-cost lower bound";
      return false;
    }
  }
  return true;
}

// The inputs of one batch check, shared by the first-failure and the
// per-query entry points.
struct Batch {
  const apgx_v1_header* header;
  Context c;
  uint64_t policy_count;
  const apgx_v1_batch_request* request;
  const apgx_v1_query* queries;
  uint32_t query_count;
  const apgx_v1_result* results;
  const apgx_v1_compact_path* paths;
  const uint32_t* states;
  uint64_t state_count;
  const apgx_v1_workspace_view* ws;
  const apgx_v1_batch_telemetry* t;
  GeneratorShape shape;
  bool full;
  uint64_t S;
  std::vector<uint32_t> result_of;  // readback position of each query's result; kNone: omitted or duplicated
};

bool Reject(apgx_v1_validation* v, uint32_t qi, uint32_t failure, const char* id, const std::string& detail) {
  std::memset(v, 0, sizeof(*v));
  v->query_index = qi;
  v->failure = failure;
  std::strncpy(v->invariant, id, sizeof(v->invariant) - 1);
  std::strncpy(v->detail, detail.c_str(), sizeof(v->detail) - 1);
  return false;
}

bool Invariant(apgx_v1_validation* v, uint32_t qi, const char* id, const std::string& detail) {
  return Reject(v, qi, APGX_V1_CHECK_INTERNAL_INVARIANT, id, detail);
}

void Accept(apgx_v1_validation* v, uint32_t qi) {
  std::memset(v, 0, sizeof(*v));
  v->valid = 1;
  v->query_index = qi;
  v->failure = APGX_V1_CHECK_OK;
}

// Checks that hold for the batch as a whole (a failure here fails every
// query): the result set names the admitted queries (:2185-2258), the readback
// kind and flat bounds, the canonical compact-path storage, and the memory and
// launch telemetry (ValidateCandidateBatchTelemetry, :1841-1955).  Fills
// result_of: results in query order map one to one; otherwise by query id,
// where an id that came back twice maps no query.
bool CheckBatch(Batch& b, apgx_v1_validation* v) {
  const uint32_t n = b.query_count;
  b.result_of.assign(n, kNone);
  bool canonical = true;
  for (uint32_t i = 0; i < n && canonical; ++i) canonical = b.results[i].query_id == b.queries[i].query_id;
  if (canonical) {
    for (uint32_t i = 0; i < n; ++i) b.result_of[i] = i;
  } else {
    std::vector<std::pair<uint64_t, uint32_t>> admitted(n), returned(n);
    for (uint32_t i = 0; i < n; ++i) {
      admitted[i] = {b.queries[i].query_id, i};
      returned[i] = {b.results[i].query_id, i};
    }
    std::sort(admitted.begin(), admitted.end());
    std::sort(returned.begin(), returned.end());
    const auto first_admitted = [&](uint64_t id) {
      return std::lower_bound(admitted.begin(), admitted.end(), std::make_pair(id, 0u));
    };
    for (const auto& r : returned) {
      const auto it = first_admitted(r.first);
      if (it == admitted.end() || it->first != r.first)
        return Invariant(v, kNone, "gpu.batch.query.identity.v1",
                         "GPU candidate batch returned an extra, foreign, or missing query result");
    }
    for (size_t k = 0; k < returned.size();) {
      size_t e = k + 1;
      while (e < returned.size() && returned[e].first == returned[k].first) ++e;
      if (e - k == 1)
        for (auto it = first_admitted(returned[k].first); it != admitted.end() && it->first == returned[k].first; ++it)
          b.result_of[it->second] = returned[k].second;
      k = e;
    }
  }

  const apgx_v1_batch_telemetry& t = *b.t;
  const apgx_v1_workspace_view& ws = *b.ws;
  const uint64_t total_states = static_cast<uint64_t>(n) * b.S;
  const bool arrays_empty = ws.label_count == 0 && ws.predecessor_count == 0 && ws.state_owner_count == 0 &&
                            ws.predecessor_owner_count == 0;
  if ((!b.full && (ws.readback_kind != APGX_V1_READBACK_COMPACT_PATHS || !arrays_empty ||
                   b.state_count > total_states || (n && b.paths == nullptr))) ||
      (b.full && (ws.readback_kind != APGX_V1_READBACK_FULL_WORKSPACE || ws.label_count != total_states ||
                  ws.predecessor_count != total_states || ws.state_owner_count != total_states ||
                  ws.predecessor_owner_count != total_states || b.state_count != 0 ||
                  (total_states && (!ws.labels || !ws.predecessors || !ws.state_owners || !ws.predecessor_owners)))))
    return Invariant(v, kNone, "gpu.batch.workspace.bounds.v1",
                     "GPU candidate batch readback kind or flat workspace bounds are invalid");
  if (!b.full) {
    uint64_t expected = 0;
    for (uint32_t i = 0; i < n; ++i) {
      if (b.paths[i].state_offset != expected)
        return Invariant(v, kNone, "gpu.batch.compact_path.bounds.v1",
                         "GPU compact sweep paths are not one canonical contiguous batch");
      expected += b.paths[i].state_count;
    }
    if (expected != b.state_count)
      return Invariant(v, kNone, "gpu.batch.compact_path.bounds.v1",
                       "GPU compact sweep path storage has trailing or missing states");
  }
  const uint64_t device_bytes = LayoutBatch(b.full, n, b.policy_count, b.c.node_count).total;
  if (t.persistent_device_bytes != b.header->estimated_persistent_device_bytes || t.batch_device_bytes != device_bytes ||
      t.batch_host_bytes != HostBytes(b.full, n, b.c.node_count, b.state_count))
    return Invariant(v, kNone, "gpu.batch.memory.accounting.v1",
                     "GPU candidate batch device or host memory telemetry disagrees with deterministic accounting");
  const apgx_v1_batch_request& rq = *b.request;
  if (t.peak_device_bytes != t.persistent_device_bytes + t.workspace_capacity_device_bytes ||
      t.peak_device_bytes < t.persistent_device_bytes || t.workspace_capacity_device_bytes < t.batch_device_bytes ||
      t.persistent_device_bytes > rq.maximum_device_bytes ||
      t.workspace_capacity_device_bytes > rq.maximum_device_bytes - t.persistent_device_bytes ||
      t.device_to_host_readback_bytes != ReadbackBytes(b.full, n, b.c.node_count, b.state_count) ||
      !std::isfinite(t.kernel_milliseconds) || t.kernel_milliseconds < 0.0 || t.chunk_rounds != b.shape.chunk_rounds ||
      t.dispatched_rounds > rq.maximum_rounds || t.finalization_launch_count > 1 ||
      (t.dispatched_rounds == 0 && t.finalization_launch_count != 1))
    return Invariant(v, kNone, "gpu.batch.telemetry.v1", "GPU candidate batch peak, timing, or chunk telemetry is invalid");
  const uint64_t readbacks =
      (static_cast<uint64_t>(t.dispatched_rounds) + b.shape.chunk_rounds - 1) / b.shape.chunk_rounds;
  if (t.kernel_launch_count != b.shape.fixed_launches + readbacks + t.finalization_launch_count ||
      t.blocking_status_readback_count != readbacks)
    return Invariant(v, kNone, "gpu.batch.telemetry.v1", "GPU candidate batch launch or blocking-readback telemetry is invalid");
  return true;
}

// One query's checks (the body of the reference's validate_query,
// :2259-2380): its envelope (ValidateCandidateBatchEnvelope, :1732-1839), its
// completion telemetry (:456-484), then the full workspace or the compact path.
// `visited` is scratch of (S + 63) / 64 words owned by the caller's worker.
bool CheckQuery(const Batch& b, uint32_t qi, std::vector<uint64_t>& visited, uint64_t* path_cost,
                apgx_v1_validation* v) {
  *path_cost = kInf;
  const uint32_t j = b.result_of[qi];
  if (j == kNone)
    return Invariant(v, qi, "gpu.batch.query.identity.v1", "GPU candidate batch omitted or duplicated a query result");
  const Context& c = b.c;
  const uint64_t S = b.S;
  const apgx_v1_batch_telemetry& t = *b.t;
  const apgx_v1_batch_request& rq = *b.request;
  const apgx_v1_header* header = b.header;
  const apgx_v1_query& q = b.queries[qi];
  const apgx_v1_result& r = b.results[j];
  const uint64_t total_states = static_cast<uint64_t>(b.query_count) * S;
  if (r.schema_version != APGX_V1_SCHEMA_VERSION || r.batch_id != q.batch_id || r.query_id != q.query_id ||
      r.workspace_owner != q.workspace_owner || r.input_ordinal != q.input_ordinal ||
      r.start_node != q.start_node || r.goal_node != q.goal_node || r.generator != q.generator ||
      r.policy_identity != q.policy_identity || r.routing_profile_fingerprint != q.routing_profile_fingerprint ||
      r.compiler_version != header->compiler_version ||
      r.source_board_content_hash != header->source_board_content_hash ||
      r.compiler_profile_fingerprint != header->compiler_profile_fingerprint ||
      r.rule_bucket_identity != header->rule_bucket_identity ||
      r.device_view_fingerprint != header->device_view_fingerprint || r.reserved[0] != 0 || r.reserved[1] != 0)
    return Reject(v, qi, APGX_V1_CHECK_VALIDATION_FAILED, "",
                  "GPU candidate batch result associations do not match its immutable query workspace");
  const apgx_v1_query_envelope& env = b.ws->queries[j];
  if (env.workspace_offset != q.workspace_offset || env.workspace_state_count != S)
    return Invariant(v, qi, "gpu.batch.workspace.bounds.v1",
                     "GPU candidate batch query-major workspace has invalid array bounds");
  const uint64_t off = env.workspace_offset;
  if (b.full) {
    if (off > total_states || env.workspace_state_count > total_states - off)
      return Invariant(v, qi, "gpu.batch.workspace.bounds.v1",
                       "GPU candidate batch query-major workspace has invalid array bounds");
    for (uint64_t k = 0; k < S; ++k)
      if (b.ws->state_owners[off + k] != q.workspace_owner || b.ws->predecessor_owners[off + k] != q.workspace_owner)
        return Invariant(v, qi, "gpu.batch.workspace.owner.v1",
                         "GPU candidate batch workspace contains foreign query ownership");
  } else {
    const apgx_v1_compact_path& p = b.paths[j];
    const bool reached = r.completion == APGX_V1_REACHED;
    const uint64_t end = p.state_offset + static_cast<uint64_t>(p.state_count);
    if (p.schema_version != APGX_V1_SCHEMA_VERSION || p.query_id != q.query_id ||
        p.workspace_owner != q.workspace_owner || p.state_count > S || end < p.state_offset || end > b.state_count ||
        (reached ? p.state_count == 0 : p.state_count != 0))
      return Invariant(v, qi, "gpu.batch.compact_path.bounds.v1",
                       "GPU compact sweep path has invalid associations or array bounds");
  }
  if (env.rounds != r.rounds || r.rounds > t.dispatched_rounds)
    return Invariant(v, qi, "gpu.batch.query.telemetry.v1",
                     "GPU candidate query round telemetry disagrees with its result header or shared dispatch envelope");
  if ((r.completion == APGX_V1_BUDGET_EXHAUSTED || r.completion == APGX_V1_CANCELLED) &&
      t.finalization_launch_count != 1)
    return Invariant(v, qi, "gpu.batch.query.telemetry.v1",
                     "GPU unfinished query completion lacks its shared finalization launch");
  const unsigned __int128 max_examined =
      static_cast<unsigned __int128>(env.rounds) *
      (b.shape.fixed_examined_work_per_round +
       static_cast<unsigned __int128>(c.node_count) * b.shape.examined_work_per_node_per_round);
  if (env.examined_work > max_examined ||
      env.heading_turn_relaxations > max_examined * b.shape.turn_relaxations_per_examined_work)
    return Invariant(v, qi, "gpu.batch.query.telemetry.v1",
                     "GPU candidate query work telemetry exceeds deterministic bounds");
  if (r.completion > APGX_V1_CANCELLED)
    return Invariant(v, qi, "gpu.completion.unknown.v1", "GPU result contains an unknown completion code");
  if (r.rounds > rq.maximum_rounds)
    return Invariant(v, qi, "", "GPU round telemetry exceeds the configured execution budget");
  if (r.completion == APGX_V1_BUDGET_EXHAUSTED && r.rounds != rq.maximum_rounds)
    return Invariant(v, qi, "", "GPU reported budget exhaustion before consuming its round budget");
  if ((r.completion == APGX_V1_REACHED || r.completion == APGX_V1_DISCONNECTED) && r.rounds == 0)
    return Invariant(v, qi, "", "GPU reported a converged completion without an execution round");

  const uint32_t start_state = q.start_node * 9u + 8u;
  if (b.full) {
    const uint64_t* lab = b.ws->labels + off;
    const uint32_t* pre = b.ws->predecessors + off;
    const bool exact = r.completion == APGX_V1_REACHED || r.completion == APGX_V1_DISCONNECTED;
    std::string why;
    const char* id = "";
    if (!AllPredecessorsValid(c, q, lab, pre, exact, &why, &id)) return Invariant(v, qi, id, why);
    if (r.completion == APGX_V1_REACHED) {
      uint64_t best = kInf;
      uint32_t best_state = kNone;
      for (uint32_t h = 0; h < 8; ++h) {
        const uint32_t st = q.goal_node * 9u + h;
        if (lab[st] < best || (lab[st] == best && st < best_state)) {
          best = lab[st];
          best_state = st;
        }
      }
      if (best == kInf || r.goal_state != best_state)
        return Invariant(v, qi, "", "GPU goal state is not the stable minimum finite heading state");
      uint32_t cursor = best_state;
      for (uint64_t steps = 0;; ++steps) {
        if (cursor >= S) return Invariant(v, qi, "", "GPU reconstruction escaped the allocated state array");
        if (cursor == start_state) break;
        if (steps >= S) return Invariant(v, qi, "", "GPU predecessor reconstruction exceeded its acyclic state bound");
        cursor = pre[cursor];
      }
      *path_cost = best;
    } else if (r.completion == APGX_V1_DISCONNECTED) {
      for (uint32_t h = 0; h < 8; ++h)
        if (lab[q.goal_node * 9u + h] != kInf)
          return Invariant(v, qi, "", "GPU reported disconnected with a finite goal label");
      if (GoalReachable(c, q))
        return Invariant(v, qi, "gpu.disconnected.cpu_reachable.v1",
                         "GPU reported disconnected for a query the CPU oracle reached under the identical policy");
    }
    Accept(v, qi);
    return true;
  }
  // Compact path (ValidateAndReconstructCompactGpuPath, :969-1131).
  const apgx_v1_compact_path& p = b.paths[j];
  if (r.completion == APGX_V1_DISCONNECTED && GoalReachable(c, q))
    return Invariant(v, qi, "gpu.disconnected.cpu_reachable.v1",
                     "GPU reported disconnected for a query the CPU oracle reached under the identical policy");
  if (r.completion != APGX_V1_REACHED) {
    Accept(v, qi);
    return true;
  }
  const uint32_t* rev = b.states + p.state_offset;
  const uint32_t n = p.state_count;
  if (rev[0] != r.goal_state || rev[n - 1] != start_state || r.goal_state >= S ||
      r.goal_state / 9u != q.goal_node || r.goal_state % 9u >= 8u)
    return Invariant(v, qi, "gpu.batch.compact_path.endpoint.v1", "path does not bind the goal to the canonical start");
  for (uint32_t k = 0; k < n; ++k) {
    if (rev[k] >= S) return Invariant(v, qi, "gpu.batch.compact_path.bounds.v1", "path state out of range");
    visited[rev[k] / 64] = 0;
  }
  uint64_t total = 0;
  for (uint32_t k = 0; k < n; ++k) {
    const uint32_t st = rev[n - 1 - k];
    const uint64_t mask = 1ull << (st % 64);
    if (visited[st / 64] & mask) return Invariant(v, qi, "gpu.batch.compact_path.cycle.v1", "repeated path state");
    visited[st / 64] |= mask;
    if (k == 0) continue;
    const uint32_t pred = rev[n - k];
    const uint32_t pnode = pred / 9u, node = st / 9u, inc = st % 9u, pinc = pred % 9u;
    if (inc >= 8u || (pinc == 8u && pred != start_state))
      return Invariant(v, qi, "gpu.batch.compact_path.heading.v1", "noncanonical incoming heading");
    if (!AdjacentLegal(c, pnode, node, inc))
      return Invariant(v, qi, "gpu.batch.compact_path.edge.v1", "path step is not one adjacent legal edge");
    uint64_t w;
    bool banned;
    if (!Transition(c, q, pnode, inc, pinc, &w, &banned)) {
      if (banned) return Invariant(v, qi, "gpu.policy.banned_predecessor.v1", "path uses a banned resource");
      return Invariant(v, qi, "gpu.batch.compact_path.edge.v1", "path step cost is invalid or overflows");
    }
    if (!CheckedAdd(total, w, &total))
      return Invariant(v, qi, "gpu.batch.compact_path.edge.v1", "path cost overflows");
  }
  *path_cost = total;
  Accept(v, qi);
  return true;
}

// Argument checks and the Batch both entry points start from.
int MakeBatch(const apgx_v1_header* header, const apgx_v1_node* nodes, uint64_t node_count,
               const apgx_v1_batch_request* request, const apgx_v1_query* queries, uint32_t query_count,
               const apgx_v1_policy_edge* policy, uint64_t policy_count, const apgx_v1_result* results,
               const apgx_v1_compact_path* paths, const uint32_t* states, uint64_t state_count,
               const apgx_v1_workspace_view* workspace, const apgx_v1_batch_telemetry* telemetry, const char* what,
               Batch* b) {
  if (header == nullptr || nodes == nullptr || request == nullptr || telemetry == nullptr || workspace == nullptr ||
      (query_count && (queries == nullptr || results == nullptr)) || (policy_count && policy == nullptr) ||
      (state_count && states == nullptr))
    return Fail(APGX_E_INTERNAL_INVARIANT, std::string(what) + ": null argument");
  if (query_count && workspace->queries == nullptr)
    return Fail(APGX_E_INTERNAL_INVARIANT, std::string(what) + ": null workspace envelopes");
  *b = Batch{header, Context{header, nodes, node_count, policy}, policy_count, request, queries, query_count,
             results, paths, states, state_count, workspace, telemetry, GeneratorShape{}, false, node_count * 9, {}};
  return APGX_OK;
}

// Validation workers for a batch of `queries`: one per 64 queries, at most 8
// (CandidateCompactValidationWorkerCountV1, include/apgar/gpu/planar_router.h:439-450).
uint32_t WorkerCount(uint32_t queries) {
  return queries == 0 ? 0 : std::min<uint32_t>(8, 1 + (queries - 1) / 64);
}

}  // namespace
}  // namespace v1check
}  // namespace apgx

extern "C" int apgx_v1_validate_batch(const apgx_v1_header* header, const apgx_v1_node* nodes,
                                      uint64_t node_count, const apgx_v1_batch_request* request,
                                      const apgx_v1_query* queries, uint32_t query_count,
                                      const apgx_v1_policy_edge* policy, uint64_t policy_count,
                                      const apgx_v1_result* results, const apgx_v1_compact_path* paths,
                                      const uint32_t* states, uint64_t state_count,
                                      const apgx_v1_workspace_view* workspace,
                                      const apgx_v1_batch_telemetry* telemetry, uint64_t* path_costs,
                                      apgx_v1_validation* out) {
  using namespace apgx;
  using namespace apgx::v1check;
  if (out == nullptr) return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_v1_validate_batch: null argument");
  Batch b;
  if (int e = MakeBatch(header, nodes, node_count, request, queries, query_count, policy, policy_count, results,
                        paths, states, state_count, workspace, telemetry, "apgx_v1_validate_batch", &b);
      e != APGX_OK)
    return e;
  for (uint32_t i = 0; i < query_count && path_costs; ++i) path_costs[i] = kInf;
  ClearError();
  if (!v1::DescribeGenerator(request->generator, &b.shape)) {
    Reject(out, kNone, APGX_V1_CHECK_VALIDATION_FAILED, "gpu.generator.unknown.v1",
           "GPU candidate batch telemetry names an unknown generator");
    return APGX_OK;
  }
  b.full = b.shape.full_workspace;
  if (!CheckBatch(b, out)) return APGX_OK;
  std::vector<uint64_t> visited((b.S + 63) / 64, 0);
  for (uint32_t qi = 0; qi < query_count; ++qi) {
    uint64_t cost;
    if (!CheckQuery(b, qi, visited, &cost, out)) return APGX_OK;
    if (path_costs) path_costs[qi] = cost;
  }
  Accept(out, kNone);
  return APGX_OK;
}

extern "C" int apgx_v1_validate_batch_queries(const apgx_v1_header* header, const apgx_v1_node* nodes,
                                              uint64_t node_count, const apgx_v1_batch_request* request,
                                              const apgx_v1_query* queries, uint32_t query_count,
                                              const apgx_v1_policy_edge* policy, uint64_t policy_count,
                                              const apgx_v1_result* results, const apgx_v1_compact_path* paths,
                                              const uint32_t* states, uint64_t state_count,
                                              const apgx_v1_workspace_view* workspace,
                                              const apgx_v1_batch_telemetry* telemetry, uint64_t* path_costs,
                                              apgx_v1_validation* verdicts, apgx_v1_validation* batch) {
  using namespace apgx;
  using namespace apgx::v1check;
  if (batch == nullptr || (query_count && verdicts == nullptr))
    return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_v1_validate_batch_queries: null argument");
  Batch b;
  if (int e = MakeBatch(header, nodes, node_count, request, queries, query_count, policy, policy_count, results,
                        paths, states, state_count, workspace, telemetry, "apgx_v1_validate_batch_queries", &b);
      e != APGX_OK)
    return e;
  for (uint32_t i = 0; i < query_count && path_costs; ++i) path_costs[i] = kInf;
  ClearError();
  bool ok = v1::DescribeGenerator(request->generator, &b.shape);
  if (!ok) {
    Reject(batch, kNone, APGX_V1_CHECK_VALIDATION_FAILED, "gpu.generator.unknown.v1",
           "GPU candidate batch telemetry names an unknown generator");
  } else {
    b.full = b.shape.full_workspace;
    ok = CheckBatch(b, batch);
  }
  if (!ok) {
    // A batch-level failure is every query's verdict (SetAdmittedBatchFailure).
    for (uint32_t i = 0; i < query_count; ++i) verdicts[i] = *batch;
    return APGX_OK;
  }
  Accept(batch, kNone);
  // Contiguous partitions, one per worker; every verdict has its own slot, so
  // the output does not depend on scheduling.  A worker that cannot be started
  // runs on the caller.
  const uint32_t workers = WorkerCount(query_count);
  std::vector<uint64_t> costs(query_count, kInf);
  const auto run = [&](uint32_t w) {
    std::vector<uint64_t> visited((b.S + 63) / 64, 0);
    const uint32_t begin = static_cast<uint32_t>(uint64_t{query_count} * w / workers);
    const uint32_t end = static_cast<uint32_t>(uint64_t{query_count} * (w + 1) / workers);
    for (uint32_t qi = begin; qi < end; ++qi) CheckQuery(b, qi, visited, &costs[qi], &verdicts[qi]);
  };
  std::vector<std::thread> threads;
  uint32_t next = 0;
  try {
    for (; next + 1 < workers; ++next) threads.emplace_back(run, next);
  } catch (const std::system_error&) {
  }
  for (; next < workers; ++next) run(next);
  for (std::thread& t : threads) t.join();
  for (uint32_t i = 0; i < query_count && path_costs; ++i) path_costs[i] = costs[i];
  return APGX_OK;
}


namespace apgx {
namespace v1route {
uint64_t RouteBytes(bool frontier, uint64_t node_count);
}  // namespace v1route
}  // namespace apgx

extern "C" int apgx_v1_validate_route(const apgx_v1_header* header, const apgx_v1_node* nodes, uint64_t node_count,
                                      const apgx_v1_route_request* request, const apgx_v1_route_result* result,
                                      const uint64_t* labels, const uint32_t* predecessors, uint64_t* path_cost,
                                      apgx_v1_validation* out) {
  using namespace apgx;
  using namespace apgx::v1;
  using namespace apgx::v1check;
  if (out == nullptr || header == nullptr || (node_count && nodes == nullptr) || request == nullptr ||
      result == nullptr)
    return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_v1_validate_route: null argument");
  std::memset(out, 0, sizeof(*out));
  out->valid = 1;
  if (path_cost) *path_cost = kInf;
  auto reject = [&](uint32_t failure, const char* invariant, const std::string& detail) {
    out->valid = 0;
    out->failure = failure;
    std::strncpy(out->invariant, invariant, sizeof(out->invariant) - 1);
    std::strncpy(out->detail, detail.c_str(), sizeof(out->detail) - 1);
    ClearError();
    return APGX_OK;
  };
  auto invariant = [&](const char* id, const std::string& detail) {
    return reject(APGX_V1_CHECK_INTERNAL_INVARIANT, id, detail);
  };
  auto failed = [&](const std::string& detail) { return reject(APGX_V1_CHECK_VALIDATION_FAILED, "", detail); };
  const uint64_t S = node_count * 9;
  const apgx_v1_kernel_telemetry& t = result->telemetry;
  const apgx_v1_route_header& rh = result->header;

  // ValidateKernelTelemetry (src/gpu/planar_router.cc:411-454).
  GeneratorShape shape{};
  if (!DescribeGenerator(request->generator, &shape))
    return reject(APGX_V1_CHECK_VALIDATION_FAILED, "gpu.generator.unknown.v1",
                  "GPU telemetry validation names an unknown planar generator");
  if (t.persistent_device_bytes != header->estimated_persistent_device_bytes)
    return invariant("", "GPU persistent-memory telemetry disagrees with deterministic accounting");
  if (t.batch_device_bytes != v1route::RouteBytes(request->generator == APGX_V1_GENERATOR_BUCKETED_FRONTIER, node_count))
    return invariant("", "GPU batch-memory telemetry disagrees with deterministic accounting");
  uint64_t peak = 0;
  if (!CheckedAdd(t.persistent_device_bytes, t.batch_device_bytes, &peak) || t.peak_device_bytes != peak)
    return invariant("", "GPU peak-memory telemetry disagrees with simultaneous owned tiers");
  if (!std::isfinite(t.kernel_milliseconds) || t.kernel_milliseconds < 0.0)
    return invariant("", "GPU kernel-time telemetry is not finite and nonnegative");

  // ValidateRawAssociations (:382-409).
  if (rh.schema_version != APGX_V1_ROUTE_RESULT_SCHEMA_VERSION)
    return failed("GPU result schema does not match the route evidence contract");
  if (rh.source_board_content_hash != header->source_board_content_hash ||
      rh.compiler_profile_fingerprint != header->compiler_profile_fingerprint ||
      rh.compiler_version != header->compiler_version || rh.rule_bucket_identity != header->rule_bucket_identity ||
      rh.device_view_fingerprint != header->device_view_fingerprint)
    return failed("GPU result associations do not match the immutable device view");
  if (rh.routing_profile_fingerprint != request->routing_profile_fingerprint)
    return failed("GPU result prepared routing context is invalid or mismatched");
  if (rh.start_node != request->start_node || rh.goal_node != request->goal_node || rh.generator != request->generator)
    return failed("GPU result endpoint or generator associations do not match the route request");

  // ValidateCompletionTelemetry (:456-484).
  const uint32_t completion = result->completion;
  if (completion > APGX_V1_CANCELLED)
    return invariant("gpu.completion.unknown.v1", "GPU result contains an unknown completion code");
  if (t.rounds > request->maximum_rounds)
    return invariant("", "GPU round telemetry exceeds the configured execution budget");
  if (completion == APGX_V1_BUDGET_EXHAUSTED && t.rounds != request->maximum_rounds)
    return invariant("", "GPU reported budget exhaustion before consuming its round budget");
  if ((completion == APGX_V1_REACHED || completion == APGX_V1_DISCONNECTED) && t.rounds == 0)
    return invariant("", "GPU reported a converged completion without an execution round");

  // ValidateAllPredecessors (:492-612): exact when converged, a lower bound
  // otherwise; the reached path re-runs it exactly below, as the reference does.
  if (result->state_count != S || labels == nullptr || predecessors == nullptr)
    return invariant("", "GPU label or predecessor array has an invalid bound");
  if (request->start_node >= node_count)
    return invariant("", "GPU start node maps outside the flattened state array");
  if (request->goal_node >= node_count) return invariant("", "GPU goal node maps outside the flattened state array");
  const Context c{header, nodes, node_count, nullptr};
  apgx_v1_query q;
  std::memset(&q, 0, sizeof(q));
  q.start_node = request->start_node;
  q.goal_node = request->goal_node;
  std::string why;
  const char* id = "";
  const bool exact = completion == APGX_V1_REACHED || completion == APGX_V1_DISCONNECTED;
  if (!AllPredecessorsValid(c, q, labels, predecessors, exact, &why, &id)) return invariant(id, why);

  if (completion == APGX_V1_REACHED) {
    // ValidateAndReconstructGpuRouteWithPolicy (:829-930): the goal state and
    // the predecessor walk from it to the start.
    uint64_t best = kInf;
    uint32_t best_state = kNone;
    for (uint32_t h = 0; h < 8; ++h) {
      const uint32_t st = request->goal_node * 9u + h;
      if (labels[st] < best || (labels[st] == best && st < best_state)) {
        best = labels[st];
        best_state = st;
      }
    }
    if (best == kInf || result->goal_state != best_state)
      return invariant("", "GPU goal state is not the stable minimum finite heading state");
    const uint32_t start_state = request->start_node * 9u + 8u;
    uint32_t cursor = best_state;
    uint64_t points = 0;
    for (;;) {
      if (cursor >= S) return invariant("", "GPU reconstruction escaped the allocated state array");
      ++points;
      if (cursor == start_state) break;
      cursor = predecessors[cursor];
      if (points > S) return invariant("", "GPU predecessor reconstruction exceeded its acyclic state bound");
    }
    if (path_cost) *path_cost = best;
  } else if (completion == APGX_V1_DISCONNECTED) {
    // ValidateDisconnectedResult (:613-627) and the CPU confirmation (:629-667).
    for (uint32_t h = 0; h < 8; ++h)
      if (labels[request->goal_node * 9ull + h] != kInf)
        return invariant("", "GPU reported disconnected with a finite goal label");
    if (GoalReachable(c, q))
      return invariant("gpu.disconnected.cpu_reachable.v1",
                       "GPU reported disconnected for a query the CPU oracle reached under the identical normalized "
                       "policy");
  }
  ClearError();
  return APGX_OK;
}

extern "C" int apgx_v1_generator_shape(uint32_t generator, apgx_v1_generator_shape_info* out) {
  using namespace apgx;
  using namespace apgx::v1;
  if (out == nullptr) return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_v1_generator_shape: null argument");
  GeneratorShape g{};
  if (!DescribeGenerator(generator, &g)) return Fail(APGX_E_UNSUPPORTED, "unknown candidate generator");
  *out = apgx_v1_generator_shape_info{g.chunk_rounds,
                                      g.fixed_launches,
                                      g.fixed_examined_work_per_round,
                                      g.examined_work_per_node_per_round,
                                      g.turn_relaxations_per_examined_work,
                                      g.full_workspace ? APGX_V1_READBACK_FULL_WORKSPACE
                                                       : APGX_V1_READBACK_COMPACT_PATHS,
                                      1u,
                                      0u};
  ClearError();
  return APGX_OK;
}

extern "C" int apgx_v1_expected_batch_bytes(uint32_t generator, uint64_t query_count, uint64_t policy_edge_count,
                                            uint64_t node_count, uint64_t compact_path_states,
                                            uint64_t* device_bytes, uint64_t* host_bytes,
                                            uint64_t* readback_bytes) {
  using namespace apgx;
  using namespace apgx::v1;
  GeneratorShape g{};
  if (!DescribeGenerator(generator, &g)) return Fail(APGX_E_UNSUPPORTED, "unknown candidate generator");
  if (device_bytes) *device_bytes = LayoutBatch(g.full_workspace, query_count, policy_edge_count, node_count).total;
  if (host_bytes) *host_bytes = HostBytes(g.full_workspace, query_count, node_count, compact_path_states);
  if (readback_bytes) *readback_bytes = ReadbackBytes(g.full_workspace, query_count, node_count, compact_path_states);
  ClearError();
  return APGX_OK;
}
