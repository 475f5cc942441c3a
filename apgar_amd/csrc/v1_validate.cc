// v1_validate.cc — host trust boundary for V1 candidate batches.
//
// Re-checks an untrusted compact-path readback against the view, the request
// and the queries before any result is used, as the reference's host router
// does (src/gpu/planar_router.cc:969-1131 compact paths, :1808-1955 query and
// batch telemetry, :629-667 CPU confirmation of "disconnected").  A failure
// names the same invariant identifiers the reference reports, so callers and
// replays can assert on them.  Nothing here produces a route: the validator
// only accepts or rejects what the device returned.
#include <algorithm>
#include <cstring>
#include <queue>
#include <string>
#include <vector>

#include "apgx.h"
#include "common.h"

namespace apgx {
namespace v1check {
namespace {

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

}  // namespace
}  // namespace v1check
}  // namespace apgx

extern "C" int apgx_v1_validate_batch(const apgx_v1_header* header, const apgx_v1_node* nodes,
                                      uint64_t node_count, const apgx_v1_batch_request* request,
                                      const apgx_v1_query* queries, uint32_t query_count,
                                      const apgx_v1_policy_edge* policy, uint64_t policy_count,
                                      const apgx_v1_result* results, const apgx_v1_compact_path* paths,
                                      const uint32_t* states, uint64_t state_count,
                                      const apgx_v1_batch_telemetry* telemetry, uint64_t* path_costs,
                                      apgx_v1_validation* out) {
  using namespace apgx;
  using namespace apgx::v1check;
  if (out == nullptr || header == nullptr || nodes == nullptr || request == nullptr || telemetry == nullptr ||
      (query_count && (queries == nullptr || results == nullptr || paths == nullptr)) ||
      (policy_count && policy == nullptr) || (state_count && states == nullptr))
    return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_v1_validate_batch: null argument");
  std::memset(out, 0, sizeof(*out));
  out->query_index = kNone;
  auto reject = [&](uint32_t qi, const char* invariant, const std::string& detail) {
    out->valid = 0;
    out->query_index = qi;
    std::strncpy(out->invariant, invariant, sizeof(out->invariant) - 1);
    std::strncpy(out->detail, detail.c_str(), sizeof(out->detail) - 1);
    ClearError();
    return APGX_OK;
  };
  const Context c{header, nodes, node_count, policy};
  const uint64_t S = node_count * 9;
  const bool frontier = request->generator == APGX_V1_GENERATOR_BUCKETED_FRONTIER;
  if (request->generator != APGX_V1_GENERATOR_BUCKETED_FRONTIER &&
      request->generator != APGX_V1_GENERATOR_HEADING_AWARE_SWEEP)
    return reject(kNone, "gpu.generator.unknown.v1", "request names an unknown generator");
  uint64_t legal_edges = 0;
  for (uint64_t i = 0; i < node_count; ++i) legal_edges += __builtin_popcount(nodes[i].legal_edges);

  // Batch telemetry identities.
  const apgx_v1_batch_telemetry& t = *telemetry;
  const uint32_t chunk = frontier ? 32u : 8u;
  const uint64_t chunks = (static_cast<uint64_t>(t.dispatched_rounds) + chunk - 1) / chunk;
  if (t.chunk_rounds != chunk || t.dispatched_rounds > request->maximum_rounds ||
      t.blocking_status_readback_count != chunks || t.kernel_launch_count != 1 + chunks + 3 ||
      t.finalization_launch_count > 1)
    return reject(kNone, "gpu.batch.telemetry.v1", "chunk, readback or launch counts disagree with the generator");
  uint64_t used = 0;
  for (uint32_t i = 0; i < query_count; ++i) used += paths[i].state_count;
  const uint64_t record_bytes = sizeof(apgx_v1_result) + sizeof(apgx_v1_compact_path);
  if (used != state_count || t.device_to_host_readback_bytes != query_count * record_bytes + used * 4 ||
      t.batch_host_bytes != query_count * (record_bytes + sizeof(apgx_v1_query)) + used * 4 ||
      t.peak_device_bytes != t.persistent_device_bytes + t.workspace_capacity_device_bytes ||
      t.workspace_capacity_device_bytes < t.batch_device_bytes ||
      t.peak_device_bytes > request->maximum_device_bytes || t.persistent_device_bytes != header->estimated_persistent_device_bytes)
    return reject(kNone, "gpu.batch.memory.accounting.v1", "byte accounting disagrees with the readback");

  std::vector<uint64_t> visited((S + 63) / 64, 0);
  for (uint32_t qi = 0; qi < query_count; ++qi) {
    const apgx_v1_query& q = queries[qi];
    const apgx_v1_result& r = results[qi];
    const apgx_v1_compact_path& p = paths[qi];
    if (path_costs) path_costs[qi] = kInf;
    if (r.schema_version != APGX_V1_SCHEMA_VERSION || r.batch_id != q.batch_id || r.query_id != q.query_id ||
        r.workspace_owner != q.workspace_owner || r.input_ordinal != q.input_ordinal ||
        r.start_node != q.start_node || r.goal_node != q.goal_node || r.generator != q.generator ||
        r.policy_identity != q.policy_identity || r.routing_profile_fingerprint != q.routing_profile_fingerprint ||
        r.compiler_version != header->compiler_version ||
        r.source_board_content_hash != header->source_board_content_hash ||
        r.compiler_profile_fingerprint != header->compiler_profile_fingerprint ||
        r.rule_bucket_identity != header->rule_bucket_identity ||
        r.device_view_fingerprint != header->device_view_fingerprint)
      return reject(qi, "gpu.batch.query.identity.v1", "result header does not echo its query and view");
    if (r.completion > APGX_V1_CANCELLED)
      return reject(qi, "gpu.completion.unknown.v1", "unknown completion value");
    const bool examined_ok = frontier ? r.examined_work <= 8ull * r.rounds
                                      : r.examined_work == static_cast<uint64_t>(r.rounds) * legal_edges;
    if (r.rounds > t.dispatched_rounds || !examined_ok || r.heading_turn_relaxations > 8 * r.examined_work ||
        (r.completion == APGX_V1_REACHED && r.rounds == 0) ||
        (r.completion == APGX_V1_BUDGET_EXHAUSTED && r.rounds < std::min(request->maximum_rounds, q.maximum_rounds)))
      return reject(qi, "gpu.batch.query.telemetry.v1", "query rounds or work disagree with the batch");
    if (p.schema_version != APGX_V1_SCHEMA_VERSION || p.query_id != q.query_id ||
        p.workspace_owner != q.workspace_owner ||
        (r.completion != APGX_V1_REACHED && p.state_count != 0) ||
        (r.completion == APGX_V1_REACHED && (p.state_count < 2 || p.state_count > S)) ||
        p.state_offset > state_count || p.state_count > state_count - p.state_offset)
      return reject(qi, "gpu.batch.compact_path.bounds.v1", "compact path header out of bounds");
    if (r.completion == APGX_V1_DISCONNECTED && GoalReachable(c, q))
      return reject(qi, "gpu.disconnected.cpu_reachable.v1", "the CPU search reaches a goal reported disconnected");
    if (r.completion != APGX_V1_REACHED) continue;
    const uint32_t* rev = states + p.state_offset;
    const uint32_t n = p.state_count;
    const uint32_t start_state = q.start_node * 9u + 8u;
    if (rev[0] != r.goal_state || rev[n - 1] != start_state || r.goal_state >= S ||
        r.goal_state / 9u != q.goal_node || r.goal_state % 9u >= 8u)
      return reject(qi, "gpu.batch.compact_path.endpoint.v1", "path does not bind the goal to the canonical start");
    for (uint32_t k = 0; k < n; ++k) {
      if (rev[k] >= S) return reject(qi, "gpu.batch.compact_path.bounds.v1", "path state out of range");
      visited[rev[k] / 64] = 0;
    }
    uint64_t total = 0;
    for (uint32_t k = 0; k < n; ++k) {
      const uint32_t st = rev[n - 1 - k];
      const uint64_t mask = 1ull << (st % 64);
      if (visited[st / 64] & mask) return reject(qi, "gpu.batch.compact_path.cycle.v1", "repeated path state");
      visited[st / 64] |= mask;
      if (k == 0) continue;
      const uint32_t pred = rev[n - k];
      const uint32_t pnode = pred / 9u, node = st / 9u, inc = st % 9u, pinc = pred % 9u;
      if (inc >= 8u || (pinc == 8u && pred != start_state))
        return reject(qi, "gpu.batch.compact_path.heading.v1", "noncanonical incoming heading");
      const apgx_v1_node& pn = nodes[pnode];
      const int64_t dx = nodes[node].lattice_x - pn.lattice_x, dy = nodes[node].lattice_y - pn.lattice_y;
      static const int kDx[8] = {1, 1, 0, -1, -1, -1, 0, 1}, kDy[8] = {0, 1, 1, 1, 0, -1, -1, -1};
      if (pn.neighbors[inc] != node || (pn.legal_edges & (1u << inc)) == 0 || dx != kDx[inc] || dy != kDy[inc] ||
          pn.layer != nodes[node].layer)
        return reject(qi, "gpu.batch.compact_path.edge.v1", "path step is not one adjacent legal edge");
      uint64_t w;
      bool banned;
      if (!Transition(c, q, pnode, inc, pinc, &w, &banned)) {
        if (banned) return reject(qi, "gpu.policy.banned_predecessor.v1", "path uses a banned resource");
        return reject(qi, "gpu.batch.compact_path.edge.v1", "path step cost is invalid or overflows");
      }
      if (!CheckedAdd(total, w, &total))
        return reject(qi, "gpu.batch.compact_path.edge.v1", "path cost overflows");
    }
    if (path_costs) path_costs[qi] = total;
  }
  out->valid = 1;
  ClearError();
  return APGX_OK;
}
