// replay.cc — candidate-batch invariant replays (gpu_batch_replay v1).
//
// A replay artifact pins a fixture, a forced generator, one readback fault, the
// execution budgets, the batch/query correlation identity, the request-local
// candidate policy and every board association, plus the expected failure
// class and invariant identifier (schemas/gpu_batch_replay/v1.md).  The runner
// follows tools/gpu_batch_replay.cc:26-94 and :226-391: parse the canonical
// envelope, verify the fixture checksum, compile the fixture into the device
// view, reject association drift before any GPU work, execute the one query,
// corrupt its readback with the fault decorator, validate it at the host trust
// boundary and compare the outcome with the expected one.
//
// The reference's fixtures are KiCad boards compiled by its Board IR and
// geometry compiler, which this library does not contain.  Its fixtures here
// are the compiler's output instead: the compiled profile scalars, the
// represented nodes with their legal-edge masks and the request endpoints, in
// a small binary record (apgx_replay_fixture_header + nodes) that
// apgx_v1_build_board turns into the same DeviceCompiledBoardV1.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <string_view>
#include <vector>

#include "apgx.h"
#include "common.h"
#include "v1_common.h"

namespace apgx {
namespace replay {
namespace {

constexpr std::string_view kAssociationError =
    "replay format, schema, generator, policy, fault, or expected outcome is unsupported";

struct FaultContract {
  std::string_view name;
  uint32_t fault;
  uint32_t generator;
  std::string_view expected_failure;
  std::string_view invariant;
};

// The thirteen v1 contracts (tools/gpu_batch_replay.cc:26-66).
constexpr FaultContract kContracts[] = {
    {"workspace_bounds", APGX_V1_FAULT_WORKSPACE_BOUNDS, APGX_V1_GENERATOR_BUCKETED_FRONTIER, "internal_invariant",
     "gpu.batch.workspace.bounds.v1"},
    {"workspace_owner", APGX_V1_FAULT_WORKSPACE_OWNER, APGX_V1_GENERATOR_BUCKETED_FRONTIER, "internal_invariant",
     "gpu.batch.workspace.owner.v1"},
    {"query_telemetry", APGX_V1_FAULT_QUERY_TELEMETRY, APGX_V1_GENERATOR_BUCKETED_FRONTIER, "internal_invariant",
     "gpu.batch.query.telemetry.v1"},
    {"memory_accounting", APGX_V1_FAULT_MEMORY_ACCOUNTING, APGX_V1_GENERATOR_BUCKETED_FRONTIER, "internal_invariant",
     "gpu.batch.memory.accounting.v1"},
    {"batch_telemetry", APGX_V1_FAULT_BATCH_TELEMETRY, APGX_V1_GENERATOR_BUCKETED_FRONTIER, "internal_invariant",
     "gpu.batch.telemetry.v1"},
    {"query_identity", APGX_V1_FAULT_QUERY_IDENTITY, APGX_V1_GENERATOR_BUCKETED_FRONTIER, "internal_invariant",
     "gpu.batch.query.identity.v1"},
    {"false_disconnected", APGX_V1_FAULT_FALSE_DISCONNECTED, APGX_V1_GENERATOR_BUCKETED_FRONTIER,
     "internal_invariant", "gpu.disconnected.cpu_reachable.v1"},
    {"compact_path_bounds", APGX_V1_FAULT_COMPACT_PATH_BOUNDS, APGX_V1_GENERATOR_HEADING_AWARE_SWEEP,
     "internal_invariant", "gpu.batch.compact_path.bounds.v1"},
    {"compact_path_endpoint", APGX_V1_FAULT_COMPACT_PATH_ENDPOINT, APGX_V1_GENERATOR_HEADING_AWARE_SWEEP,
     "internal_invariant", "gpu.batch.compact_path.endpoint.v1"},
    {"compact_path_cycle", APGX_V1_FAULT_COMPACT_PATH_CYCLE, APGX_V1_GENERATOR_HEADING_AWARE_SWEEP,
     "internal_invariant", "gpu.batch.compact_path.cycle.v1"},
    {"compact_path_heading", APGX_V1_FAULT_COMPACT_PATH_HEADING, APGX_V1_GENERATOR_HEADING_AWARE_SWEEP,
     "internal_invariant", "gpu.batch.compact_path.heading.v1"},
    {"compact_path_edge", APGX_V1_FAULT_COMPACT_PATH_EDGE, APGX_V1_GENERATOR_HEADING_AWARE_SWEEP,
     "internal_invariant", "gpu.batch.compact_path.edge.v1"},
    {"unauthenticated_cuda_producer", APGX_REPLAY_FAULT_UNAUTHENTICATED_PRODUCER,
     APGX_V1_GENERATOR_BUCKETED_FRONTIER, "unsupported", "candidate.builder.gpu_producer_authentication.v1"},
};

// The 34 keys in canonical order (tools/gpu_batch_replay.cc:131-166).
constexpr std::string_view kKeys[] = {
    "format=", "schema_version=", "device_candidate_batch_schema_version=", "fixture=",
    "fixture_checksum_fnv1a64=", "generator=", "fault=", "layer=", "maximum_rounds=",
    "maximum_workspace_states_per_query=", "maximum_frontier_states_per_query=", "maximum_device_bytes=",
    "maximum_host_bytes=", "batch_id=", "query_id=", "input_ordinal=", "policy_schema_version=", "objective=",
    "deterministic_seed=", "candidate_ordinal=", "orthogonal_step_surcharge=", "diagonal_step_surcharge=",
    "bend_surcharge=", "banned_resources=", "resource_penalties=", "expected_policy_identity=",
    "expected_board_content_hash=", "expected_compiler_profile_fingerprint=", "expected_compiler_version=",
    "expected_rule_bucket_identity=", "expected_routing_profile_fingerprint=", "expected_device_view_fingerprint=",
    "expected_failure=", "expected_invariant=",
};
constexpr size_t kKeyCount = sizeof(kKeys) / sizeof(kKeys[0]);

uint64_t Fnv1a(std::string_view bytes) {
  StableHash h;
  for (char c : bytes) h.Byte(static_cast<uint8_t>(c));
  return h.Finish();
}

bool ValidUtf8(std::string_view s) {
  size_t i = 0;
  while (i < s.size()) {
    const uint8_t c = static_cast<uint8_t>(s[i]);
    if (c < 0x80) {
      ++i;
      continue;
    }
    size_t n;
    uint32_t cp;
    if ((c & 0xE0) == 0xC0) {
      n = 1;
      cp = c & 0x1F;
    } else if ((c & 0xF0) == 0xE0) {
      n = 2;
      cp = c & 0x0F;
    } else if ((c & 0xF8) == 0xF0) {
      n = 3;
      cp = c & 0x07;
    } else {
      return false;
    }
    if (n >= s.size() - i) return false;  // the n continuation bytes must follow
    for (size_t k = 1; k <= n; ++k) {
      const uint8_t t = static_cast<uint8_t>(s[i + k]);
      if ((t & 0xC0) != 0x80) return false;
      cp = (cp << 6) | (t & 0x3F);
    }
    // Overlong forms, surrogates and values above U+10FFFF are invalid.
    if ((n == 1 && cp < 0x80) || (n == 2 && cp < 0x800) || (n == 3 && cp < 0x10000) || cp > 0x10FFFF ||
        (cp >= 0xD800 && cp <= 0xDFFF))
      return false;
    i += n + 1;
  }
  return true;
}

// Canonical unsigned decimal 0|[1-9][0-9]* within `max`.
bool ParseUnsigned(std::string_view text, uint64_t max, uint64_t* out) {
  if (text.empty() || (text.size() > 1 && text[0] == '0')) return false;
  uint64_t v = 0;
  for (char c : text) {
    if (c < '0' || c > '9') return false;
    const uint64_t d = static_cast<uint64_t>(c - '0');
    if (v > (max - d) / 10) return false;
    v = v * 10 + d;
  }
  *out = v;
  return true;
}

// ParseCanonicalReplayEnvelope (src/tooling/replay.cc:11-80) over one
// format's keys in canonical order.
bool ParseEnvelope(std::string_view contents, const std::string_view* keys, size_t key_count,
                   std::vector<std::string_view>* values, std::string* error) {
  if (!ValidUtf8(contents)) {
    *error = "canonical replay must be valid UTF-8";
    return false;
  }
  if (contents.find('\0') != std::string_view::npos || contents.find('\r') != std::string_view::npos) {
    *error = "canonical replay must not contain NUL or CR bytes";
    return false;
  }
  if (contents.empty() || contents.back() != '\n') {
    *error = "canonical replay must end every line with LF";
    return false;
  }
  constexpr std::string_view kChecksumKey = "checksum_fnv1a64=";
  const std::string_view body = contents.substr(0, contents.size() - 1);
  const size_t sep = body.rfind('\n');
  if (sep == std::string_view::npos || body.substr(sep + 1).substr(0, kChecksumKey.size()) != kChecksumKey) {
    *error = "missing final replay checksum";
    return false;
  }
  const std::string_view payload = contents.substr(0, sep + 1);
  uint64_t recorded = 0;
  if (!ParseUnsigned(body.substr(sep + 1 + kChecksumKey.size()), ~0ull, &recorded) || Fnv1a(payload) != recorded) {
    *error = "replay payload checksum mismatch";
    return false;
  }
  size_t offset = 0;
  values->clear();
  for (size_t k = 0; k < key_count; ++k) {
    const std::string_view key = keys[k];
    const size_t end = payload.find('\n', offset);
    if (end == std::string_view::npos) {
      *error = "replay payload is truncated";
      return false;
    }
    const std::string_view line = payload.substr(offset, end - offset);
    if (line.substr(0, key.size()) != key) {
      *error = "replay fields are missing or out of canonical order";
      return false;
    }
    values->push_back(line.substr(key.size()));
    offset = end + 1;
  }
  if (offset != payload.size()) {
    *error = "replay payload has unknown fields";
    return false;
  }
  return true;
}

void Copy(char* dst, size_t cap, std::string_view src) {
  const size_t n = std::min(cap - 1, src.size());
  std::memcpy(dst, src.data(), n);
  dst[n] = '\0';
}

// ParseArtifact (tools/gpu_batch_replay.cc:129-226).
bool Parse(std::string_view contents, apgx_batch_replay* a, std::string* error) {
  std::vector<std::string_view> v;
  if (!ParseEnvelope(contents, kKeys, kKeyCount, &v, error)) return false;
  const FaultContract* contract = nullptr;
  for (const FaultContract& c : kContracts)
    if (c.name == v[6]) contract = &c;
  const int generator = v[5] == "bucketed_frontier"     ? static_cast<int>(APGX_V1_GENERATOR_BUCKETED_FRONTIER)
                        : v[5] == "heading_aware_sweep" ? static_cast<int>(APGX_V1_GENERATOR_HEADING_AWARE_SWEEP)
                                                        : -1;
  if (v[0] != "apgar_gpu_batch_invariant_replay" || v[1] != "1" || v[2] != "1" || generator < 0 ||
      contract == nullptr || static_cast<uint32_t>(generator) != contract->generator || v[16] != "1" ||
      v[17] != "base_scalar_cost" || v[23] != "none" || v[24] != "none" || v[32] != contract->expected_failure ||
      v[33] != contract->invariant) {
    *error = std::string(kAssociationError);
    return false;
  }
  std::memset(a, 0, sizeof(*a));
  Copy(a->fixture, sizeof(a->fixture), v[3]);
  a->generator = static_cast<uint32_t>(generator);
  a->fault = contract->fault;
  a->expected_failure = contract->expected_failure == "unsupported" ? APGX_REPLAY_EXPECT_UNSUPPORTED
                                                                    : APGX_REPLAY_EXPECT_INTERNAL_INVARIANT;
  Copy(a->expected_invariant, sizeof(a->expected_invariant), v[33]);
  uint64_t u = 0;
  constexpr uint64_t k32 = 0xFFFFFFFFull, k64 = ~0ull;
  auto u32 = [&](size_t i, uint32_t* out) {
    if (!ParseUnsigned(v[i], k32, &u)) return false;
    *out = static_cast<uint32_t>(u);
    return true;
  };
  auto u64 = [&](size_t i, uint64_t* out) { return ParseUnsigned(v[i], k64, out); };
  uint32_t policy_schema = 0;
  if (!u64(4, &a->fixture_checksum) || !u32(7, &a->layer) || !u32(8, &a->maximum_rounds) ||
      a->maximum_rounds == 0 || !u64(9, &a->maximum_workspace_states_per_query) ||
      !u64(10, &a->maximum_frontier_states_per_query) || !u64(11, &a->maximum_device_bytes) ||
      !u64(12, &a->maximum_host_bytes) || !u64(13, &a->batch_id) || a->batch_id == 0 || !u64(14, &a->query_id) ||
      a->query_id == 0 || !u32(15, &a->input_ordinal) || !u32(16, &policy_schema) ||
      !u64(18, &a->deterministic_seed) || !u32(19, &a->candidate_ordinal) ||
      !u64(20, &a->orthogonal_step_surcharge) || !u64(21, &a->diagonal_step_surcharge) ||
      !u64(22, &a->bend_surcharge) || !u64(25, &a->expected_policy_identity) ||
      !u64(26, &a->expected_board_content_hash) || !u64(27, &a->expected_compiler_profile_fingerprint) ||
      !u32(28, &a->expected_compiler_version) || !u64(29, &a->expected_rule_bucket_identity) ||
      !u64(30, &a->expected_routing_profile_fingerprint) || !u64(31, &a->expected_device_view_fingerprint)) {
    *error = "replay contains an invalid integer field";
    return false;
  }
  a->policy_schema_version = policy_schema;
  return true;
}

// The 14 keys of the single-route artifact in canonical order
// (tools/gpu_replay_parser.cc:13-29).
constexpr std::string_view kRouteKeys[] = {
    "format=", "schema_version=", "fixture=", "fixture_checksum_fnv1a64=", "generator=", "fault=", "layer=",
    "maximum_rounds=", "expected_board_content_hash=", "expected_profile_fingerprint=",
    "expected_routing_profile_fingerprint=", "expected_device_view_fingerprint=", "expected_failure=",
    "expected_invariant=",
};

// ParseGpuReplayArtifactV1 (tools/gpu_replay_parser.cc:11-56): v1 has one
// contract, the frontier with the goal-predecessor self cycle reported as an
// internal invariant.
bool ParseRoute(std::string_view contents, apgx_route_replay* a, std::string* error) {
  std::vector<std::string_view> v;
  if (!ParseEnvelope(contents, kRouteKeys, sizeof(kRouteKeys) / sizeof(kRouteKeys[0]), &v, error)) return false;
  if (v[0] != "apgar_gpu_invariant_replay" || v[1] != "1" || v[4] != "bucketed_frontier" ||
      v[5] != "goal_predecessor_self_cycle" || v[12] != "internal_invariant" ||
      v[13] != "gpu.predecessor.self_reference.v1") {
    *error = "replay format, schema, generator, fault, or expected outcome is unsupported";
    return false;
  }
  std::memset(a, 0, sizeof(*a));
  Copy(a->fixture, sizeof(a->fixture), v[2]);
  Copy(a->expected_invariant, sizeof(a->expected_invariant), v[13]);
  a->generator = APGX_V1_GENERATOR_BUCKETED_FRONTIER;
  a->fault = APGX_V1_ROUTE_FAULT_GOAL_PREDECESSOR_SELF_CYCLE;
  uint64_t layer = 0, rounds = 0;
  constexpr uint64_t k32 = 0xFFFFFFFFull, k64 = ~0ull;
  if (!ParseUnsigned(v[3], k64, &a->fixture_checksum) || !ParseUnsigned(v[6], k32, &layer) ||
      !ParseUnsigned(v[7], k32, &rounds) || rounds == 0 || !ParseUnsigned(v[8], k64, &a->expected_board_content_hash) ||
      !ParseUnsigned(v[9], k64, &a->expected_profile_fingerprint) ||
      !ParseUnsigned(v[10], k64, &a->expected_routing_profile_fingerprint) ||
      !ParseUnsigned(v[11], k64, &a->expected_device_view_fingerprint)) {
    *error = "replay contains an invalid integer field";
    return false;
  }
  a->layer = static_cast<uint32_t>(layer);
  a->maximum_rounds = static_cast<uint32_t>(rounds);
  return true;
}

// A fixture compiled into its device view, with the request's endpoints.
struct Fixture {
  apgx_replay_fixture_header fh;
  apgx_v1_header header;
  std::vector<apgx_v1_layer_range> layers;
  std::vector<apgx_v1_node> nodes;
  std::vector<apgx_v1_run> runs;
  std::vector<uint32_t> run_nodes;
  uint32_t start = APGX_V1_INVALID_INDEX, goal = APGX_V1_INVALID_INDEX;
};

// Checksum, record structure and layer of a fixture, its compilation by
// apgx_v1_build_board and its endpoints.  Returns 0, or the stage at which it
// was rejected (the runner's exit code is then 2) with the message.
int CompileFixture(const uint8_t* fixture, uint64_t fixture_length, uint64_t checksum, uint32_t layer, Fixture* fx,
                   std::string* why) {
  const std::string_view bytes(reinterpret_cast<const char*>(fixture ? fixture : reinterpret_cast<const uint8_t*>("")),
                               fixture_length);
  if (fixture == nullptr || Fnv1a(bytes) != checksum) {
    *why = "fixture is missing or its checksum differs";
    return APGX_REPLAY_STAGE_FIXTURE;
  }
  apgx_replay_fixture_header& fh = fx->fh;
  if (fixture_length < sizeof(fh)) {
    *why = "fixture is truncated";
    return APGX_REPLAY_STAGE_FIXTURE;
  }
  std::memcpy(&fh, fixture, sizeof(fh));
  if (std::memcmp(fh.magic, APGX_REPLAY_FIXTURE_MAGIC, sizeof(fh.magic)) != 0 || fh.schema_version != 1 ||
      fh.node_count > (fixture_length - sizeof(fh)) / sizeof(apgx_v1_compiled_node) ||
      fixture_length != sizeof(fh) + fh.node_count * sizeof(apgx_v1_compiled_node)) {
    *why = "fixture is not an apgx V1 replay fixture";
    return APGX_REPLAY_STAGE_FIXTURE;
  }
  std::vector<apgx_v1_compiled_node> cnodes(fh.node_count);
  if (fh.node_count)
    std::memcpy(cnodes.data(), fixture + sizeof(fh), fh.node_count * sizeof(apgx_v1_compiled_node));
  if (fh.layer != layer) {
    *why = "replay layer differs from the versioned fixture case";
    return APGX_REPLAY_STAGE_FIXTURE;
  }
  // Device-view compilation (BuildDeviceCompiledBoardV1).
  apgx_v1_board* board = nullptr;
  if (apgx_v1_build_board(&fh.profile, cnodes.data(), fh.node_count, &board) != APGX_OK) {
    *why = std::string("device-view compilation failed: ") + apgx_last_error();
    return APGX_REPLAY_STAGE_FIXTURE;
  }
  uint64_t nl = 0, nn = 0, nr = 0, nrn = 0;
  apgx_v1_board_sizes(board, &nl, &nn, &nr, &nrn);
  fx->layers.resize(nl);
  fx->nodes.resize(nn);
  fx->runs.resize(nr);
  fx->run_nodes.resize(nrn);
  apgx_v1_board_copy(board, &fx->header, fx->layers.data(), fx->nodes.data(), fx->runs.data(), fx->run_nodes.data());
  apgx_v1_board_release(board);
  for (uint64_t i = 0; i < nn; ++i) {
    const apgx_v1_node& n = fx->nodes[i];
    if (n.layer != fh.layer) continue;
    if (n.lattice_x == fh.start_x && n.lattice_y == fh.start_y) fx->start = static_cast<uint32_t>(i);
    if (n.lattice_x == fh.goal_x && n.lattice_y == fh.goal_y) fx->goal = static_cast<uint32_t>(i);
  }
  if (fx->start == APGX_V1_INVALID_INDEX || fx->goal == APGX_V1_INVALID_INDEX || fx->start == fx->goal) {
    *why = "fixture endpoints are not distinct represented nodes";
    return APGX_REPLAY_STAGE_FIXTURE;
  }
  return 0;
}

}  // namespace

// FingerprintCandidateGenerationPolicy (src/routing/candidate_policy.cc:139-160)
// for the v1 replay policy: base_scalar_cost (objective byte 0) with empty ban
// and penalty sets.
uint64_t PolicyIdentity(const apgx_batch_replay& a) {
  StableHash h;
  h.String("APGAR-CANDIDATE-POLICY-V1");
  h.U32(a.policy_schema_version);
  h.Byte(0);
  h.U64(a.deterministic_seed);
  h.U32(a.candidate_ordinal);
  h.U64(a.orthogonal_step_surcharge);
  h.U64(a.diagonal_step_surcharge);
  h.U64(a.bend_surcharge);
  h.U64(0);
  h.U64(0);
  return h.Finish();
}

}  // namespace replay
}  // namespace apgx

extern "C" int apgx_batch_replay_parse(const char* text, uint64_t length, apgx_batch_replay* out, char* error,
                                       uint64_t error_capacity) {
  using namespace apgx;
  if (out == nullptr || (length && text == nullptr)) return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_batch_replay_parse: null argument");
  std::string why;
  const bool ok = replay::Parse(std::string_view(text ? text : "", length), out, &why);
  if (error && error_capacity) replay::Copy(error, error_capacity, ok ? std::string_view() : std::string_view(why));
  ClearError();
  return ok ? APGX_OK : APGX_E_UNSUPPORTED;
}

extern "C" int apgx_route_replay_parse(const char* text, uint64_t length, apgx_route_replay* out, char* error,
                                       uint64_t error_capacity) {
  using namespace apgx;
  if (out == nullptr || (length && text == nullptr)) return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_route_replay_parse: null argument");
  std::string why;
  const bool ok = replay::ParseRoute(std::string_view(text ? text : "", length), out, &why);
  if (error && error_capacity) replay::Copy(error, error_capacity, ok ? std::string_view() : std::string_view(why));
  ClearError();
  return ok ? APGX_OK : APGX_E_UNSUPPORTED;
}

extern "C" uint64_t apgx_batch_replay_policy_identity(const apgx_batch_replay* artifact) {
  return artifact ? apgx::replay::PolicyIdentity(*artifact) : 0;
}

namespace {

struct Outcome {
  apgx_replay_outcome* out;
  int finish(int exit_code, uint32_t stage, const std::string& message, const char* invariant = "") {
    out->exit_code = exit_code;
    out->stage = stage;
    apgx::replay::Copy(out->message, sizeof(out->message), message);
    apgx::replay::Copy(out->invariant, sizeof(out->invariant), invariant);
    apgx::ClearError();
    return APGX_OK;
  }
};

}  // namespace

extern "C" int apgx_batch_replay_run(const char* text, uint64_t text_length, const uint8_t* fixture,
                                     uint64_t fixture_length, uint32_t flags, apgx_replay_outcome* out) {
  using namespace apgx;
  if (out == nullptr || (text_length && text == nullptr) || (fixture_length && fixture == nullptr))
    return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_batch_replay_run: null argument");
  std::memset(out, 0, sizeof(*out));
  Outcome o{out};
  apgx_batch_replay a;
  std::string why;
  if (!replay::Parse(std::string_view(text ? text : "", text_length), &a, &why))
    return o.finish(2, APGX_REPLAY_STAGE_PARSE, why);
  replay::Fixture fx;
  if (int code = replay::CompileFixture(fixture, fixture_length, a.fixture_checksum, a.layer, &fx, &why))
    return o.finish(2, static_cast<uint32_t>(code), why);
  const apgx_replay_fixture_header& fh = fx.fh;
  const apgx_v1_header& header = fx.header;
  auto& layers = fx.layers;
  auto& nodes = fx.nodes;
  auto& runs = fx.runs;
  auto& run_nodes = fx.run_nodes;
  const uint64_t nl = layers.size(), nn = nodes.size(), nr = runs.size(), nrn = run_nodes.size();
  // Association gate, before any device work (tools/gpu_batch_replay.cc:271-287).
  const uint64_t policy_identity = replay::PolicyIdentity(a);
  if (header.source_board_content_hash != a.expected_board_content_hash ||
      header.compiler_profile_fingerprint != a.expected_compiler_profile_fingerprint ||
      header.compiler_version != a.expected_compiler_version ||
      header.rule_bucket_identity != a.expected_rule_bucket_identity ||
      fh.routing_profile_fingerprint != a.expected_routing_profile_fingerprint ||
      header.device_view_fingerprint != a.expected_device_view_fingerprint ||
      policy_identity != a.expected_policy_identity) {
    char msg[256];
    std::snprintf(msg, sizeof(msg),
                  "replay semantic associations differ: board=%llu compiler_profile=%llu compiler_version=%u "
                  "rule_bucket=%llu routing_profile=%llu device=%llu policy=%llu",
                  static_cast<unsigned long long>(header.source_board_content_hash),
                  static_cast<unsigned long long>(header.compiler_profile_fingerprint), header.compiler_version,
                  static_cast<unsigned long long>(header.rule_bucket_identity),
                  static_cast<unsigned long long>(fh.routing_profile_fingerprint),
                  static_cast<unsigned long long>(header.device_view_fingerprint),
                  static_cast<unsigned long long>(policy_identity));
    return o.finish(2, APGX_REPLAY_STAGE_ASSOCIATION, msg);
  }
  const uint32_t start = fx.start, goal = fx.goal;
  if (a.fault == APGX_REPLAY_FAULT_UNAUTHENTICATED_PRODUCER)
    return o.finish(2, APGX_REPLAY_STAGE_ASSOCIATION,
                    "candidate construction and producer authentication are outside this library");
  if (flags & APGX_REPLAY_DRY_RUN) return o.finish(0, APGX_REPLAY_STAGE_ASSOCIATION, "associations verified");

  // One immutable query under the artifact's policy and budgets.
  const uint64_t S = nn * 9;
  apgx_v1_query q;
  std::memset(&q, 0, sizeof(q));
  q.schema_version = APGX_V1_SCHEMA_VERSION;
  q.input_ordinal = a.input_ordinal;
  q.batch_id = a.batch_id;
  q.query_id = a.query_id;
  q.workspace_owner = a.query_id;
  q.policy_identity = policy_identity;
  q.routing_profile_fingerprint = fh.routing_profile_fingerprint;
  q.start_node = start;
  q.goal_node = goal;
  q.workspace_offset = 0;
  q.workspace_state_count = S;
  q.frontier_state_capacity = a.generator == APGX_V1_GENERATOR_BUCKETED_FRONTIER ? S : 0;
  q.orthogonal_step_surcharge = a.orthogonal_step_surcharge;
  q.diagonal_step_surcharge = a.diagonal_step_surcharge;
  q.bend_surcharge = a.bend_surcharge;
  q.maximum_rounds = a.maximum_rounds;
  q.generator = static_cast<uint8_t>(a.generator);
  apgx_v1_batch_request req{APGX_V1_SCHEMA_VERSION, a.generator, a.batch_id, a.maximum_rounds, 0,
                            a.maximum_workspace_states_per_query, a.maximum_frontier_states_per_query,
                            a.maximum_device_bytes, a.maximum_host_bytes};
  apgx_v1_view* view = nullptr;
  if (apgx_v1_view_upload(&header, layers.data(), nl, nodes.data(), nn, runs.data(), nr, run_nodes.data(), nrn,
                          &view) != APGX_OK)
    return o.finish(1, APGX_REPLAY_STAGE_EXECUTE, std::string("view upload failed: ") + apgx_last_error());
  apgx_v1_pending* pending = nullptr;
  const int rc = apgx_v1_batch_execute(view, &req, &q, 1, nullptr, 0, nullptr, &pending);
  apgx_v1_view_release(view);
  if (rc != APGX_OK)
    return o.finish(1, APGX_REPLAY_STAGE_EXECUTE, std::string("batch failed before query classification: ") +
                                                     apgx_last_error());
  apgx_v1_pending_inject_fault(pending, a.fault);
  apgx_v1_result result;
  apgx_v1_compact_path path;
  apgx_v1_batch_telemetry tel;
  uint64_t used = 0;
  apgx_v1_batch_readback(pending, &result, &path, nullptr, 0, &used, nullptr);
  std::vector<uint32_t> states(used);
  apgx_v1_batch_readback(pending, &result, &path, states.data(), used, &used, &tel);
  uint32_t kind = 0;
  uint64_t total = 0;
  apgx_v1_query_envelope env;
  apgx_v1_batch_readback_workspace(pending, &kind, &env, nullptr, nullptr, nullptr, nullptr, 0, &total);
  std::vector<uint64_t> labels(total), sown(total), pown(total);
  std::vector<uint32_t> preds(total);
  apgx_v1_batch_readback_workspace(pending, nullptr, nullptr, labels.data(), preds.data(), sown.data(), pown.data(),
                                   total, &total);
  apgx_v1_pending_release(pending);
  const apgx_v1_workspace_view ws{kind, 0, &env, labels.data(), total, preds.data(), total, sown.data(), total,
                                  pown.data(), total};
  apgx_v1_validation check;
  uint64_t cost = 0;
  apgx_v1_validate_batch(&header, nodes.data(), nn, &req, &q, 1, nullptr, 0, &result, &path, states.data(), used, &ws,
                         &tel, &cost, &check);
  const bool reproduced = !check.valid && check.failure == APGX_V1_CHECK_INTERNAL_INVARIANT &&
                          std::string_view(check.invariant) == a.expected_invariant;
  if (!reproduced) {
    std::string msg = check.valid ? "the readback validated; expected GPU batch invariant was not reproduced"
                                  : std::string("expected GPU batch invariant was not reproduced: ") + check.detail;
    return o.finish(1, APGX_REPLAY_STAGE_REPRODUCE, msg, check.invariant);
  }
  apgx_device_info info;
  std::memset(&info, 0, sizeof(info));
  apgx_device_info_query(&info);
  char msg[256];
  std::snprintf(msg, sizeof(msg),
                "reproduced=internal_invariant invariant=%s artifact_schema=1 backend=%s device=\"%s\" arch=%s",
                a.expected_invariant, info.backend, info.device_name, info.gcn_arch_name);
  return o.finish(0, APGX_REPLAY_STAGE_REPRODUCE, msg, check.invariant);
}

extern "C" int apgx_route_replay_run(const char* text, uint64_t text_length, const uint8_t* fixture,
                                     uint64_t fixture_length, uint32_t flags, apgx_replay_outcome* out) {
  using namespace apgx;
  if (out == nullptr || (text_length && text == nullptr) || (fixture_length && fixture == nullptr))
    return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_route_replay_run: null argument");
  std::memset(out, 0, sizeof(*out));
  Outcome o{out};
  apgx_route_replay a;
  std::string why;
  if (!replay::ParseRoute(std::string_view(text ? text : "", text_length), &a, &why))
    return o.finish(2, APGX_REPLAY_STAGE_PARSE, why);
  replay::Fixture fx;
  if (int code = replay::CompileFixture(fixture, fixture_length, a.fixture_checksum, a.layer, &fx, &why))
    return o.finish(2, static_cast<uint32_t>(code), why);
  // Association gate (tools/gpu_replay_runner.cc:78-89), before any device work.
  const bool board = fx.header.source_board_content_hash != a.expected_board_content_hash;
  const bool profile = fx.header.compiler_profile_fingerprint != a.expected_profile_fingerprint;
  const bool routing = fx.fh.routing_profile_fingerprint != a.expected_routing_profile_fingerprint;
  const bool device = fx.header.device_view_fingerprint != a.expected_device_view_fingerprint;
  if (board || profile || routing || device) {
    std::string msg = "replay semantic association mismatch:";
    if (board) msg += " board";
    if (profile) msg += " compiler_profile";
    if (routing) msg += " routing_profile";
    if (device) msg += " device_view";
    return o.finish(2, APGX_REPLAY_STAGE_ASSOCIATION, msg);
  }
  if (flags & APGX_REPLAY_DRY_RUN) return o.finish(0, APGX_REPLAY_STAGE_ASSOCIATION, "associations verified");
  apgx_v1_view* view = nullptr;
  if (apgx_v1_view_upload(&fx.header, fx.layers.data(), fx.layers.size(), fx.nodes.data(), fx.nodes.size(),
                          fx.runs.data(), fx.runs.size(), fx.run_nodes.data(), fx.run_nodes.size(), &view) != APGX_OK)
    return o.finish(1, APGX_REPLAY_STAGE_EXECUTE, std::string("view upload failed: ") + apgx_last_error());
  const apgx_v1_route_request req{a.generator, fx.start, fx.goal, a.maximum_rounds, fx.fh.routing_profile_fingerprint,
                                  ~0ull};
  apgx_v1_route_pending* pending = nullptr;
  if (apgx_v1_route_execute(view, &req, nullptr, &pending) != APGX_OK) {
    const std::string err = apgx_last_error();
    apgx_v1_view_release(view);
    return o.finish(1, APGX_REPLAY_STAGE_EXECUTE, "route failed before validation: " + err);
  }
  apgx_v1_route_pending_inject_fault(pending, a.fault);
  const uint64_t S = fx.nodes.size() * 9;
  std::vector<uint64_t> labels(S);
  std::vector<uint32_t> preds(S);
  apgx_v1_route_result result;
  const int rc = apgx_v1_route_readback(pending, &result, labels.data(), preds.data(), S);
  apgx_v1_route_pending_release(pending);
  apgx_v1_view_release(view);
  if (rc != APGX_OK)
    return o.finish(1, APGX_REPLAY_STAGE_EXECUTE, std::string("route readback failed: ") + apgx_last_error());
  apgx_v1_validation check;
  uint64_t cost = 0;
  apgx_v1_validate_route(&fx.header, fx.nodes.data(), fx.nodes.size(), &req, &result, labels.data(), preds.data(),
                         &cost, &check);
  const bool reproduced = !check.valid && check.failure == APGX_V1_CHECK_INTERNAL_INVARIANT &&
                          std::string_view(check.invariant) == a.expected_invariant;
  if (!reproduced) return o.finish(1, APGX_REPLAY_STAGE_REPRODUCE, "expected GPU invariant failure was not reproduced",
                                   check.invariant);
  apgx_device_info info;
  std::memset(&info, 0, sizeof(info));
  apgx_device_info_query(&info);
  char msg[256];
  std::snprintf(msg, sizeof(msg),
                "reproduced=internal_invariant invariant=%s artifact_schema=1 backend=%s device=\"%s\" arch=%s",
                a.expected_invariant, info.backend, info.device_name, info.gcn_arch_name);
  return o.finish(0, APGX_REPLAY_STAGE_REPRODUCE, msg, check.invariant);
}
