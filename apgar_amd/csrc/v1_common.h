// v1_common.h — shared definitions of the V1 (reference record) path.
#pragma once

#include <cstdint>
#include <string_view>

#include "apgx.h"

namespace apgx {

// Byte-stable FNV-1a with little-endian field encoding and length-prefixed
// strings: the hash every reference fingerprint and replay checksum uses
// (include/apgar/board_ir/stable_hash.h:12-53).
class StableHash {
 public:
  void Byte(uint8_t v) {
    value_ ^= v;
    value_ *= 1099511628211ull;
  }
  void U32(uint32_t v) {
    for (int i = 0; i < 4; ++i) Byte(static_cast<uint8_t>(v >> (8 * i)));
  }
  void U64(uint64_t v) {
    for (int i = 0; i < 8; ++i) Byte(static_cast<uint8_t>(v >> (8 * i)));
  }
  void I64(int64_t v) { U64(static_cast<uint64_t>(v)); }
  void String(std::string_view s) {
    U64(s.size());
    for (char c : s) Byte(static_cast<uint8_t>(c));
  }
  uint64_t Finish() const { return value_; }

 private:
  uint64_t value_ = 14695981039346656037ull;
};

namespace v1 {

constexpr uint64_t kMaximumRepresentedNodes = 10'000'000;  // compiled_board.h:21
constexpr int kDx[8] = {1, 1, 0, -1, -1, -1, 0, 1};
constexpr int kDy[8] = {0, 1, 1, 1, 0, -1, -1, -1};

// Heading class of a direction: 1 horizontal, 2 vertical, 4 diagonal.
inline uint32_t HeadingFor(int dir) { return kDy[dir] == 0 ? 1u : (kDx[dir] == 0 ? 2u : 4u); }

uint64_t Fingerprint(const apgx_v1_header& h, const apgx_v1_layer_range* layers, uint64_t nl,
                     const apgx_v1_node* nodes, uint64_t nn, const apgx_v1_run* runs, uint64_t nr,
                     const uint32_t* run_nodes, uint64_t nrn);
uint64_t PersistentBytes(uint64_t nl, uint64_t nn, uint64_t nr, uint64_t nrn);

// Execution and telemetry shape of each generator as this backend runs it
// (the role of DescribePlanarGenerator, include/apgar/gpu/planar_router.h:71-99):
// both generators launch once per chunk of rounds; fixed launches are init +
// predecessor selection (frontier) or init + 3 compact-path kernels (sweep).
struct GeneratorShape {
  uint32_t chunk_rounds;
  uint32_t fixed_launches;
  uint32_t fixed_examined_work_per_round;
  uint32_t examined_work_per_node_per_round;
  uint32_t turn_relaxations_per_examined_work;
  bool full_workspace;
};

// Chunk sizes are kCandidateFrontierChunkRounds = 32 and
// kCandidateSweepChunkRounds = 8 (include/apgar/gpu/planar_router.h:36-37).
inline bool DescribeGenerator(uint32_t generator, GeneratorShape* out) {
  if (generator == APGX_V1_GENERATOR_BUCKETED_FRONTIER) {
    *out = GeneratorShape{32, 2, 8, 0, 0, true};
    return true;
  }
  if (generator == APGX_V1_GENERATOR_HEADING_AWARE_SWEEP) {
    *out = GeneratorShape{8, 4, 0, 8, 8, false};
    return true;
  }
  return false;
}

// Section offsets of a batch's single device block (256-byte aligned).
// Common: queries, policy edges, results, labels.  Sweep: compact path
// headers, packed path states, departures.  Frontier: predecessors, state and
// predecessor owners, closed flags.
struct BatchLayout {
  uint64_t queries, policy, results, labels;
  uint64_t paths, packed, departures;
  uint64_t predecessors, state_owners, predecessor_owners, closed;
  uint64_t total;
};

inline BatchLayout LayoutBatch(bool frontier, uint64_t query_count, uint64_t policy_count, uint64_t node_count) {
  auto align = [](uint64_t v) { return (v + 255) & ~255ull; };
  const uint64_t states = query_count * node_count * 9;
  BatchLayout l{};
  l.queries = 0;
  l.policy = align(query_count * sizeof(apgx_v1_query));
  l.results = align(l.policy + (policy_count ? policy_count : 1) * sizeof(apgx_v1_policy_edge));
  l.labels = align(l.results + query_count * sizeof(apgx_v1_result));
  const uint64_t tail = align(l.labels + states * 8);
  if (frontier) {
    l.predecessors = tail;
    l.state_owners = align(l.predecessors + states * 4);
    l.predecessor_owners = align(l.state_owners + states * 8);
    l.closed = align(l.predecessor_owners + states * 8);
    l.total = l.closed + states;
  } else {
    l.paths = tail;
    l.packed = align(l.paths + query_count * sizeof(apgx_v1_compact_path));
    l.departures = align(l.packed + states * 4);
    l.total = l.departures + query_count * node_count * 8 * 8;
  }
  return l;
}

// Bytes of one state in the full-workspace readback: label u64, predecessor
// u32, state owner u64, predecessor owner u64 (src/gpu/planar_router.cc:1918-1920).
constexpr uint64_t kFullWorkspaceBytesPerState = 8 + 4 + 2 * 8;

// Readback and host bytes of a batch (deterministic accounting).
inline uint64_t ReadbackBytes(bool full_workspace, uint64_t query_count, uint64_t node_count,
                              uint64_t compact_states) {
  return full_workspace ? query_count * sizeof(apgx_v1_result) +
                              query_count * node_count * 9 * kFullWorkspaceBytesPerState
                        : query_count * (sizeof(apgx_v1_result) + sizeof(apgx_v1_compact_path)) +
                              compact_states * 4;
}
inline uint64_t HostBytes(bool full_workspace, uint64_t query_count, uint64_t node_count, uint64_t compact_states) {
  return ReadbackBytes(full_workspace, query_count, node_count, compact_states) + query_count * sizeof(apgx_v1_query);
}

}  // namespace v1
}  // namespace apgx
