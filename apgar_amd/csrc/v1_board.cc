// v1_board.cc — host side of the V1 device view: flattening of compiled
// sparse nodes into DeviceCompiledBoardV1 records and its FNV-1a fingerprint.
#include <algorithm>
#include <cstring>
#include <limits>
#include <string>
#include <tuple>
#include <unordered_map>
#include <vector>

#include "apgx.h"
#include "common.h"
#include "v1_common.h"

static_assert(sizeof(apgx_v1_header) == 112, "DeviceCompiledHeaderV1 layout");
static_assert(sizeof(apgx_v1_layer_range) == 48, "DeviceLayerRangeV1 layout");
static_assert(sizeof(apgx_v1_node) == 56, "DeviceNodeV1 layout");
static_assert(sizeof(apgx_v1_run) == 12, "DeviceRunV1 layout");
static_assert(sizeof(apgx_v1_policy_edge) == 16, "DeviceCandidatePolicyEdgeV1 layout");
static_assert(sizeof(apgx_v1_query) == 120, "DeviceCandidateBatchQueryV1 layout");
static_assert(sizeof(apgx_v1_result) == 120, "DeviceCandidateBatchResultV1 layout");
static_assert(sizeof(apgx_v1_compact_path) == 32, "DeviceCandidateCompactPathV1 layout");
static_assert(offsetof(apgx_v1_header, heading_mask) == 100, "header field offsets");
static_assert(offsetof(apgx_v1_query, maximum_rounds) == 112, "query field offsets");
static_assert(offsetof(apgx_v1_result, batch_id) == 32, "result field offsets");

struct apgx_v1_board {
  apgx_v1_header header{};
  std::vector<apgx_v1_layer_range> layers;
  std::vector<apgx_v1_node> nodes;
  std::vector<apgx_v1_run> runs;
  std::vector<uint32_t> run_nodes;
};

namespace apgx {
namespace v1 {

uint64_t Fingerprint(const apgx_v1_header& h, const apgx_v1_layer_range* layers, uint64_t nl,
                     const apgx_v1_node* nodes, uint64_t nn, const apgx_v1_run* runs, uint64_t nr,
                     const uint32_t* run_nodes, uint64_t nrn) {
  // Field order of ComputeDeviceCompiledBoardFingerprintV1 (src/gpu/device_board.cc:82-134).
  StableHash s;
  s.String("APGAR-DEVICE-COMPILED-BOARD-V1");
  s.U32(h.schema_version);
  s.U32(h.compiler_version);
  s.U64(h.source_board_content_hash);
  s.U64(h.compiler_profile_fingerprint);
  s.U64(h.rule_bucket_identity);
  s.U64(h.represented_nodes);
  s.U64(h.represented_states);
  s.U64(h.estimated_persistent_device_bytes);
  s.I64(h.lattice_origin_x);
  s.I64(h.lattice_origin_y);
  s.I64(h.lattice_step);
  s.U32(h.orthogonal_step);
  s.U32(h.diagonal_step);
  s.U32(h.bend);
  s.Byte(h.heading_mask);
  s.U64(nl);
  for (uint64_t i = 0; i < nl; ++i) {
    s.U32(layers[i].layer);
    s.U32(layers[i].node_offset);
    s.U32(layers[i].node_count);
    s.I64(layers[i].minimum_lattice_x);
    s.I64(layers[i].maximum_lattice_x);
    s.I64(layers[i].minimum_lattice_y);
    s.I64(layers[i].maximum_lattice_y);
  }
  s.U64(nn);
  for (uint64_t i = 0; i < nn; ++i) {
    s.I64(nodes[i].lattice_x);
    s.I64(nodes[i].lattice_y);
    s.U32(nodes[i].layer);
    s.Byte(nodes[i].legal_edges);
    for (uint32_t nb : nodes[i].neighbors) s.U32(nb);
  }
  s.U64(nr);
  for (uint64_t i = 0; i < nr; ++i) {
    s.U32(runs[i].node_offset);
    s.U32(runs[i].node_count);
    s.Byte(runs[i].direction);
  }
  s.U64(nrn);
  for (uint64_t i = 0; i < nrn; ++i) s.U32(run_nodes[i]);
  return s.Finish();
}

uint64_t PersistentBytes(uint64_t nl, uint64_t nn, uint64_t nr, uint64_t nrn) {
  return sizeof(apgx_v1_header) + nl * sizeof(apgx_v1_layer_range) + nn * sizeof(apgx_v1_node) +
         nr * sizeof(apgx_v1_run) + nrn * sizeof(uint32_t);
}

namespace {

int64_t FloorDiv(int64_t a, int64_t b) {
  int64_t q = a / b;
  if (a % b < 0) --q;
  return q;
}

struct Key {
  uint32_t layer;
  int64_t x, y;
  bool operator==(const Key& o) const { return layer == o.layer && x == o.x && y == o.y; }
};
struct KeyHash {
  size_t operator()(const Key& k) const {
    StableHash h;
    h.U32(k.layer);
    h.I64(k.x);
    h.I64(k.y);
    return static_cast<size_t>(h.Finish());
  }
};

}  // namespace

// Flattening of BuildDeviceCompiledBoardV1: nodes in sparse-tile order (tile key
// (layer, tile x, tile y), then row-major local index), neighbour ids per
// direction, per-layer ranges, maximal runs ordered by (direction, start node).
int Build(const apgx_v1_compiled_profile& p, const apgx_v1_compiled_node* in, uint64_t n,
          apgx_v1_board* b) {
  if (n == 0 || n > kMaximumRepresentedNodes)
    return Fail(APGX_E_INTERNAL_INVARIANT, "represented-node count is outside schema v1 bounds");
  if (static_cast<unsigned __int128>(n) * 9 >= APGX_V1_INVALID_INDEX)
    return Fail(APGX_E_RESOURCE_EXHAUSTED, "flattened heading-state count exceeds the uint32 index space");
  if (p.tile_width_nodes == 0 || p.tile_height_nodes == 0 || p.lattice_step <= 0)
    return Fail(APGX_E_INTERNAL_INVARIANT, "compiled profile tile or step is not positive");
  struct Ordered {
    uint32_t layer;
    int64_t tx, ty;
    uint64_t local;
    uint64_t input;
  };
  std::vector<Ordered> ord(n);
  const int64_t tw = p.tile_width_nodes, th = p.tile_height_nodes;
  for (uint64_t i = 0; i < n; ++i) {
    const int64_t tx = FloorDiv(in[i].lattice_x, tw), ty = FloorDiv(in[i].lattice_y, th);
    const uint64_t lx = static_cast<uint64_t>(in[i].lattice_x - tx * tw);
    const uint64_t ly = static_cast<uint64_t>(in[i].lattice_y - ty * th);
    ord[i] = Ordered{in[i].layer, tx, ty, ly * static_cast<uint64_t>(tw) + lx, i};
  }
  std::sort(ord.begin(), ord.end(), [](const Ordered& a, const Ordered& c) {
    return std::tie(a.layer, a.tx, a.ty, a.local) < std::tie(c.layer, c.tx, c.ty, c.local);
  });
  b->nodes.resize(n);
  std::unordered_map<Key, uint32_t, KeyHash> index;
  index.reserve(n * 2);
  for (uint64_t i = 0; i < n; ++i) {
    const apgx_v1_compiled_node& c = in[ord[i].input];
    apgx_v1_node& d = b->nodes[i];
    std::memset(&d, 0, sizeof(d));
    d.lattice_x = c.lattice_x;
    d.lattice_y = c.lattice_y;
    d.layer = c.layer;
    d.legal_edges = c.legal_edges;
    for (uint32_t& nb : d.neighbors) nb = APGX_V1_INVALID_INDEX;
    if (!index.emplace(Key{c.layer, c.lattice_x, c.lattice_y}, static_cast<uint32_t>(i)).second)
      return Fail(APGX_E_INTERNAL_INVARIANT, "compiled sparse tiles contain a duplicate global node identity");
  }
  for (uint64_t i = 0; i < n; ++i) {
    apgx_v1_node& d = b->nodes[i];
    for (int dir = 0; dir < 8; ++dir) {
      const bool legal = (d.legal_edges >> dir) & 1u;
      if ((p.heading_mask & HeadingFor(dir)) == 0) {
        if (legal) return Fail(APGX_E_INTERNAL_INVARIANT, "compiled mask enables a direction excluded by its profile");
        continue;
      }
      if (!legal) continue;
      auto it = index.find(Key{d.layer, d.lattice_x + kDx[dir], d.lattice_y + kDy[dir]});
      if (it == index.end())
        return Fail(APGX_E_INTERNAL_INVARIANT, "compiled legal edge exits the represented sparse field");
      d.neighbors[dir] = it->second;
    }
  }
  for (uint64_t i = 0; i < n; ++i) {
    const apgx_v1_node& d = b->nodes[i];
    for (int dir = 0; dir < 8; ++dir) {
      if (!((d.legal_edges >> dir) & 1u)) continue;
      const apgx_v1_node& t = b->nodes[d.neighbors[dir]];
      const int opp = (dir + 4) % 8;
      if (!((t.legal_edges >> opp) & 1u) || t.neighbors[opp] != i)
        return Fail(APGX_E_INTERNAL_INVARIANT, "flattened directional edge lacks its compiled reverse edge");
    }
  }
  for (uint64_t begin = 0; begin < n;) {
    const uint32_t layer = b->nodes[begin].layer;
    apgx_v1_layer_range r{};
    r.layer = layer;
    r.node_offset = static_cast<uint32_t>(begin);
    r.minimum_lattice_x = r.maximum_lattice_x = b->nodes[begin].lattice_x;
    r.minimum_lattice_y = r.maximum_lattice_y = b->nodes[begin].lattice_y;
    uint64_t end = begin;
    while (end < n && b->nodes[end].layer == layer) {
      r.minimum_lattice_x = std::min(r.minimum_lattice_x, b->nodes[end].lattice_x);
      r.maximum_lattice_x = std::max(r.maximum_lattice_x, b->nodes[end].lattice_x);
      r.minimum_lattice_y = std::min(r.minimum_lattice_y, b->nodes[end].lattice_y);
      r.maximum_lattice_y = std::max(r.maximum_lattice_y, b->nodes[end].lattice_y);
      ++end;
    }
    r.node_count = static_cast<uint32_t>(end - begin);
    b->layers.push_back(r);
    begin = end;
  }
  for (int dir = 0; dir < 8; ++dir) {
    const int opp = (dir + 4) % 8;
    for (uint64_t start = 0; start < n; ++start) {
      const apgx_v1_node& s = b->nodes[start];
      if (!((s.legal_edges >> dir) & 1u)) continue;
      const uint32_t prev = s.neighbors[opp];
      if (prev != APGX_V1_INVALID_INDEX && ((b->nodes[prev].legal_edges >> dir) & 1u)) continue;
      const uint64_t offset = b->run_nodes.size();
      uint32_t cursor = static_cast<uint32_t>(start);
      b->run_nodes.push_back(cursor);
      while ((b->nodes[cursor].legal_edges >> dir) & 1u) {
        const uint32_t next = b->nodes[cursor].neighbors[dir];
        if (next == APGX_V1_INVALID_INDEX || b->run_nodes.size() - offset > n)
          return Fail(APGX_E_INTERNAL_INVARIANT, "directional run reconstruction encountered an invalid edge or cycle");
        b->run_nodes.push_back(next);
        cursor = next;
      }
      const uint64_t count = b->run_nodes.size() - offset;
      if (offset >= APGX_V1_INVALID_INDEX || count > std::numeric_limits<uint32_t>::max())
        return Fail(APGX_E_RESOURCE_EXHAUSTED, "directional run storage exceeds schema v1 index bounds");
      apgx_v1_run run{};
      run.node_offset = static_cast<uint32_t>(offset);
      run.node_count = static_cast<uint32_t>(count);
      run.direction = static_cast<uint8_t>(dir);
      b->runs.push_back(run);
    }
  }
  apgx_v1_header& h = b->header;
  std::memset(&h, 0, sizeof(h));
  h.schema_version = APGX_V1_SCHEMA_VERSION;
  h.compiler_version = p.compiler_version;
  h.source_board_content_hash = p.source_board_content_hash;
  h.compiler_profile_fingerprint = p.compiler_profile_fingerprint;
  h.rule_bucket_identity = p.rule_bucket_identity;
  h.represented_nodes = n;
  h.represented_states = n * 9;
  h.lattice_origin_x = p.lattice_origin_x;
  h.lattice_origin_y = p.lattice_origin_y;
  h.lattice_step = p.lattice_step;
  h.orthogonal_step = p.orthogonal_step;
  h.diagonal_step = p.diagonal_step;
  h.bend = p.bend;
  h.heading_mask = static_cast<uint8_t>(p.heading_mask);
  h.estimated_persistent_device_bytes =
      PersistentBytes(b->layers.size(), b->nodes.size(), b->runs.size(), b->run_nodes.size());
  h.device_view_fingerprint = Fingerprint(h, b->layers.data(), b->layers.size(), b->nodes.data(),
                                          b->nodes.size(), b->runs.data(), b->runs.size(),
                                          b->run_nodes.data(), b->run_nodes.size());
  return APGX_OK;
}

}  // namespace v1
}  // namespace apgx

extern "C" int apgx_v1_build_board(const apgx_v1_compiled_profile* profile,
                                   const apgx_v1_compiled_node* nodes, uint64_t node_count,
                                   apgx_v1_board** out) {
  using namespace apgx;
  if (profile == nullptr || out == nullptr || (node_count && nodes == nullptr))
    return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_v1_build_board: null argument");
  *out = nullptr;
  auto* b = new (std::nothrow) apgx_v1_board();
  if (b == nullptr) return Fail(APGX_E_RESOURCE_EXHAUSTED, "apgx_v1_build_board: host allocation");
  if (int s = v1::Build(*profile, nodes, node_count, b)) {
    delete b;
    return s;
  }
  *out = b;
  ClearError();
  return APGX_OK;
}

extern "C" int apgx_v1_board_sizes(const apgx_v1_board* b, uint64_t* nl, uint64_t* nn, uint64_t* nr,
                                   uint64_t* nrn) {
  if (b == nullptr || !nl || !nn || !nr || !nrn)
    return apgx::Fail(APGX_E_INTERNAL_INVARIANT, "apgx_v1_board_sizes: null argument");
  *nl = b->layers.size();
  *nn = b->nodes.size();
  *nr = b->runs.size();
  *nrn = b->run_nodes.size();
  apgx::ClearError();
  return APGX_OK;
}

extern "C" int apgx_v1_board_copy(const apgx_v1_board* b, apgx_v1_header* header,
                                  apgx_v1_layer_range* layers, apgx_v1_node* nodes,
                                  apgx_v1_run* runs, uint32_t* run_nodes) {
  if (b == nullptr || header == nullptr)
    return apgx::Fail(APGX_E_INTERNAL_INVARIANT, "apgx_v1_board_copy: null argument");
  *header = b->header;
  if (layers) std::memcpy(layers, b->layers.data(), b->layers.size() * sizeof(apgx_v1_layer_range));
  if (nodes) std::memcpy(nodes, b->nodes.data(), b->nodes.size() * sizeof(apgx_v1_node));
  if (runs) std::memcpy(runs, b->runs.data(), b->runs.size() * sizeof(apgx_v1_run));
  if (run_nodes) std::memcpy(run_nodes, b->run_nodes.data(), b->run_nodes.size() * 4);
  apgx::ClearError();
  return APGX_OK;
}

extern "C" void apgx_v1_board_release(apgx_v1_board* b) { delete b; }

extern "C" uint64_t apgx_v1_view_fingerprint(const apgx_v1_header* header,
                                             const apgx_v1_layer_range* layers, uint64_t nl,
                                             const apgx_v1_node* nodes, uint64_t nn,
                                             const apgx_v1_run* runs, uint64_t nr,
                                             const uint32_t* run_nodes, uint64_t nrn) {
  if (header == nullptr) return 0;
  return apgx::v1::Fingerprint(*header, layers, nl, nodes, nn, runs, nr, run_nodes, nrn);
}
