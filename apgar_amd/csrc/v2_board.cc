// v2_board.cc — the DeviceDenseBoardV2 record: a sealed, fingerprinted dense
// board (header + cost plane) for the V2 model of multi-layer lattices, vias,
// windows and multi-pin nets, which the reference's V1 records cannot express
// (planar heading states, uniform step costs, two terminals).  It plays the
// role DeviceCompiledBoardV1 plays for V1 (src/gpu/device_board.cc:82-325):
// the producer seals the board with its board-level associations, and the
// consumer re-derives every derived field and the fingerprint before upload.
// V1 records are untouched by it.  The byte layout is fixed little endian and
// pinned by the static_asserts below.
#include <cstddef>
#include <cstring>
#include <string>

#include "apgx.h"
#include "common.h"
#include "v1_common.h"

static_assert(sizeof(apgx_v2_board_header) == 128, "DeviceDenseBoardV2 header is 128 bytes");
static_assert(offsetof(apgx_v2_board_header, device_view_fingerprint) == 32, "header layout");
static_assert(offsetof(apgx_v2_board_header, width) == 72, "header layout");
static_assert(offsetof(apgx_v2_board_header, free_cells) == 104, "header layout");
static_assert(sizeof(apgx_net_query) == 32 && offsetof(apgx_net_query, window) == 16, "net query layout");
static_assert(sizeof(apgx_net_result) == 32, "net result layout");
static_assert(sizeof(apgx_connection) == 24 && offsetof(apgx_connection, cell_offset) == 8, "connection layout");

namespace apgx {
namespace v2 {
namespace {

struct Derived {
  uint32_t min_cost = 0, max_cost = 0;
  uint64_t free_cells = 0;
};

uint64_t Cells(const apgx_v2_board_header& h) {
  return static_cast<uint64_t>(h.width) * h.height * h.layers;
}

int CheckShape(const apgx_v2_board_header& h, uint64_t plane_bytes) {
  if (h.schema_version != APGX_LATTICE_SCHEMA_VERSION)
    return Fail(APGX_E_UNSUPPORTED, "dense board schema version is not supported");
  apgx_lattice_desc d{APGX_LATTICE_SCHEMA_VERSION, h.width, h.height, h.layers, h.cost_bytes, h.via_cost, {0, 0}};
  if (d.width == 0 || d.height == 0 || d.layers == 0 || d.layers > APGX_MAX_LAYERS)
    return Fail(APGX_E_INTERNAL_INVARIANT, "dense board rejected: dimensions out of range");
  if (d.cost_bytes != 2 && d.cost_bytes != 4)
    return Fail(APGX_E_INTERNAL_INVARIANT, "dense board rejected: cost_bytes must be 2 or 4");
  if (d.via_cost > APGX_MAX_VIA_COST) return Fail(APGX_E_INTERNAL_INVARIANT, "dense board rejected: via cost");
  const unsigned __int128 cells = static_cast<unsigned __int128>(h.width) * h.height * h.layers;
  if (cells >= 0xFFFFFFFFull)
    return Fail(APGX_E_RESOURCE_EXHAUSTED, "dense board rejected: cell count exceeds the uint32 index space");
  if (plane_bytes != Cells(h) * h.cost_bytes)
    return Fail(APGX_E_INTERNAL_INVARIANT, "dense board rejected: plane size disagrees with the header");
  if (h.lattice_step <= 0) return Fail(APGX_E_INTERNAL_INVARIANT, "dense board rejected: lattice step");
  for (uint8_t b : h.reserved)
    if (b != 0) return Fail(APGX_E_INTERNAL_INVARIANT, "dense board rejected: reserved bytes");
  if (h.reserved0 != 0) return Fail(APGX_E_INTERNAL_INVARIANT, "dense board rejected: reserved bytes");
  return APGX_OK;
}

int Derive(const apgx_v2_board_header& h, const void* plane, Derived* out) {
  const uint64_t cells = Cells(h);
  uint32_t mn = 0xFFFFFFFFu, mx = 0;
  uint64_t free_cells = 0;
  for (uint64_t c = 0; c < cells; ++c) {
    const uint32_t v = h.cost_bytes == 2 ? static_cast<const uint16_t*>(plane)[c] : static_cast<const uint32_t*>(plane)[c];
    if (v == 0) continue;
    if (v > APGX_MAX_CELL_COST)
      return Fail(APGX_E_INTERNAL_INVARIANT, "dense board rejected: cell cost exceeds APGX_MAX_CELL_COST");
    ++free_cells;
    mn = v < mn ? v : mn;
    mx = v > mx ? v : mx;
  }
  out->min_cost = free_cells ? mn : 0;
  out->max_cost = mx;
  out->free_cells = free_cells;
  return APGX_OK;
}

uint64_t Fingerprint(const apgx_v2_board_header& h, const void* plane) {
  StableHash s;
  s.String("APGAR-DEVICE-DENSE-BOARD-V2");
  s.U32(h.schema_version);
  s.U32(h.compiler_version);
  s.U64(h.source_board_content_hash);
  s.U64(h.compiler_profile_fingerprint);
  s.U64(h.rule_bucket_identity);
  s.U64(h.estimated_persistent_device_bytes);
  s.I64(h.lattice_origin_x);
  s.I64(h.lattice_origin_y);
  s.I64(h.lattice_step);
  s.U32(h.width);
  s.U32(h.height);
  s.U32(h.layers);
  s.U32(h.cost_bytes);
  s.U32(h.via_cost);
  s.U32(h.min_cell_cost);
  s.U32(h.max_cell_cost);
  s.U64(h.free_cells);
  const uint64_t bytes = Cells(h) * h.cost_bytes;
  const uint8_t* p = static_cast<const uint8_t*>(plane);
  for (uint64_t i = 0; i < bytes; ++i) s.Byte(p[i]);
  return s.Finish();
}

}  // namespace
}  // namespace v2
}  // namespace apgx

extern "C" uint64_t apgx_v2_board_fingerprint(const apgx_v2_board_header* header, const void* plane) {
  if (header == nullptr || plane == nullptr) return 0;
  return apgx::v2::Fingerprint(*header, plane);
}

extern "C" int apgx_v2_board_seal(apgx_v2_board_header* header, const void* plane, uint64_t plane_bytes) {
  using namespace apgx;
  if (header == nullptr || plane == nullptr) return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_v2_board_seal: null argument");
  if (int s = v2::CheckShape(*header, plane_bytes)) return s;
  v2::Derived d;
  if (int s = v2::Derive(*header, plane, &d)) return s;
  header->min_cell_cost = d.min_cost;
  header->max_cell_cost = d.max_cost;
  header->free_cells = d.free_cells;
  header->estimated_persistent_device_bytes = sizeof(apgx_v2_board_header) + plane_bytes;
  header->device_view_fingerprint = v2::Fingerprint(*header, plane);
  ClearError();
  return APGX_OK;
}

extern "C" int apgx_v2_board_validate(const apgx_v2_board_header* header, const void* plane, uint64_t plane_bytes) {
  using namespace apgx;
  if (header == nullptr || plane == nullptr)
    return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_v2_board_validate: null argument");
  const apgx_v2_board_header& h = *header;
  if (int s = v2::CheckShape(h, plane_bytes)) return s;
  v2::Derived d;
  if (int s = v2::Derive(h, plane, &d)) return s;
  if (d.min_cost != h.min_cell_cost || d.max_cost != h.max_cell_cost || d.free_cells != h.free_cells)
    return Fail(APGX_E_INTERNAL_INVARIANT, "dense board rejected: cost summary disagrees with the plane");
  if (h.estimated_persistent_device_bytes != sizeof(apgx_v2_board_header) + plane_bytes)
    return Fail(APGX_E_INTERNAL_INVARIANT, "dense board rejected: persistent byte formula");
  if (v2::Fingerprint(h, plane) != h.device_view_fingerprint)
    return Fail(APGX_E_INTERNAL_INVARIANT, "dense board rejected: device view fingerprint mismatch");
  ClearError();
  return APGX_OK;
}

extern "C" int apgx_lattice_create_v2(const apgx_v2_board_header* header, const void* plane, uint64_t plane_bytes,
                                      apgx_lattice** out) {
  using namespace apgx;
  if (out == nullptr) return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_lattice_create_v2: null argument");
  *out = nullptr;
  if (int s = apgx_v2_board_validate(header, plane, plane_bytes)) return s;
  const apgx_lattice_desc d{APGX_LATTICE_SCHEMA_VERSION, header->width, header->height, header->layers,
                            header->cost_bytes, header->via_cost, {0, 0}};
  return apgx_lattice_create(&d, plane, out);
}
