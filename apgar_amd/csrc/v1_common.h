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

}  // namespace v1
}  // namespace apgx
