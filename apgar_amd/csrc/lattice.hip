// lattice.hip — batched exact shortest-path search over dense multi-layer
// lattices on gfx950 (MI355X).
//
// Design (what the reference's heading-aware sweep becomes on CDNA4):
//  * Every net owns a window-local label plane of packed 32-bit words
//      word = (distance << 3) | predecessor_direction
//    Unsigned min over these words is the lexicographic (distance, predecessor)
//    min.  Direction codes are ordered like the predecessor's cell index
//    (layer below < south < west < east < north < layer above), so at the fixed
//    point the low bits name the lowest-index predecessor that achieves the
//    distance: the canonical predecessor of gpu/cuda/planar_router.cu:1084-1165,
//    restated for dense lattices.  Distance 0 with code 7 marks a source.
//  * The window is cut into TX x TY x L tiles.  A workgroup stages one tile plus
//    a one-cell halo of its four neighbours in LDS, then repeats directional
//    sweeps (E, W, N, S serial min-plus run scans, then via scans up and down
//    the layer stack) until the tile no longer changes.  Changed border cells
//    activate the neighbouring tile for the next round.
//  * Active tiles of all nets form one device work list per round.  A
//    persistent round kernel pops items with an atomic head; pushes are deduped
//    by a per-tile round stamp (atomicMax).  Rounds are separated by kernel
//    boundaries, which make every label written in round r visible in round
//    r+1 on all XCDs without fine-grained memory.
//  * Tiles start uninitialised; a tile's labels are trusted by its neighbours
//    only from the round after its first visit, so label planes never need a
//    full-plane initialisation pass.
//  * After each round a phase kernel takes every net with no queued work (its
//    labels are the exact distances from its current tree), connects the
//    nearest unconnected pin along the predecessor chain, re-seeds that path as
//    sources and re-queues the touched tiles: multi-pin nets keep their labels
//    between connection steps (labels only decrease), so later steps relax
//    only what the new tree cells improve.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "apgx.h"
#include "common.h"
#include "lattice_internal.h"

namespace apgx {
namespace {

constexpr uint32_t kInfD = 0x1FFFFFFFu;        // saturated distance
constexpr uint32_t kInf = kInfD << 3;          // packed "unreached" word
constexpr uint32_t kNever = 0xFFFFFFFFu;       // tile never initialised
constexpr uint32_t kSourceWord = 7u;           // distance 0, no predecessor
constexpr uint32_t kFromBelow = 0, kFromSouth = 1, kFromWest = 2, kFromEast = 3,
                   kFromNorth = 4, kFromAbove = 5;
constexpr uint32_t kSearching = 0xFFu;
constexpr uint32_t kFinalWalk = 0xFEu;  // last pin settled; path written after the round loop
constexpr int kThreads = 256;
constexpr uint32_t kDefaultSweepSchedule = 2;
// Defaults picked by a sweep on four_layer_1024 (max sweeps 3..64 x bucket
// width 96..576): 10 sweeps per register-tile visit and a bucket of 48 mean
// cell costs give 12.0k queries/s against 11.3k for (64, 64 mean costs).
constexpr uint32_t kRegDefaultSweeps = 10;
constexpr double kDefaultDeltaCells = 48.0;  // apgx_route_options.frontier_delta 0
// With landmarks (four_layer_1024, 1024 pairs, 16 fields, neighbour keys
// bounded at their entry edge): tiles ordered by distance + the whole bound,
// register-tile visits up to 64 sweeps: 26.6k queries/s against 23.9k for
// 15/16 of the bound and 32 sweeps, 14.0k without landmarks.  (With the
// coarser whole-tile bound, 15/16 was best: the whole bound put a route's
// tiles in one bucket.)
constexpr uint32_t kDefaultLandmarkWeight = 16;
constexpr uint32_t kRegLandmarkSweeps = 64;
constexpr int kMaxTileCells = 16384;
constexpr int kKeep = kMaxTileCells / kThreads;  // labels kept in registers per thread

struct NetStatic {
  uint32_t x0, y0, w, h;
  uint32_t tiles_x, tiles_y;
  uint32_t tile_offset;
  uint32_t pin_offset;
  uint32_t pin_count;
  uint32_t conn_offset;
  uint64_t label_offset;
  // Element offset of this net's cost plane from SearchArgs::cost: nets of
  // one search may each read their own plane (an allocator's speculative
  // group routes every net on its own snapshot); 0 for a shared plane.
  uint64_t cost_offset;
};

struct NetState {
  uint32_t status;
  uint32_t connected;    // bit per pin
  uint32_t n_connected;  // connections made
  uint32_t cost;
  uint32_t last_push;    // highest round this net queued a tile for
  uint32_t rounds;
  uint32_t last_counted;
  uint32_t pad;
};

struct DevConnection {
  uint32_t pin_index;
  uint32_t distance;
  uint64_t cell_offset;
  uint32_t cell_count;
  uint32_t reserved;
};

struct SearchArgs {
  const void* cost;
  uint32_t W, H, L, via;
  const NetStatic* nets;
  NetState* state;
  uint32_t net_count;
  const uint32_t* tile_net;
  const uint32_t* pins;
  uint32_t* labels;
  uint32_t* tile_stamp;
  uint32_t* tile_init;
  uint32_t* wl[2];
  uint32_t* counters;  // count[r] at 2r, head[r] at 2r+1
  unsigned long long* stats;
  DevConnection* conns;
  uint32_t* cells;
  unsigned long long* cell_cursor;
  unsigned long long cell_capacity;
  // Device-resident round number: kernels read it at entry and the phase
  // kernel's last block advances it, so one captured graph of a chunk of
  // rounds replays unchanged for every chunk.
  uint32_t* round_ptr;  // [0] round, [1] phase-kernel block completion count
  uint32_t max_sweeps;
  // Distance-ordered scheduling (delta-stepping over tiles).  tile_key holds,
  // per round parity, the smallest distance that can newly enter a queued
  // tile; net_minkey the smallest key queued for a net.  A tile is processed
  // only while its key is within `delta` of its net's smallest pending key.
  uint32_t* tile_key;    // 2 * tile_count
  uint32_t* net_minkey;  // 2 * net_count
  // With landmarks, tile_key orders tiles by distance plus lm_weight / 16 of
  // the goal bound, and tile_fkey / net_minf carry distance plus the whole
  // bound, which alone decides when a pin is settled.  (Without landmarks the
  // two keys are equal and only tile_key / net_minkey are kept.)
  uint32_t* tile_fkey;   // 2 * tile_count
  uint32_t* net_minf;    // 2 * net_count
  uint32_t lm_weight;
  uint32_t tile_count;
  uint32_t delta;
  uint32_t min_edge;     // smallest edge weight: bounds path lengths by distance / min_edge + 1
  uint32_t goal_cost;    // smallest in-layer step cost for the goal bound; 0 turns it off
  // 1: a via move into cell v weighs via + cost[v] (every entry into a cell pays
  // that cell's cost, as the allocator's per-cell congestion needs); 0: via.
  uint32_t via_pays_cell;
  // Register tiles: every sweep_mode-th sweep (0, k, 2k, ...) is full (via
  // exchange and stability vote); the others relax in-layer edges only.
  uint32_t sweep_mode;
  // ALT bound: lm_pin[pin slot * lm_count + k] = d(landmark k, pin) and
  // lm_colmax[(k * lm_columns + column) * kLmParts + part] = the largest
  // finite d(landmark k, c) over the cells c of a 64 x 64 column (all layers):
  // part 0 the whole column, 1..4 its west, east, south and north edge
  // lines; lm_count 0: off.
  const uint32_t* lm_pin;
  const uint32_t* lm_colmax;
  uint32_t lm_count;
  uint32_t lm_columns_x;
  uint32_t lm_columns;
  // 1: no pin is settled; each net ends when its queue is empty.
  uint32_t explore;
};

constexpr uint32_t kLmColumn = 64;  // landmark column side (cells)
constexpr uint32_t kLmParts = 5;    // per column: whole, west, east, south, north edge
constexpr uint32_t kLmMinTiles = 4;  // smallest window (in tiles) that uses landmarks

template <typename CostT>
__device__ __forceinline__ uint32_t LoadCost(const void* cost, uint64_t index) {
  return static_cast<const CostT*>(cost)[index];
}

// Queues `tile` for `round` with entry key `key` (a lower bound, minus the
// entry cost, of any label the pending change can create in the tile).
__device__ __forceinline__ void Push(const SearchArgs& a, uint32_t tile, uint32_t net,
                                     uint32_t round, uint32_t key, uint32_t fkey) {
  atomicMin(&a.tile_key[(round & 1) * a.tile_count + tile], key);
  atomicMin(&a.net_minkey[(round & 1) * a.net_count + net], key);
  if (a.lm_count) {
    atomicMin(&a.tile_fkey[(round & 1) * a.tile_count + tile], fkey);
    atomicMin(&a.net_minf[(round & 1) * a.net_count + net], fkey);
  }
  if (atomicMax(&a.tile_stamp[tile], round) < round) {
    const uint32_t slot = atomicAdd(&a.counters[2 * round], 1u);
    a.wl[round & 1][slot] = tile;
    atomicMax(&a.state[net].last_push, round);
  }
}

// Wave-aggregated Push, called by every lane of a wave (want = false for lanes
// with nothing to queue), all for one round.  Keys and the once-per-round
// stamp stay per lane; the lanes that won their stamp take their worklist
// slots from one atomic per wave, ordered by a ballot prefix count.
__device__ __forceinline__ void PushWave(const SearchArgs& a, bool want, uint32_t tile, uint32_t net,
                                         uint32_t round, uint32_t key, uint32_t fkey) {
  bool fresh = false;
  if (want) {
    atomicMin(&a.tile_key[(round & 1) * a.tile_count + tile], key);
    atomicMin(&a.net_minkey[(round & 1) * a.net_count + net], key);
    if (a.lm_count) {
      atomicMin(&a.tile_fkey[(round & 1) * a.tile_count + tile], fkey);
      atomicMin(&a.net_minf[(round & 1) * a.net_count + net], fkey);
    }
    fresh = atomicMax(&a.tile_stamp[tile], round) < round;
  }
  const uint64_t mask = __ballot(fresh);
  if (mask == 0) return;
  const int lane = static_cast<int>(threadIdx.x & 63u);
  const int leader = __builtin_ctzll(mask);
  uint32_t base = 0;
  if (lane == leader) base = atomicAdd(&a.counters[2 * round], static_cast<uint32_t>(__popcll(mask)));
  base = __shfl(base, leader, 64);
  if (fresh) {
    a.wl[round & 1][base + static_cast<uint32_t>(__popcll(mask & ((1ull << lane) - 1ull)))] = tile;
    atomicMax(&a.state[net].last_push, round);
  }
}

// Lower bound on the distance to the nearest unconnected pin of the net from
// the cells a queued change can enter first: with side < 0 any cell of the
// window rectangle [x0, x1] x [y0, y1] (lattice coordinates, every layer);
// with side 0..3 only its west, east, south or north edge line, the one a
// change from the neighbouring tile crosses.  (A change entering at cell c'
// creates labels whose priority is at least its exit distance + 1 + the
// bound at c', and the consistency of the per-cell bound carries that to
// every later cell, so the entry cells are all the key has to cover.)  It is
// the minimum over those pins of two per-pin bounds:
//  - goal_cost * Manhattan distance: each in-layer step costs at least
//    goal_cost and vias cost nothing in the bound;
//  - ALT: for a landmark l with d(l, p) finite, d(c, p) >= d(l, p) - d(l, c)
//    (triangle inequality of the metric the fields were built on, which no
//    search metric undercuts), and d(l, c) <= the largest finite d(l, .) over
//    a set of cells holding the entry cells: the entry edge line of the one
//    64 x 64 column that holds the rectangle when that edge lies on the
//    column's own edge, else every column the rectangle touches (at most
//    2 x 2).  Cells unreachable from l cannot reach p (adjacency is
//    symmetric), so leaving them out of the maxima is sound.
// Each per-cell bound is consistent (it drops by at most the weight of an
// edge), and so are their maximum over landmarks, the minimum over pins and
// the clamp below kInfD; adding the bound to tile keys therefore orders tiles
// by A* priority without changing any label: a change queued with key g can
// only create labels whose own priority exceeds g + bound.
__device__ __forceinline__ uint32_t GoalBound(const SearchArgs& a, const NetStatic& ns, uint32_t connected,
                                              uint32_t x0, uint32_t x1, uint32_t y0, uint32_t y1, int side) {
  if (a.goal_cost == 0 && a.lm_count == 0) return 0;
  // Windows of at most kLmMinTiles tiles search too little for the landmark
  // loop to pay; the choice is fixed per net, so its keys stay consistent.
  const uint32_t lm_count = ns.tiles_x * ns.tiles_y > kLmMinTiles ? a.lm_count : 0u;
  const uint32_t cx0 = x0 / kLmColumn, cx1 = x1 / kLmColumn;
  const uint32_t cy0 = y0 / kLmColumn, cy1 = y1 / kLmColumn;
  // The entry edge's own maximum applies when the rectangle lies in one
  // column and its entry edge is on that column's edge.
  uint32_t part = 0;
  if (side >= 0 && cx0 == cx1 && cy0 == cy1) {
    const bool on_edge = side == 0   ? x0 == cx0 * kLmColumn
                         : side == 1 ? x1 == min(cx0 * kLmColumn + kLmColumn, a.W) - 1
                         : side == 2 ? y0 == cy0 * kLmColumn
                                     : y1 == min(cy0 * kLmColumn + kLmColumn, a.H) - 1;
    if (on_edge) part = 1u + static_cast<uint32_t>(side);
  }
  // Entry cells for the Manhattan part.
  const uint32_t ex0 = side == 1 ? x1 : x0, ex1 = side == 0 ? x0 : x1;
  const uint32_t ey0 = side == 3 ? y1 : y0, ey1 = side == 2 ? y0 : y1;
  // Per landmark, the largest distance over the entry cells: loaded once, then
  // shared by every pin (kept in registers: the loop bound is a constant).
  uint32_t far[APGX_MAX_LANDMARKS];
#pragma unroll
  for (uint32_t k = 0; k < APGX_MAX_LANDMARKS; ++k) {
    far[k] = 0;
    if (k < lm_count) {
      const uint32_t* col = a.lm_colmax + static_cast<uint64_t>(k) * a.lm_columns * kLmParts;
      if (part != 0) {
        far[k] = col[(cy0 * a.lm_columns_x + cx0) * kLmParts + part];
      } else {
        for (uint32_t cy = cy0; cy <= cy1; ++cy)
          for (uint32_t cx = cx0; cx <= cx1; ++cx) far[k] = max(far[k], col[(cy * a.lm_columns_x + cx) * kLmParts]);
      }
    }
  }
  uint64_t best = ~0ull;
  for (uint32_t p = 0; p < ns.pin_count; ++p) {
    if (connected & (1u << p)) continue;
    const uint32_t cell = a.pins[ns.pin_offset + p];
    const uint32_t x = cell % a.W, y = (cell / a.W) % a.H;
    const uint32_t dx = x < ex0 ? ex0 - x : (x > ex1 ? x - ex1 : 0u);
    const uint32_t dy = y < ey0 ? ey0 - y : (y > ey1 ? y - ey1 : 0u);
    uint64_t b = static_cast<uint64_t>(dx + dy) * a.goal_cost;
    const uint32_t* dp = a.lm_pin + static_cast<uint64_t>(ns.pin_offset + p) * a.lm_count;
#pragma unroll
    for (uint32_t k = 0; k < APGX_MAX_LANDMARKS; ++k) {
      if (k >= lm_count) break;
      const uint32_t to_pin = dp[k];
      if (to_pin < kInfD && to_pin > far[k]) b = max(b, static_cast<uint64_t>(to_pin - far[k]));
    }
    best = min(best, b);
  }
  return best == ~0ull ? 0u : static_cast<uint32_t>(min(best, static_cast<uint64_t>(kInfD)));
}

// Tile keys from the smallest distance a change starts from and the goal
// bound of the tile it enters; both stay below kInfD (kInfD means nothing
// queued).  x: ordering key, y: settling key (equal without landmarks).
// lm_weight / 16 of a consistent bound is still consistent, so the ordering
// key is also a valid lower bound; weighting it keeps the tiles along a route
// in distance order where the full ALT bound would put them all in one bucket.
__device__ __forceinline__ uint2 TileKeys(const SearchArgs& a, const NetStatic& ns, uint32_t connected,
                                          uint32_t tx, uint32_t ty, uint32_t tw, uint32_t th, uint32_t dist,
                                          int side) {
  const uint32_t x0 = ns.x0 + tx * tw, y0 = ns.y0 + ty * th;
  const uint32_t x1 = min(x0 + tw, ns.x0 + ns.w) - 1, y1 = min(y0 + th, ns.y0 + ns.h) - 1;
  const uint64_t b = GoalBound(a, ns, connected, x0, x1, y0, y1, side);
  const uint64_t cap = kInfD - 1;
  const uint32_t f = static_cast<uint32_t>(min(static_cast<uint64_t>(dist) + b, cap));
  if (a.lm_count == 0) return make_uint2(f, f);
  const uint32_t k = static_cast<uint32_t>(min(static_cast<uint64_t>(dist) + ((b * a.lm_weight) >> 4), cap));
  return make_uint2(k, f);
}

// Admission of a popped item: items of finished nets are dropped, items whose
// key is beyond the net's bucket (smallest pending key + delta) move to the
// next round unprocessed.  The round's key slot is cleared for reuse two
// rounds later.
__device__ __forceinline__ bool AdmitItem(const SearchArgs& a, uint32_t r, uint32_t g) {
  uint32_t* slot = &a.tile_key[(r & 1) * a.tile_count + g];
  const uint32_t key = *slot;
  *slot = kInfD;
  uint32_t fkey = key;
  if (a.lm_count) {
    uint32_t* fslot = &a.tile_fkey[(r & 1) * a.tile_count + g];
    fkey = *fslot;
    *fslot = kInfD;
  }
  const uint32_t n = a.tile_net[g];
  if (a.state[n].status != kSearching) return false;
  const uint32_t base = a.net_minkey[(r & 1) * a.net_count + n];
  const uint32_t limit = base >= kInfD - a.delta ? kInfD : base + a.delta;
  if (key > limit) {
    Push(a, g, n, r + 1, key, fkey);
    return false;
  }
  return true;
}

// Pops the next admitted item of this round (thread 0 only); kNever when the
// round's list is exhausted.
__device__ __forceinline__ uint32_t PopAdmitted(const SearchArgs& a, uint32_t r, uint32_t count) {
  for (;;) {
    const uint32_t i = atomicAdd(&a.counters[2 * r + 1], 1u);
    if (i >= count) return kNever;
    const uint32_t g = a.wl[r & 1][i];
    if (AdmitItem(a, r, g)) return g;
  }
}

__device__ __forceinline__ uint32_t Step(uint32_t carry_d, uint32_t cost, uint32_t code) {
  return (min(carry_d + cost, kInfD) << 3) | code;
}

// Search statistics, slots of SearchArgs::stats (u64):
//   0 tile visits, 1 window cells of the visited tiles, 2 sweep iterations,
//   3 relaxations (edges relaxed into free cells by the sweeps executed:
//     4 in-layer per free cell and sweep, plus one per adjacent layer on full
//     sweeps), 4 distinct tiles visited, 5 layer sweeps executed (summed over
//     the tile's layers; register tiles skip sweeps of quiet layers),
//   6 full sweeps (via exchange and stability test), 7 cell-buffer overflow
//   flag, 8 full sweeps whose test found a change, 9 free cells of the visited
//   tiles.
constexpr int kStatSlots = 16;
__device__ __forceinline__ void AddVisitStats(const SearchArgs& a, uint64_t cells, uint32_t slots,
                                              uint64_t layer_sweeps, uint32_t full, uint32_t changing,
                                              uint32_t free_cells, uint64_t relaxations, bool first) {
  atomicAdd(&a.stats[0], 1ull);
  atomicAdd(&a.stats[1], static_cast<unsigned long long>(cells));
  atomicAdd(&a.stats[2], static_cast<unsigned long long>(slots));
  atomicAdd(&a.stats[3], static_cast<unsigned long long>(relaxations));
  if (first) atomicAdd(&a.stats[4], 1ull);
  atomicAdd(&a.stats[5], static_cast<unsigned long long>(layer_sweeps));
  atomicAdd(&a.stats[6], static_cast<unsigned long long>(full));
  atomicAdd(&a.stats[8], static_cast<unsigned long long>(changing));
  atomicAdd(&a.stats[9], static_cast<unsigned long long>(free_cells));
}

// One tile visit: stage, sweep to local convergence, write back, activate.
template <int TX, int TY>
__device__ void ProcessTile(const SearchArgs& a, uint32_t r, uint32_t g, uint32_t* s_lab, uint32_t* s_cost,
                            uint32_t* s_flags, uint32_t cost_bytes) {
  constexpr int PX = TX + 3;   // odd LDS pitch: row scans are bank-conflict free
  constexpr int CPX = TX + 1;
  const int tid = threadIdx.x;
  const int L = static_cast<int>(a.L);
  const uint32_t n = a.tile_net[g];
  const NetStatic ns = a.nets[n];
  const uint32_t t = g - ns.tile_offset;
  const uint32_t tx = t % ns.tiles_x, ty = t / ns.tiles_x;
  const uint32_t bx = tx * TX, by = ty * TY;
  const int nx = static_cast<int>(min(static_cast<uint32_t>(TX), ns.w - bx));
  const int ny = static_cast<int>(min(static_cast<uint32_t>(TY), ns.h - by));
  uint32_t* lab = a.labels + ns.label_offset;
  const uint64_t w = ns.w, h = ns.h;
  const bool initd = a.tile_init[g] != kNever;
  auto ready = [&](uint32_t ntx, uint32_t nty) {
    const uint32_t v = a.tile_init[ns.tile_offset + nty * ns.tiles_x + ntx];
    return v != kNever && v < r;
  };
  const bool has_w = tx > 0 && ready(tx - 1, ty);
  const bool has_e = tx + 1 < ns.tiles_x && ready(tx + 1, ty);
  const bool has_s = ty > 0 && ready(tx, ty - 1);
  const bool has_n = ty + 1 < ns.tiles_y && ready(tx, ty + 1);

  const int cells = L * TY * TX;
  uint32_t orig[kKeep];
  uint32_t free_cells = 0, edges = 0;  // this thread's free cells and the edges a sweep relaxes into them
  const bool vias = L > 1 && a.via != 0;
#pragma unroll
  for (int k = 0; k < kKeep; ++k) {
    const int i = tid + k * kThreads;
    orig[k] = kInf;
    if (i < cells) {
      const int x = i % TX, y = (i / TX) % TY, l = i / (TX * TY);
      uint32_t c = 0, v = kInf;
      if (x < nx && y < ny) {
        const uint64_t gidx =
            ns.cost_offset + (static_cast<uint64_t>(l) * a.H + ns.y0 + by + y) * a.W + ns.x0 + bx + x;
        c = cost_bytes == 2 ? LoadCost<uint16_t>(a.cost, gidx) : LoadCost<uint32_t>(a.cost, gidx);
        if (initd) v = lab[(l * h + by + y) * w + bx + x];
        if (c != 0) {
          ++free_cells;
          edges += 4u + (vias ? (l > 0 ? 1u : 0u) + (l + 1 < L ? 1u : 0u) : 0u);
        }
      }
      s_cost[(l * TY + y) * CPX + x] = c ? c : kInfD;
      s_lab[(l * (TY + 2) + y + 1) * PX + x + 1] = v;
      orig[k] = v;
    }
  }
  for (int i = tid; i < L * TY; i += kThreads) {
    const int y = i % TY, l = i / TY;
    uint32_t vw = kInf, ve = kInf;
    if (y < ny) {
      const uint64_t row = (l * h + by + y) * w;
      if (has_w) vw = lab[row + bx - 1];
      if (has_e) ve = lab[row + bx + TX];
    }
    s_lab[(l * (TY + 2) + y + 1) * PX] = vw;
    s_lab[(l * (TY + 2) + y + 1) * PX + TX + 1] = ve;
  }
  for (int i = tid; i < L * TX; i += kThreads) {
    const int x = i % TX, l = i / TX;
    uint32_t vs = kInf, vn = kInf;
    if (x < nx) {
      if (has_s) vs = lab[(l * h + by - 1) * w + bx + x];
      if (has_n) vn = lab[(l * h + by + TY) * w + bx + x];
    }
    s_lab[(l * (TY + 2)) * PX + x + 1] = vs;
    s_lab[(l * (TY + 2) + TY + 1) * PX + x + 1] = vn;
  }
  __syncthreads();

  uint32_t sweeps = 0;
  bool converged = false;
  while (sweeps < a.max_sweeps) {
    bool changed = false;
    // Row scans: east (+x) then west (-x), one row per thread.
    for (int i = tid; i < L * TY; i += kThreads) {
      const int y = i % TY, l = i / TY;
      uint32_t* row = s_lab + (l * (TY + 2) + y + 1) * PX;
      const uint32_t* crow = s_cost + (l * TY + y) * CPX;
      uint32_t carry = row[0] >> 3;
#pragma unroll 8
      for (int x = 0; x < TX; ++x) {
        const uint32_t v = row[x + 1];
        const uint32_t nv = min(v, Step(carry, crow[x], kFromWest));
        changed |= nv != v;
        row[x + 1] = nv;
        carry = nv >> 3;
      }
      carry = row[TX + 1] >> 3;
#pragma unroll 8
      for (int x = TX - 1; x >= 0; --x) {
        const uint32_t v = row[x + 1];
        const uint32_t nv = min(v, Step(carry, crow[x], kFromEast));
        changed |= nv != v;
        row[x + 1] = nv;
        carry = nv >> 3;
      }
    }
    __syncthreads();
    // Column scans: north (+y) then south (-y), one column per thread.
    for (int i = tid; i < L * TX; i += kThreads) {
      const int x = i % TX, l = i / TX;
      uint32_t* col = s_lab + l * (TY + 2) * PX + x + 1;
      const uint32_t* ccol = s_cost + l * TY * CPX + x;
      uint32_t carry = col[0] >> 3;
#pragma unroll 8
      for (int y = 0; y < TY; ++y) {
        const uint32_t v = col[(y + 1) * PX];
        const uint32_t nv = min(v, Step(carry, ccol[y * CPX], kFromSouth));
        changed |= nv != v;
        col[(y + 1) * PX] = nv;
        carry = nv >> 3;
      }
      carry = col[(TY + 1) * PX] >> 3;
#pragma unroll 8
      for (int y = TY - 1; y >= 0; --y) {
        const uint32_t v = col[(y + 1) * PX];
        const uint32_t nv = min(v, Step(carry, ccol[y * CPX], kFromNorth));
        changed |= nv != v;
        col[(y + 1) * PX] = nv;
        carry = nv >> 3;
      }
    }
    // Via scans up and down each (x, y) stack.
    if (L > 1 && a.via != 0) {
      __syncthreads();
      for (int i = tid; i < TX * TY; i += kThreads) {
        const int x = i % TX, y = i / TX;
        uint32_t* stack = s_lab + (y + 1) * PX + x + 1;
        const uint32_t* cstack = s_cost + y * CPX + x;
        constexpr int kLab = (TY + 2) * PX, kCost = TY * CPX;
        uint32_t carry = stack[0] >> 3;
        for (int l = 1; l < L; ++l) {
          const uint32_t v = stack[l * kLab];
          const uint32_t ce = cstack[l * kCost];
          const uint32_t c = ce >= kInfD ? kInfD : a.via + (a.via_pays_cell ? ce : 0u);
          const uint32_t nv = min(v, Step(carry, c, kFromBelow));
          changed |= nv != v;
          stack[l * kLab] = nv;
          carry = nv >> 3;
        }
        for (int l = L - 2; l >= 0; --l) {
          const uint32_t v = stack[l * kLab];
          const uint32_t ce = cstack[l * kCost];
          const uint32_t c = ce >= kInfD ? kInfD : a.via + (a.via_pays_cell ? ce : 0u);
          const uint32_t nv = min(v, Step(carry, c, kFromAbove));
          changed |= nv != v;
          stack[l * kLab] = nv;
          carry = nv >> 3;
        }
      }
    }
    ++sweeps;
    if (!__syncthreads_or(changed)) {
      converged = true;
      break;
    }
  }

#pragma unroll
  for (int k = 0; k < kKeep; ++k) {
    const int i = tid + k * kThreads;
    if (i < cells) {
      const int x = i % TX, y = (i / TX) % TY, l = i / (TX * TY);
      if (x < nx && y < ny) {
        const uint32_t v = s_lab[(l * (TY + 2) + y + 1) * PX + x + 1];
        if (v != orig[k] || !initd) lab[(l * h + by + y) * w + bx + x] = v;
        if ((v >> 3) != (orig[k] >> 3)) {
          // Wake test, as in RegRoundKernel: the neighbour's cell across the
          // border is queued only if this edge beats its halo label (the
          // staged ring, untouched by the sweeps; kInf when not yet written).
          const uint32_t d = v >> 3;
          const uint64_t gidx =
              ns.cost_offset + (static_cast<uint64_t>(l) * a.H + ns.y0 + by + y) * a.W + ns.x0 + bx + x;
          auto wakes = [&](int64_t off, uint32_t halo, uint32_t code) {
            const uint64_t cell = static_cast<uint64_t>(static_cast<int64_t>(gidx) + off);
            const uint32_t c = cost_bytes == 2 ? LoadCost<uint16_t>(a.cost, cell) : LoadCost<uint32_t>(a.cost, cell);
            return c != 0 && Step(d, c, code) < halo;
          };
          const uint32_t* ring = s_lab + l * (TY + 2) * PX;
          if (x == 0 && tx > 0 && wakes(-1, ring[(y + 1) * PX], kFromEast)) atomicMin(&s_flags[0], d);
          if (x == nx - 1 && tx + 1 < ns.tiles_x && wakes(1, ring[(y + 1) * PX + TX + 1], kFromWest))
            atomicMin(&s_flags[1], d);
          if (y == 0 && ty > 0 && wakes(-static_cast<int64_t>(a.W), ring[x + 1], kFromNorth))
            atomicMin(&s_flags[2], d);
          if (y == ny - 1 && ty + 1 < ns.tiles_y && wakes(static_cast<int64_t>(a.W), ring[(TY + 1) * PX + x + 1], kFromSouth))
            atomicMin(&s_flags[3], d);
        }
      }
    }
  }
  if (free_cells) {
    atomicAdd(&s_flags[4], free_cells);
    atomicAdd(&s_flags[5], edges);
  }
  __syncthreads();
  if (tid == 0) {
    if (!initd) a.tile_init[g] = r;
    const uint32_t next = r + 1;
    const uint32_t conn = a.state[n].connected;
    if (s_flags[0] < kInfD && tx > 0) {
      const uint2 k = TileKeys(a, ns, conn, tx - 1, ty, TX, TY, s_flags[0], 1);
      Push(a, g - 1, n, next, k.x, k.y);
    }
    if (s_flags[1] < kInfD && tx + 1 < ns.tiles_x) {
      const uint2 k = TileKeys(a, ns, conn, tx + 1, ty, TX, TY, s_flags[1], 0);
      Push(a, g + 1, n, next, k.x, k.y);
    }
    if (s_flags[2] < kInfD && ty > 0) {
      const uint2 k = TileKeys(a, ns, conn, tx, ty - 1, TX, TY, s_flags[2], 3);
      Push(a, g - ns.tiles_x, n, next, k.x, k.y);
    }
    if (s_flags[3] < kInfD && ty + 1 < ns.tiles_y) {
      const uint2 k = TileKeys(a, ns, conn, tx, ty + 1, TX, TY, s_flags[3], 2);
      Push(a, g + ns.tiles_x, n, next, k.x, k.y);
    }
    if (!converged) Push(a, g, n, next, 0, 0);
    if (atomicMax(&a.state[n].last_counted, r) < r) atomicAdd(&a.state[n].rounds, 1u);
    // Every sweep of an LDS tile is full: all layers, vias, one stability test.
    AddVisitStats(a, static_cast<uint64_t>(nx) * ny * L, sweeps, static_cast<uint64_t>(sweeps) * L,
                  sweeps, converged ? sweeps - 1 : sweeps, s_flags[4],
                  static_cast<uint64_t>(sweeps) * s_flags[5], !initd);
  }
}

template <int TX, int TY>
__global__ __launch_bounds__(kThreads) void RoundKernel(SearchArgs a, uint32_t cost_bytes) {
  extern __shared__ uint32_t smem[];
  constexpr int PX = TX + 3;
  uint32_t* s_lab = smem;
  uint32_t* s_cost = smem + static_cast<int>(a.L) * (TY + 2) * PX;
  __shared__ uint32_t s_item;
  __shared__ uint32_t s_flags[6];  // smallest changed border distance per side; free cells, sweep edges
  const uint32_t r = __builtin_amdgcn_readfirstlane(*a.round_ptr);
  const uint32_t count = a.counters[2 * r];
  for (;;) {
    if (threadIdx.x == 0) {
      s_item = PopAdmitted(a, r, count);
      for (int k = 0; k < 4; ++k) s_flags[k] = kInfD;
      s_flags[4] = 0;
      s_flags[5] = 0;
    }
    __syncthreads();
    const uint32_t g = s_item;
    if (g == kNever) break;
    ProcessTile<TX, TY>(a, r, g, s_lab, s_cost, s_flags, cost_bytes);
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Register-resident tile kernel (64 x 64 x L tiles, one wave per layer).
//
// Wave l holds layer l of the tile column-major in VGPRs: lane x keeps the 64
// packed labels of column x and the 64 pre-shifted entry costs as signed u16
// pairs (cost << 3, or 0xFFFF for a blocked or out-of-window cell, so cell
// costs are limited to kRegMaxCost).  A relaxation step is one saturating add
// of the sign-extended half (all ones for a blocked cell saturates), an and_or
// that swaps in the direction code and a min, with no memory access in the
// dependency chain:
//   north/south  serial over the lane's own 64 registers;
//   east/west    after a wave-private LDS transpose (pitch 65: conflict free),
//                serial over the transposed row, then transposed back;
//   vias         every wave publishes its layer in LDS, then relaxes each cell
//                from the layers below and above (one hop per sweep; sweeps
//                repeat until the tile is stable, so chains of vias converge).
// Stability test: labels only decrease, so a lane's column is unchanged by a
// sweep exactly when the sums of its 64 low and of its 64 high label halves
// (each below 2^22, no wrap) are both unchanged.
// ---------------------------------------------------------------------------
constexpr int kRT = 64;          // tile side
constexpr int kRP = kRT + 1;     // LDS pitch of the transpose buffers
constexpr int kEdgeRows = 16;    // rows loaded together when staging a window-edge tile
constexpr uint32_t kBlocked8 = 0xFFFFFFFFu;
constexpr uint32_t kBlockedHalf = 0xFFFFu;  // packed cost half of a blocked cell
constexpr uint32_t kRegMaxCost = 4095;      // (cost << 3) must stay a positive int16
constexpr uint32_t kRegWideMaxCost = 32767; // WIDE tiles: the unshifted cost must stay a positive int16

__device__ __forceinline__ uint32_t Relax8(uint32_t carry, uint32_t c8, uint32_t code) {
  return (__builtin_elementwise_add_sat(carry, c8) & ~7u) | code;
}

__device__ __forceinline__ void WaveSync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Packed cost pair helpers: each 32-bit register holds two 16-bit entries,
// cost << 3 for a free cell and 0xFFFF for a blocked one; sign extension gives
// the pre-shifted cost or all ones (both fold into the SDWA operand of the add).
__device__ __forceinline__ uint32_t CostLo(uint32_t pair) {
  return static_cast<uint32_t>(static_cast<int32_t>(static_cast<int16_t>(pair & 0xFFFFu)));
}
__device__ __forceinline__ uint32_t CostHi(uint32_t pair) {
  return static_cast<uint32_t>(static_cast<int32_t>(pair) >> 16);
}
// WIDE tiles hold the unshifted cost (up to kRegWideMaxCost) and shift it at
// each relaxation: one more VALU op, off the carry chain (it does not depend on
// the carry).  A blocked half (-1) shifts to 0xFFFFFFF8, which saturates the
// add for every label (all labels are at least 7), as all ones does.
template <bool WIDE>
__device__ __forceinline__ uint32_t PackCost(uint32_t c) {
  return c ? (WIDE ? c : c << 3) : kBlockedHalf;
}
// Makes the packed cost registers opaque to the compiler, so that values
// derived from them are recomputed after this point instead of kept live.
__device__ __forceinline__ void RegBarrier(uint32_t (&pairs)[32]) {
#pragma unroll
  for (int i = 0; i < 32; ++i) asm volatile("" : "+v"(pairs[i]));
}
template <bool WIDE>
__device__ __forceinline__ uint32_t Cost8(uint32_t pair, bool hi) {
  const uint32_t s = hi ? CostHi(pair) : CostLo(pair);
  return WIDE ? s << 3 : s;
}

// PAYS: vias also pay the entered cell's cost (allocator searches); a
// template parameter so the plain kernel's register allocation is untouched.
// WIDE: costs up to kRegWideMaxCost (allocator snapshots under a high present
// factor), held unshifted.
template <int L, bool PAYS, bool WIDE = false>
__global__ __launch_bounds__(64 * L) __attribute__((amdgpu_waves_per_eu(2)))
void RegRoundKernel(SearchArgs a, uint32_t cost_bytes) {
  extern __shared__ uint32_t smem[];
  uint32_t* s_t = smem;                       // L transpose/publish buffers, kRT * kRP each
  __shared__ uint32_t s_item[2];  // current / prefetched item (parity of the iteration)
  __shared__ uint32_t s_side[5];   // smallest changed border distance per side (W, E, S, N), tile min
  __shared__ uint32_t s_changed[2 * L];  // stability votes, double-buffered by vote parity
  __shared__ uint32_t s_wstat[3 * L];    // per layer: free cells, sweeps executed, relaxations
  const int l = __builtin_amdgcn_readfirstlane(static_cast<int>(threadIdx.x >> 6));  // wave-uniform layer
  const uint32_t r = __builtin_amdgcn_readfirstlane(*a.round_ptr);
  const uint32_t count = a.counters[2 * r];
  // The last wave pops the next admitted item during the current item's
  // epilogue, so the pop's dependent atomics overlap the pushes.
  const bool popper = threadIdx.x == 64 * (L - 1);
  if (popper) s_item[0] = PopAdmitted(a, r, count);
  for (uint32_t iter = 0;; ++iter) {
    if (threadIdx.x == 0) {
      for (int k = 0; k < 5; ++k) s_side[k] = kInfD;
    }
    __syncthreads();
    const uint32_t g = s_item[iter & 1];
    if (g == kNever) break;
    // The lane index is opaque per tile: every LDS and global offset derived
    // from it is recomputed per tile rather than hoisted out of this loop
    // (64 hoisted offsets would stay live across the whole loop).
    int lane = static_cast<int>(threadIdx.x & 63);
    asm volatile("" : "+v"(lane));
    uint32_t* T = s_t + l * kRT * kRP;
    const uint32_t n = a.tile_net[g];
    const NetStatic ns = a.nets[n];
    const uint32_t t = g - ns.tile_offset;
    const uint32_t tx = t % ns.tiles_x, ty = t / ns.tiles_x;
    const uint32_t bx = tx * kRT, by = ty * kRT;
    const int nx = static_cast<int>(min(static_cast<uint32_t>(kRT), ns.w - bx));
    const int ny = static_cast<int>(min(static_cast<uint32_t>(kRT), ns.h - by));
    const uint32_t w = ns.w;
    // Uniform bases; per-cell offsets fit in 32 bits (cells < 2^32).
    uint32_t* lbase = a.labels + ns.label_offset + (static_cast<uint64_t>(l) * ns.h + by) * w + bx;
    const uint64_t cbase_index = ns.cost_offset + (static_cast<uint64_t>(l) * a.H + ns.y0 + by) * a.W + ns.x0 + bx;
    const uint16_t* cbase16 = static_cast<const uint16_t*>(a.cost) + cbase_index;
    const uint32_t* cbase32 = static_cast<const uint32_t*>(a.cost) + cbase_index;
    const bool initd = a.tile_init[g] != kNever;
    auto ready = [&](uint32_t ntx, uint32_t nty) {
      const uint32_t v = a.tile_init[ns.tile_offset + nty * ns.tiles_x + ntx];
      return v != kNever && v < r;
    };
    const bool has_w = tx > 0 && ready(tx - 1, ty);
    const bool has_e = tx + 1 < ns.tiles_x && ready(tx + 1, ty);
    const bool has_s = ty > 0 && ready(tx, ty - 1);
    const bool has_n = ty + 1 < ns.tiles_y && ready(tx, ty + 1);

    const bool xin = lane < nx;
    uint32_t lab[kRT], cst[kRT / 2], cstr[kRT / 2];
    if (nx == kRT && ny == kRT) {
      // Full tile: unrolled coalesced row loads straight into registers.
#pragma unroll
      for (int y = 0; y < kRT; ++y) {
        const uint32_t c = cost_bytes == 2 ? cbase16[y * a.W + lane] : cbase32[y * a.W + lane];
        const uint32_t e = PackCost<WIDE>(c);
        if (y & 1) cst[y >> 1] |= e << 16; else cst[y >> 1] = e;
      }
      if (initd) {
#pragma unroll
        for (int y = 0; y < kRT; ++y) lab[y] = lbase[y * w + lane];
      } else {
#pragma unroll
        for (int y = 0; y < kRT; ++y) lab[y] = kInf;
      }
#pragma unroll
      for (int i = 0; i < kRT / 2; ++i) {
        T[(2 * i) * kRP + lane] = cst[i] & 0xFFFFu;
        T[(2 * i + 1) * kRP + lane] = cst[i] >> 16;
      }
      WaveSync();
#pragma unroll
      for (int i = 0; i < kRT / 2; ++i) cstr[i] = T[lane * kRP + 2 * i] | (T[lane * kRP + 2 * i + 1] << 16);
      WaveSync();
#pragma unroll
      for (int y = 0; y < kRT; ++y) T[y * kRP + lane] = lab[y];
      WaveSync();
    } else {
      // Window-edge tile: predicated runtime loops stage rows through the
      // wave's LDS buffer (out-of-window cells masked to blocked / unreached),
      // then unrolled conflict-free reads fill registers.  No per-row
      // predicate lives in the unrolled register code.
      // Rows go in groups of kEdgeRows: the group's loads are all issued
      // before the first is used (a uniform test per row, of the cost width or
      // of the tile's initialisation, would branch and serialise them).
      const uint32_t xc = min(static_cast<uint32_t>(lane), static_cast<uint32_t>(nx - 1));
#pragma unroll 1
      for (int y0 = 0; y0 < kRT; y0 += kEdgeRows) {
        uint32_t c[kEdgeRows];
        if (cost_bytes == 2) {
#pragma unroll
          for (int k = 0; k < kEdgeRows; ++k) c[k] = cbase16[static_cast<uint32_t>(min(y0 + k, ny - 1)) * a.W + xc];
        } else {
#pragma unroll
          for (int k = 0; k < kEdgeRows; ++k) c[k] = cbase32[static_cast<uint32_t>(min(y0 + k, ny - 1)) * a.W + xc];
        }
#pragma unroll
        for (int k = 0; k < kEdgeRows; ++k)
          T[(y0 + k) * kRP + lane] = PackCost<WIDE>((xin && y0 + k < ny) ? c[k] : 0u);
      }
      WaveSync();
#pragma unroll
      for (int i = 0; i < kRT / 2; ++i) {
        cst[i] = T[(2 * i) * kRP + lane] | (T[(2 * i + 1) * kRP + lane] << 16);
        cstr[i] = T[lane * kRP + 2 * i] | (T[lane * kRP + 2 * i + 1] << 16);
      }
      WaveSync();
      if (initd) {
#pragma unroll 1
        for (int y0 = 0; y0 < kRT; y0 += kEdgeRows) {
          uint32_t v[kEdgeRows];
#pragma unroll
          for (int k = 0; k < kEdgeRows; ++k) v[k] = lbase[static_cast<uint32_t>(min(y0 + k, ny - 1)) * w + xc];
#pragma unroll
          for (int k = 0; k < kEdgeRows; ++k) T[(y0 + k) * kRP + lane] = (xin && y0 + k < ny) ? v[k] : kInf;
        }
      } else {
#pragma unroll 8
        for (int y = 0; y < kRT; ++y) T[y * kRP + lane] = kInf;
      }
      WaveSync();
#pragma unroll
      for (int y = 0; y < kRT; ++y) lab[y] = T[y * kRP + lane];
    }
    // Borders before the sweeps (distances compared after).
    const uint32_t orig_s = T[lane];
    const uint32_t orig_n = T[(ny - 1) * kRP + lane];
    const uint32_t orig_w = T[lane * kRP];
    const uint32_t orig_e = T[lane * kRP + nx - 1];
    const uint32_t halo_s = (has_s && xin) ? lbase[static_cast<int64_t>(lane) - static_cast<int64_t>(w)] : kInf;
    const uint32_t halo_n = (has_n && xin) ? lbase[lane + kRT * w] : kInf;
    const uint32_t halo_w = (has_w && lane < ny) ? lbase[static_cast<int64_t>(lane) * w - 1] : kInf;
    const uint32_t halo_e = (has_e && lane < ny) ? lbase[lane * w + kRT] : kInf;
    WaveSync();
    const uint32_t via8 = a.via << 3;

    uint32_t sum_lo = 0, sum_hi = 0;  // label half sums of the lane's column
#pragma unroll
    for (int y = 0; y < kRT; ++y) {
      sum_lo += lab[y] & 0xFFFFu;
      sum_hi += lab[y] >> 16;
    }
    // Vias from a missing layer get an all-ones cost (the add saturates).
    const uint32_t via_b = l > 0 ? via8 : kBlocked8;
    const uint32_t via_a = l + 1 < L ? via8 : kBlocked8;
    const uint32_t* below = s_t + (l > 0 ? l - 1 : l) * kRT * kRP;
    const uint32_t* above = s_t + (l + 1 < L ? l + 1 : l) * kRT * kRP;

    uint32_t sweeps = 0, votes = 0, executed = 0, changing = 0;
    bool converged = false;
    // A quiet layer: unchanged since the last vote, as are the layers next to
    // it.  Its labels are then a fixed point of the in-layer passes (halos are
    // fixed for the visit) and of its via inputs, so it skips the in-layer-only
    // sweeps up to the next full sweep, whose via exchange is where a change
    // of a neighbouring layer reaches it.  Every wave still meets the same
    // barriers: those belong to full sweeps only.
    bool quiet = false;
    const uint32_t period = max(a.sweep_mode, 1u);
#pragma unroll 1
    while (sweeps < a.max_sweeps) {
      // Keep the packed costs packed: without this the compiler unpacks all
      // 128 entries once outside the loop and runs out of registers.
#pragma unroll
      for (int i = 0; i < kRT / 2; ++i) asm volatile("" : "+v"(cst[i]), "+v"(cstr[i]));
      uint32_t carry = halo_s;
#pragma unroll
      for (int y = 0; y < kRT; ++y) {
        const uint32_t c8 = Cost8<WIDE>(cst[y >> 1], y & 1);
        lab[y] = min(lab[y], Relax8(carry, c8, kFromSouth));
        carry = lab[y];
      }
      if constexpr (WIDE) RegBarrier(cst);  // recompute the shifted costs: no 64 live copies
      carry = halo_n;
#pragma unroll
      for (int y = kRT - 1; y >= 0; --y) {
        const uint32_t c8 = Cost8<WIDE>(cst[y >> 1], y & 1);
        lab[y] = min(lab[y], Relax8(carry, c8, kFromNorth));
        carry = lab[y];
      }
#pragma unroll
      for (int y = 0; y < kRT; ++y) T[y * kRP + lane] = lab[y];
      WaveSync();
#pragma unroll
      for (int x = 0; x < kRT; ++x) lab[x] = T[lane * kRP + x];
      carry = halo_w;
#pragma unroll
      for (int x = 0; x < kRT; ++x) {
        const uint32_t c8 = Cost8<WIDE>(cstr[x >> 1], x & 1);
        lab[x] = min(lab[x], Relax8(carry, c8, kFromWest));
        carry = lab[x];
      }
      if constexpr (WIDE) RegBarrier(cstr);
      carry = halo_e;
#pragma unroll
      for (int x = kRT - 1; x >= 0; --x) {
        const uint32_t c8 = Cost8<WIDE>(cstr[x >> 1], x & 1);
        lab[x] = min(lab[x], Relax8(carry, c8, kFromEast));
        carry = lab[x];
      }
      WaveSync();
#pragma unroll
      for (int x = 0; x < kRT; ++x) T[lane * kRP + x] = lab[x];
      WaveSync();
#pragma unroll
      for (int y = 0; y < kRT; ++y) lab[y] = T[y * kRP + lane];
      // In-layer-only sweep (sweeps not divisible by the period): no via
      // exchange, no vote and no barrier.  Other waves read this wave's buffer
      // only between the barriers of a full sweep, so overwriting it here is
      // safe; the next full sweep's vote compares against the sums of the last
      // full sweep and so covers this sweep's changes too.
      if (sweeps % period != 0) {
        ++sweeps;
        ++executed;
        continue;
      }
      if (a.via != 0) {  // the host passes via = 0 for single-layer lattices
        __syncthreads();  // every layer published in its buffer
        if constexpr (PAYS) {
          if constexpr (WIDE) RegBarrier(cst);
          // Via weight via + cost of the entered cell: the packed cost half
          // (all ones when blocked) saturates the sum for blocked cells.
#pragma unroll
          for (int y = 0; y < kRT; ++y) {
            const uint32_t c8 = Cost8<WIDE>(cst[y >> 1], y & 1);
            const uint32_t wb = __builtin_elementwise_add_sat(via_b, c8);
            const uint32_t wa = __builtin_elementwise_add_sat(via_a, c8);
            const uint32_t fb = __builtin_elementwise_add_sat(below[y * kRP + lane], wb) & ~7u;  // kFromBelow = 0
            const uint32_t fa = Relax8(above[y * kRP + lane], wa, kFromAbove);
            lab[y] = min(lab[y], min(fb, fa));
          }
        } else {
#pragma unroll
          for (int y = 0; y < kRT; ++y) {
            // All ones for a blocked cell (sign of its packed cost half): the
            // via adds then saturate, so the cell keeps its label.  Arithmetic
            // rather than a branch keeps both LDS reads unconditional.
            const uint32_t blk = (y & 1) ? static_cast<uint32_t>(static_cast<int32_t>(cst[y >> 1]) >> 31)
                                         : static_cast<uint32_t>(static_cast<int32_t>(cst[y >> 1] << 16) >> 31);
            const uint32_t fb = __builtin_elementwise_add_sat(below[y * kRP + lane], via_b | blk) & ~7u;  // kFromBelow = 0
            const uint32_t fa = Relax8(above[y * kRP + lane], via_a | blk, kFromAbove);
            lab[y] = min(lab[y], min(fb, fa));
          }
        }
      }
      uint32_t new_lo = 0, new_hi = 0;
#pragma unroll
      for (int y = 0; y < kRT; ++y) {
        new_lo += lab[y] & 0xFFFFu;
        new_hi += lab[y] >> 16;
      }
      const bool lane_changed = new_lo != sum_lo || new_hi != sum_hi;
      sum_lo = new_lo;
      sum_hi = new_hi;
      bool any;
      {
        // One barrier per vote: every wave finished its via reads before it,
        // so the next sweep may overwrite the published buffers, and the
        // next vote writes the other half of s_changed (a wave cannot reach
        // that vote before every wave has read this one: the next full
        // sweep's first barrier holds it).
        uint32_t* flags = s_changed + (votes & 1u) * L;
        const bool wave_any = __any(lane_changed);
        if (lane == 0) flags[l] = wave_any ? 1u : 0u;
        __syncthreads();
        any = false;
#pragma unroll
        for (int k = 0; k < L; ++k) any |= flags[k] != 0;
        quiet = __builtin_amdgcn_readfirstlane(
                    (flags[l] | (l > 0 ? flags[l - 1] : 0u) | (l + 1 < L ? flags[l + 1] : 0u)) == 0 ? 1 : 0) != 0;
        ++votes;
      }
      ++sweeps;
      ++executed;
      if (!any) {
        converged = true;
        break;
      }
      ++changing;
      if (quiet) sweeps += period - 1;  // skip the in-layer-only sweeps up to the next full one
    }

    // Write back through the tile buffer; detect border distance changes.
    if constexpr (L > 1) __syncthreads();  // other layers may still read this buffer
#pragma unroll
    for (int y = 0; y < kRT; ++y) T[y * kRP + lane] = lab[y];
    WaveSync();
    // Free cells of this layer of the tile, for the statistics (blocked and
    // out-of-window cells hold the blocked cost half).  Counted here, after
    // the sweeps: counting at staging time costs the sweep loop registers.
    uint32_t free_l = 0;
#pragma unroll
    for (int i = 0; i < kRT / 2; ++i)
      free_l += ((cst[i] & 0xFFFFu) != kBlockedHalf ? 1u : 0u) + ((cst[i] >> 16) != kBlockedHalf ? 1u : 0u);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) free_l += __shfl_xor(free_l, off, 64);
    free_l = __builtin_amdgcn_readfirstlane(free_l);
    uint32_t col_min = kInf;
#pragma unroll
    for (int y = 0; y < kRT; ++y) col_min = min(col_min, lab[y]);
    if (!converged) atomicMin(&s_side[4], col_min >> 3);
    {
      // A border cell whose distance changed wakes the neighbouring tile only
      // if relaxing the edge into the neighbour's cell beats that cell's label
      // as read at the start of this visit (the halo; kInf when the neighbour
      // was not written before this round).  Labels only decrease and a tile's
      // labels are written by its own visit alone, so the current label is no
      // higher than the halo: the test may wake a tile needlessly but never
      // misses a change the neighbour's fixed point depends on.
      const uint32_t ws = T[lane], wn = T[(ny - 1) * kRP + lane];
      const uint32_t ww = T[lane * kRP], we = T[lane * kRP + nx - 1];
      auto cost8 = [&](int64_t off) -> uint32_t {
        const uint32_t c = cost_bytes == 2 ? cbase16[off] : cbase32[off];
        return c ? c << 3 : kBlocked8;
      };
      if (xin && ty > 0 && (ws >> 3) != (orig_s >> 3) &&
          Relax8(ws, cost8(static_cast<int64_t>(lane) - static_cast<int64_t>(a.W)), kFromNorth) < halo_s)
        atomicMin(&s_side[2], ws >> 3);
      if (xin && ty + 1 < ns.tiles_y && (wn >> 3) != (orig_n >> 3) &&
          Relax8(wn, cost8(static_cast<int64_t>(kRT) * a.W + lane), kFromSouth) < halo_n)
        atomicMin(&s_side[3], wn >> 3);
      if (lane < ny && tx > 0 && (ww >> 3) != (orig_w >> 3) &&
          Relax8(ww, cost8(static_cast<int64_t>(lane) * a.W - 1), kFromEast) < halo_w)
        atomicMin(&s_side[0], ww >> 3);
      if (lane < ny && tx + 1 < ns.tiles_x && (we >> 3) != (orig_e >> 3) &&
          Relax8(we, cost8(static_cast<int64_t>(lane) * a.W + kRT), kFromWest) < halo_e)
        atomicMin(&s_side[1], we >> 3);
    }
    if (xin) {
#pragma unroll 8
      for (int y = 0; y < ny; ++y) lbase[y * w + lane] = T[y * kRP + lane];
    }
    if (lane == 0) {
      const uint32_t nvia = a.via != 0 ? (l > 0 ? 1u : 0u) + (l + 1 < L ? 1u : 0u) : 0u;
      s_wstat[l] = free_l;
      s_wstat[L + l] = executed;
      s_wstat[2 * L + l] = free_l * (4u * executed + nvia * votes);
    }
    __syncthreads();
    // Epilogue: lanes 0..4 of the first wave issue the four neighbour pushes
    // and the self push together; lane 5 counts the net's round, lane 6 the
    // statistics; the popper fetches the next item meanwhile.
    if (threadIdx.x < 64) {  // the first wave
      const uint32_t next = r + 1;
      const int k = threadIdx.x;
      if (k == 0 && !initd) a.tile_init[g] = r;
      bool want = false;
      uint32_t target = g, key = kInfD, fkey = kInfD;
      if (k < 5) {
        bool exists;
        uint32_t ttx = tx, tty = ty;
        int side = -1;  // the target's edge a change from this tile enters through
        switch (k) {
          case 0: exists = tx > 0; target = g - 1; ttx = tx - 1; side = 1; break;
          case 1: exists = tx + 1 < ns.tiles_x; target = g + 1; ttx = tx + 1; side = 0; break;
          case 2: exists = ty > 0; target = g - ns.tiles_x; tty = ty - 1; side = 3; break;
          case 3: exists = ty + 1 < ns.tiles_y; target = g + ns.tiles_x; tty = ty + 1; side = 2; break;
          default: exists = !converged; target = g; break;
        }
        want = s_side[k] < kInfD && exists;
        if (want) {
          const uint2 kk = TileKeys(a, ns, a.state[n].connected, ttx, tty, kRT, kRT, s_side[k], side);
          key = kk.x;
          fkey = kk.y;
        }
      }
      PushWave(a, want, target, n, next, key, fkey);  // lanes 0..4: four neighbours and the tile itself
      if (k == 5) {
        if (atomicMax(&a.state[n].last_counted, r) < r) atomicAdd(&a.state[n].rounds, 1u);
      } else if (k == 6) {
        uint32_t free_cells = 0, layer_sweeps = 0;
        uint64_t relax = 0;
#pragma unroll
        for (int q = 0; q < L; ++q) {
          free_cells += s_wstat[q];
          layer_sweeps += s_wstat[L + q];
          relax += s_wstat[2 * L + q];
        }
        AddVisitStats(a, static_cast<uint64_t>(nx) * ny * L, min(sweeps, a.max_sweeps), layer_sweeps, votes,
                      changing, free_cells, relax, !initd);
      }
    }
    if (popper) s_item[(iter + 1) & 1] = PopAdmitted(a, r, count);
  }
}

size_t RegTileLdsBytes(uint32_t layers) {
  return static_cast<size_t>(layers) * (kRT * kRP + 2 * kRT) * 4;
}

template <bool PAYS, bool WIDE>
const void* RegKernelForT(uint32_t layers) {
  switch (layers) {
    case 1: return reinterpret_cast<const void*>(&RegRoundKernel<1, PAYS, WIDE>);
    case 2: return reinterpret_cast<const void*>(&RegRoundKernel<2, PAYS, WIDE>);
    case 3: return reinterpret_cast<const void*>(&RegRoundKernel<3, PAYS, WIDE>);
    case 4: return reinterpret_cast<const void*>(&RegRoundKernel<4, PAYS, WIDE>);
    case 5: return reinterpret_cast<const void*>(&RegRoundKernel<5, PAYS, WIDE>);
    case 6: return reinterpret_cast<const void*>(&RegRoundKernel<6, PAYS, WIDE>);
    case 7: return reinterpret_cast<const void*>(&RegRoundKernel<7, PAYS, WIDE>);
    case 8: return reinterpret_cast<const void*>(&RegRoundKernel<8, PAYS, WIDE>);
    default: return nullptr;
  }
}

// Wide register tiles are built for allocator searches only (vias paying the
// entered cell), where snapshot costs grow with the present factor; plain
// searches with costs above kRegMaxCost run LDS tiles.
const void* RegKernelFor(uint32_t layers, bool via_pays_cell, uint32_t max_cost) {
  if (!via_pays_cell) return RegKernelForT<false, false>(layers);
  return max_cost > kRegMaxCost ? RegKernelForT<true, true>(layers) : RegKernelForT<true, false>(layers);
}

__device__ __forceinline__ uint64_t LocalIndex(const NetStatic& ns, uint32_t W, uint32_t H,
                                               uint32_t cell) {
  const uint32_t x = cell % W, y = (cell / W) % H, l = cell / (W * H);
  return (static_cast<uint64_t>(l) * ns.h + (y - ns.y0)) * ns.w + (x - ns.x0);
}

template <int TX, int TY>
__device__ __forceinline__ uint32_t TileOf(const NetStatic& ns, uint32_t W, uint32_t H,
                                           uint32_t cell) {
  const uint32_t x = cell % W - ns.x0, y = (cell / W) % H - ns.y0;
  return ns.tile_offset + (y / TY) * ns.tiles_x + x / TX;
}

// Seeds pin 0 of every searchable net: initialises its tile (all unreached,
// the pin a source) and queues it for round 1.  One workgroup per net.
template <int TX, int TY>
__global__ __launch_bounds__(kThreads) void SeedKernel(SearchArgs a) {
  const uint32_t n = blockIdx.x;
  if (n >= a.net_count || a.state[n].status != kSearching) return;
  const NetStatic ns = a.nets[n];
  const uint32_t pin = a.pins[ns.pin_offset];
  const uint32_t g = TileOf<TX, TY>(ns, a.W, a.H, pin);
  const uint32_t t = g - ns.tile_offset;
  const uint32_t bx = (t % ns.tiles_x) * TX, by = (t / ns.tiles_x) * TY;
  const uint32_t nx = min(static_cast<uint32_t>(TX), ns.w - bx);
  const uint32_t ny = min(static_cast<uint32_t>(TY), ns.h - by);
  const uint64_t pin_local = LocalIndex(ns, a.W, a.H, pin);
  uint32_t* lab = a.labels + ns.label_offset;
  for (uint32_t i = threadIdx.x; i < a.L * TY * TX; i += kThreads) {
    const uint32_t x = i % TX, y = (i / TX) % TY, l = i / (TX * TY);
    if (x < nx && y < ny) {
      const uint64_t idx = (static_cast<uint64_t>(l) * ns.h + by + y) * ns.w + bx + x;
      lab[idx] = idx == pin_local ? kSourceWord : kInf;
    }
  }
  if (threadIdx.x == 0) {
    a.tile_init[g] = 0;
    a.state[n].connected = 1u;
    Push(a, g, n, 1u, 0u, 0u);
  }
}

__device__ __forceinline__ uint32_t Neighbour(uint32_t cell, uint32_t code, uint32_t W,
                                              uint32_t plane) {
  switch (code) {
    case kFromBelow: return cell - plane;
    case kFromSouth: return cell - W;
    case kFromWest: return cell - 1;
    case kFromEast: return cell + 1;
    case kFromNorth: return cell + W;
    default: return cell + plane;
  }
}

template <int TX, int TY>
__device__ void WalkConnection(const SearchArgs& a, uint32_t r, const NetStatic& ns, NetState& st,
                               DevConnection& c, uint32_t distance, uint32_t start, bool reseed, uint32_t n);

// Runs after every round.  Nets without queued work hold exact distances from
// their current tree: connect the nearest unconnected pin (ties: lowest pin
// index) along the predecessor chain, make that path part of the tree and
// queue its tiles for the next round.  One wave per net: every lane reads the
// net's state, lane 0 writes it, and the walk is wave-cooperative.
template <int TX, int TY>
__device__ void PhaseNet(const SearchArgs& a, uint32_t r, uint32_t n);

template <int TX, int TY>
__global__ void PhaseKernel(SearchArgs a) {
  const uint32_t r = __builtin_amdgcn_readfirstlane(*a.round_ptr) + 1;  // the round new work goes to
  const uint32_t n = blockIdx.x;  // one 64-lane workgroup per net
  if (n < a.net_count) PhaseNet<TX, TY>(a, r, n);
  // The last block to finish advances the device round: every block read it
  // above, before its completion count.
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(&a.round_ptr[1], 1u) == gridDim.x - 1) {
      a.round_ptr[1] = 0;
      atomicAdd(&a.round_ptr[0], 1u);
    }
  }
}

template <int TX, int TY>
__device__ void PhaseNet(const SearchArgs& a, uint32_t r, uint32_t n) {
  const bool lead = (threadIdx.x & 63u) == 0;
  // The key buffer read by the round just executed collects pushes of the
  // round after the next one.
  if (lead) {
    a.net_minkey[((r - 1) & 1) * a.net_count + n] = kInfD;
    if (a.lm_count) a.net_minf[((r - 1) & 1) * a.net_count + n] = kInfD;
  }
  NetState& st = a.state[n];
  if (st.status != kSearching) return;
  const uint32_t pending = a.net_minkey[(r & 1) * a.net_count + n];
  if (a.explore) {
    if (lead && pending >= kInfD) st.status = APGX_NET_DISCONNECTED;
    return;
  }
  const NetStatic ns = a.nets[n];
  uint32_t* lab = a.labels + ns.label_offset;
  const uint32_t plane = a.W * a.H;
  uint32_t best = kInfD, best_pin = kNever;
  for (uint32_t p = 0; p < ns.pin_count; ++p) {
    if (st.connected & (1u << p)) continue;
    const uint32_t cell = a.pins[ns.pin_offset + p];
    const uint32_t g = TileOf<TX, TY>(ns, a.W, a.H, cell);
    const uint32_t d = a.tile_init[g] != kNever ? lab[LocalIndex(ns, a.W, a.H, cell)] >> 3 : kInfD;
    if (d < best) {
      best = d;
      best_pin = p;
    }
  }
  // Pending work can only create labels above its settling keys, so the
  // nearest pin is settled (its distance and every label on its predecessor
  // chain exact) once its distance is at most the smallest pending one.
  if (best_pin == kNever) {
    if (lead && pending >= kInfD) st.status = APGX_NET_DISCONNECTED;
    return;
  }
  if (best > (a.lm_count ? a.net_minf[(r & 1) * a.net_count + n] : pending)) return;
  const uint32_t k = st.n_connected;
  const uint32_t cost = st.cost, connected = st.connected;
  const bool last = k + 2 == ns.pin_count;
  DevConnection& c = a.conns[ns.conn_offset + k];
  WaveSync();  // every lane has read the state before lane 0 updates it
  if (lead) {
    c.pin_index = best_pin;
    c.distance = best;
    c.cell_offset = 0;
    c.cell_count = 0;
    c.reserved = 0;
    st.cost = cost + best;
    st.connected = connected | (1u << best_pin);
    st.n_connected = k + 1;
    // Nothing depends on the last path: it is written after the round loop so
    // this (per-round) kernel stays short.
    if (last) st.status = kFinalWalk;
  }
  if (last) return;
  WalkConnection<TX, TY>(a, r, ns, st, c, best, a.pins[ns.pin_offset + best_pin], true, n);
}

// Writes the predecessor chain of a settled pin back to the tree (distance 0)
// in one pass into a reservation of min(distance / min_edge + 1, window cells)
// cells, which bounds the chain since every step costs at least min_edge; when
// `reseed`, the path cells become sources of the next connection step and
// their tiles are queued.  The step bound and window checks guard against a
// corrupted word only.
//
// One wave walks a path, segment by segment: the wave loads the TX x TY plane
// of the tile and layer holding the current cell into LDS with coalesced row
// loads, lane 0 follows the chain inside it at LDS latency, recording the
// cells' segment indices, until the chain leaves the segment or ends, and all
// lanes then write the segment's path cells and reseed labels together.  The
// tiles to queue (each change of tile along the path) collect in LDS and are
// pushed by all lanes together.  A path thus costs one global round trip per
// segment instead of one per cell, with the same cells, checks and pushes as
// a cell-by-cell walk.
__device__ __forceinline__ uint64_t ShflU64(uint64_t v, int src) {
  const uint32_t lo = __shfl(static_cast<uint32_t>(v), src, 64);
  const uint32_t hi = __shfl(static_cast<uint32_t>(v >> 32), src, 64);
  return (static_cast<uint64_t>(hi) << 32) | lo;
}

struct WalkState {
  uint32_t x, y, l;  // current cell (lattice coordinates)
  uint32_t pushed;   // last queued tile
  uint64_t len;      // cells written
  uint32_t done, bad;
  uint32_t npush;    // tiles waiting in s_push
};
constexpr uint32_t kWalkPushes = 256;

template <int TX, int TY>
__device__ void WalkConnection(const SearchArgs& a, uint32_t r, const NetStatic& ns, NetState& st,
                               DevConnection& c, uint32_t distance, uint32_t start, bool reseed, uint32_t n) {
  __shared__ uint32_t s_seg[TX * TY];
  __shared__ uint16_t s_list[TX * TY];  // segment-local indices of the chain, in walk order
  __shared__ uint32_t s_push[kWalkPushes];
  __shared__ WalkState s_walk;
  __shared__ uint32_t s_count, s_reseeded;  // chain cells in this segment, and how many are reseeded
  const int lane = static_cast<int>(threadIdx.x & 63u);
  uint32_t* lab = a.labels + ns.label_offset;
  const uint32_t plane = a.W * a.H;
  const uint64_t bound = min(static_cast<uint64_t>(distance) / max(a.min_edge, 1u) + 1,
                             static_cast<uint64_t>(ns.w) * ns.h * a.L);
  unsigned long long offset = 0;
  if (lane == 0) offset = atomicAdd(a.cell_cursor, static_cast<unsigned long long>(bound));
  offset = ShflU64(offset, 0);
  if (offset + bound > a.cell_capacity) {
    if (lane == 0) {
      st.status = APGX_NET_INVALID;
      atomicOr(&a.stats[7], 1ull);  // reported to the caller as resource exhaustion
    }
    return;
  }
  if (lane == 0) s_walk = WalkState{start % a.W, (start / a.W) % a.H, start / plane, kNever, 0, 0, 0, 0};
  WaveSync();
  for (;;) {
    const WalkState w = s_walk;
    if (w.npush == kWalkPushes || ((w.done || w.bad) && w.npush)) {
      for (uint32_t b = 0; b < w.npush; b += 64) {
        const uint32_t i = b + static_cast<uint32_t>(lane);
        PushWave(a, i < w.npush, s_push[min(i, w.npush - 1)], n, r, 0u, 0u);
      }
      WaveSync();
      if (lane == 0) s_walk.npush = 0;
      WaveSync();
      continue;
    }
    if (w.done || w.bad) break;
    // Segment: the tile (window-relative grid) and layer of the current cell.
    const uint32_t bx = ((w.x - ns.x0) / TX) * TX, by = ((w.y - ns.y0) / TY) * TY;
    const uint32_t nx = min(static_cast<uint32_t>(TX), ns.w - bx), ny = min(static_cast<uint32_t>(TY), ns.h - by);
    const uint32_t* seg = lab + (static_cast<uint64_t>(w.l) * ns.h + by) * ns.w + bx;
    for (uint32_t y = 0; y < ny; ++y)
      for (uint32_t x = static_cast<uint32_t>(lane); x < nx; x += 64) s_seg[y * TX + x] = seg[y * ns.w + x];
    WaveSync();
    const uint32_t x_lo = ns.x0 + bx, y_lo = ns.y0 + by;
    // Lane 0 follows the chain inside the segment, recording each cell's
    // segment index: a step that stays in the segment (same layer, inside the
    // tile) needs no other check.  The chase ends at a cell at distance 0 or at
    // the first step that leaves the segment (a via, a tile edge or a bad
    // code), which then takes the window and layer checks.  A chain is simple
    // (distances fall along it), so a step inside the segment after all TX * TY
    // cells were visited marks a corrupted chain.
    if (lane == 0) {
      WalkState u = w;
      // The step is computed without branches (selects on the code), so a
      // cell costs the LDS read and one exit test.
      const uint32_t end = ny * TX;
      uint32_t i = (u.y - y_lo) * TX + (u.x - x_lo), cnt = 0, word = 0, next = kNever;
      for (;;) {
        word = s_seg[i];
        s_list[cnt++] = static_cast<uint16_t>(i);
        const uint32_t code = word & 7u, col = i % TX;
        next = kNever;  // the step leaves the segment (a via, an edge or a bad code)
        next = (code == kFromSouth && i >= TX) ? i - TX : next;
        next = (code == kFromNorth && i + TX < end) ? i + TX : next;
        next = (code == kFromWest && col > 0) ? i - 1 : next;
        next = (code == kFromEast && col + 1 < nx) ? i + 1 : next;
        if ((word >> 3) == 0 || next == kNever || cnt == TX * TY) break;
        i = next;
      }
      const bool cycle = (word >> 3) != 0 && next != kNever;  // only at cnt == TX * TY
      u.x = x_lo + i % TX;
      u.y = y_lo + i / TX;
      // Path cells written and cells reseeded (each cell that takes a valid
      // step onward): the walk writes at most `bound` cells (its reservation).
      const uint64_t room = w.len < bound ? bound - w.len : 0;
      uint32_t reseeded;
      if (cnt > room) {
        cnt = static_cast<uint32_t>(room);  // each of these stepped inside the segment
        reseeded = cnt;
        u.bad = 1;
      } else if ((word >> 3) == 0) {
        reseeded = cnt - 1;
        u.done = 1;
      } else if (cycle) {
        reseeded = cnt;
        u.bad = 1;
      } else {
        const uint32_t code = word & 7u;
        const bool ok = code <= kFromAbove && (code != kFromBelow || u.l > 0) &&
                        (code != kFromAbove || u.l + 1 < a.L) && (code != kFromWest || u.x > ns.x0) &&
                        (code != kFromEast || u.x + 1 < ns.x0 + ns.w) && (code != kFromSouth || u.y > ns.y0) &&
                        (code != kFromNorth || u.y + 1 < ns.y0 + ns.h);
        if (!ok) {
          reseeded = cnt - 1;
          u.bad = 1;
        } else {
          reseeded = cnt;
          switch (code) {
            case kFromBelow: --u.l; break;
            case kFromSouth: --u.y; break;
            case kFromWest: --u.x; break;
            case kFromEast: ++u.x; break;
            case kFromNorth: ++u.y; break;
            default: ++u.l; break;
          }
        }
      }
      // Every cell of a segment lies in one tile: it is queued once, unless it
      // is the tile queued last.
      if (reseed && reseeded > 0) {
        const uint32_t g = ns.tile_offset + (by / TY) * ns.tiles_x + bx / TX;
        if (g != u.pushed) {
          s_push[u.npush++] = g;
          u.pushed = g;
        }
      }
      u.len = w.len + cnt;
      s_count = cnt;
      s_reseeded = reseed ? reseeded : 0;
      s_walk = u;
    }
    WaveSync();
    {
      // Every lane writes its share of the chain: the path cells, and the
      // reseeded cells as sources of the next connection step.
      const uint32_t cnt = s_count, reseeded = s_reseeded;
      for (uint32_t i = static_cast<uint32_t>(lane); i < cnt; i += 64) {
        const uint32_t li = s_list[i];
        const uint32_t lx = li % TX, ly = li / TX;
        a.cells[offset + w.len + i] = (w.l * a.H + y_lo + ly) * a.W + x_lo + lx;
        if (i < reseeded) lab[(static_cast<uint64_t>(w.l) * ns.h + by + ly) * ns.w + bx + lx] = kSourceWord;
      }
    }
    WaveSync();
  }
  if (lane == 0) {
    if (s_walk.bad) {
      st.status = APGX_NET_INVALID;
    } else {
      c.cell_offset = offset;
      c.cell_count = static_cast<uint32_t>(s_walk.len);
    }
  }
}

// After the round loop: paths of nets whose last pin settled.  One thread per net.
template <int TX, int TY>
__global__ void FinalWalkKernel(SearchArgs a) {
  const uint32_t n = blockIdx.x;  // one 64-lane workgroup per net
  if (n >= a.net_count) return;
  NetState& st = a.state[n];
  if (st.status != kFinalWalk) return;
  const NetStatic ns = a.nets[n];
  DevConnection& c = a.conns[ns.conn_offset + st.n_connected - 1];
  const uint32_t distance = c.distance, pin = c.pin_index;
  WaveSync();  // every lane has read the state before lane 0 updates it
  if ((threadIdx.x & 63u) == 0) st.status = APGX_NET_REACHED;
  WalkConnection<TX, TY>(a, 0u, ns, st, c, distance, a.pins[ns.pin_offset + pin], false, n);
}

// Label-range check of a finished search.  Distances saturate at kInfD, which
// reads as unreached, so a true distance of 2^29 - 1 or more would surface as
// an unreached cell.  A reached pin's distance is always exact (every label
// on its shortest path is no larger than its own), so the one result that can
// be wrong is "disconnected".  When a net ends disconnected its queue is
// empty and its labels are the fixed point of the saturated relaxation over
// the window; there a free cell is left unreached next to a reached free
// neighbour exactly when that edge's relaxation saturated.  One block row per
// listed net (grid.y), grid-stride over the window's cells; flags[i] gets 1 if
// net list[i] saturated, 2 if a relaxable edge is left (not a fixed point).
__device__ __forceinline__ uint32_t LabelDistanceAt(const SearchArgs& a, const NetStatic& ns, uint32_t tw,
                                                    uint32_t th, uint32_t x, uint32_t y, uint32_t l) {
  const uint32_t g = ns.tile_offset + (y / th) * ns.tiles_x + x / tw;
  if (a.tile_init[g] == kNever) return kInfD;
  return min(a.labels[ns.label_offset + (static_cast<uint64_t>(l) * ns.h + y) * ns.w + x] >> 3, kInfD);
}

__global__ __launch_bounds__(256) void SaturationCheckKernel(SearchArgs a, uint32_t cost_bytes, uint32_t tw,
                                                             uint32_t th, const uint32_t* __restrict__ list,
                                                             uint32_t* __restrict__ flags) {
  const NetStatic ns = a.nets[list[blockIdx.y]];
  const uint64_t cells = static_cast<uint64_t>(ns.w) * ns.h * a.L;
  auto cost_at = [&](uint32_t x, uint32_t y, uint32_t l) -> uint32_t {
    const uint64_t gi = ns.cost_offset + (static_cast<uint64_t>(l) * a.H + ns.y0 + y) * a.W + ns.x0 + x;
    return cost_bytes == 2 ? LoadCost<uint16_t>(a.cost, gi) : LoadCost<uint32_t>(a.cost, gi);
  };
  uint32_t found = 0;
  for (uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x; i < cells;
       i += static_cast<uint64_t>(gridDim.x) * blockDim.x) {
    const uint32_t x = static_cast<uint32_t>(i % ns.w), y = static_cast<uint32_t>((i / ns.w) % ns.h);
    const uint32_t l = static_cast<uint32_t>(i / (static_cast<uint64_t>(ns.w) * ns.h));
    const uint32_t c = cost_at(x, y, l);
    if (c == 0 || LabelDistanceAt(a, ns, tw, th, x, y, l) < kInfD) continue;
    // Every free in-window neighbour u with a finite label relaxes into this cell.
    auto probe = [&](uint32_t ux, uint32_t uy, uint32_t ul, uint32_t weight) {
      if (cost_at(ux, uy, ul) == 0) return;
      const uint32_t du = LabelDistanceAt(a, ns, tw, th, ux, uy, ul);
      if (du >= kInfD) return;
      found |= static_cast<uint64_t>(du) + weight >= kInfD ? 1u : 2u;
    };
    if (x > 0) probe(x - 1, y, l, c);
    if (x + 1 < ns.w) probe(x + 1, y, l, c);
    if (y > 0) probe(x, y - 1, l, c);
    if (y + 1 < ns.h) probe(x, y + 1, l, c);
    if (a.via != 0) {
      const uint32_t vw = a.via + (a.via_pays_cell ? c : 0u);
      if (l > 0) probe(x, y, l - 1, vw);
      if (l + 1 < a.L) probe(x, y, l + 1, vw);
    }
  }
  if (found) atomicOr(&flags[blockIdx.y], found);
}

// Per pin slot and landmark: d(landmark k, pin) from the landmark fields
// (kInfD for a slot whose cell lies outside the lattice: its net is invalid).
__global__ void LandmarkPinKernel(const uint32_t* __restrict__ pins, uint64_t pin_count,
                                  const uint32_t* __restrict__ dist, uint64_t cells, uint32_t count,
                                  uint32_t* __restrict__ out) {
  const uint64_t i = blockIdx.x * static_cast<uint64_t>(blockDim.x) + threadIdx.x;
  if (i >= pin_count * count) return;
  const uint32_t cell = pins[i / count];
  out[i] = cell < cells ? dist[(i % count) * cells + cell] : kInfD;
}

// Landmark fields from an exploring search whose net k (window: the whole
// lattice, so label index = cell index) started at landmark k.  One workgroup
// per (64 x 64 column, landmark): writes d(landmark, cell) for the column's
// cells (kInfD where the cell is blocked, unreached or its tile never
// initialised) and the largest finite distance (0 if none) over the whole
// column and over each of its four edge lines to colmax.
__global__ __launch_bounds__(256) void LandmarkExtractKernel(const uint32_t* __restrict__ labels,
                                                             const uint32_t* __restrict__ tile_init,
                                                             const NetStatic* __restrict__ nets, uint32_t W,
                                                             uint32_t H, uint32_t L, uint32_t tx, uint32_t ty,
                                                             uint32_t columns_x, uint32_t* __restrict__ dist,
                                                             uint32_t* __restrict__ colmax) {
  __shared__ uint32_t s_max[kLmParts][4];
  const uint32_t k = blockIdx.y, column = blockIdx.x;
  const NetStatic ns = nets[k];
  const uint32_t x0 = (column % columns_x) * kLmColumn, y0 = (column / columns_x) * kLmColumn;
  const uint32_t nx = min(kLmColumn, W - x0), ny = min(kLmColumn, H - y0);
  const uint64_t cells = static_cast<uint64_t>(W) * H * L;
  uint32_t m[kLmParts] = {0, 0, 0, 0, 0};
  for (uint32_t i = threadIdx.x; i < kLmColumn * kLmColumn * L; i += blockDim.x) {
    const uint32_t x = i % kLmColumn, y = (i / kLmColumn) % kLmColumn, l = i / (kLmColumn * kLmColumn);
    if (x >= nx || y >= ny) continue;
    const uint32_t X = x0 + x, Y = y0 + y;
    const uint64_t c = (static_cast<uint64_t>(l) * H + Y) * W + X;
    const uint32_t g = ns.tile_offset + (Y / ty) * ns.tiles_x + X / tx;
    uint32_t d = kInfD;
    if (tile_init[g] != kNever) d = min(labels[ns.label_offset + c] >> 3, kInfD);
    dist[k * cells + c] = d;
    if (d < kInfD) {
      m[0] = max(m[0], d);
      if (x == 0) m[1] = max(m[1], d);
      if (x == nx - 1) m[2] = max(m[2], d);
      if (y == 0) m[3] = max(m[3], d);
      if (y == ny - 1) m[4] = max(m[4], d);
    }
  }
#pragma unroll
  for (uint32_t part = 0; part < kLmParts; ++part) {
    uint32_t v = m[part];
    for (int off = 32; off > 0; off >>= 1) v = max(v, static_cast<uint32_t>(__shfl_xor(static_cast<int>(v), off, 64)));
    if ((threadIdx.x & 63u) == 0) s_max[part][threadIdx.x >> 6] = v;
  }
  __syncthreads();
  if (threadIdx.x < kLmParts) {
    uint32_t v = 0;
    for (uint32_t w = 0; w < blockDim.x / 64; ++w) v = max(v, s_max[threadIdx.x][w]);
    colmax[(static_cast<uint64_t>(k) * gridDim.x + column) * kLmParts + threadIdx.x] = v;
  }
}

struct TileShape {
  int tx, ty;
};

size_t TileLdsBytes(TileShape s, uint32_t layers) {
  return static_cast<size_t>(layers) * ((s.ty + 2) * (s.tx + 3) + s.ty * (s.tx + 1)) * 4;
}

struct Kernels {
  const void* round;
  const void* seed;
  const void* phase;
  const void* final_walk;
};

template <int TX, int TY>
Kernels KernelsFor() {
  return Kernels{reinterpret_cast<const void*>(&RoundKernel<TX, TY>),
                 reinterpret_cast<const void*>(&SeedKernel<TX, TY>),
                 reinterpret_cast<const void*>(&PhaseKernel<TX, TY>),
                 reinterpret_cast<const void*>(&FinalWalkKernel<TX, TY>)};
}

bool KernelsForShape(TileShape s, Kernels* k) {
  if (s.tx == 64 && s.ty == 64) { *k = KernelsFor<64, 64>(); return true; }
  if (s.tx == 64 && s.ty == 32) { *k = KernelsFor<64, 32>(); return true; }
  if (s.tx == 32 && s.ty == 32) { *k = KernelsFor<32, 32>(); return true; }
  if (s.tx == 32 && s.ty == 16) { *k = KernelsFor<32, 16>(); return true; }
  return false;
}

}  // namespace
}  // namespace apgx

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

struct apgx_search {
  apgx_lattice* lattice = nullptr;
  uint32_t net_count = 0;
  std::vector<apgx_net_query> nets;
  std::vector<uint32_t> conn_offset;
  apgx::TileShape shape{};
  void* d_labels = nullptr;
  size_t label_bytes = 0;
  void* d_block = nullptr;   // every other device buffer, one allocation
  size_t block_bytes = 0;
  apgx::NetState* d_state = nullptr;
  apgx::DevConnection* d_conns = nullptr;
  uint32_t* d_cells = nullptr;
  unsigned long long* d_cursor = nullptr;
  const apgx::NetStatic* d_nets = nullptr;  // for ExtractLandmarks
  const uint32_t* d_tile_init = nullptr;
  uint64_t total_connections = 0;
  apgx_route_stats stats{};
  // host copies filled at the end of the search
  std::vector<apgx::NetState> state;
  std::vector<apgx::DevConnection> conns;
  std::vector<uint32_t> cells;
  bool read = false;
};

namespace apgx {
namespace {

int ValidateDesc(const apgx_lattice_desc* d) {
  if (d == nullptr) return Fail(APGX_E_INTERNAL_INVARIANT, "lattice descriptor is null");
  if (d->schema_version != APGX_LATTICE_SCHEMA_VERSION)
    return Fail(APGX_E_UNSUPPORTED, "lattice schema version is not supported");
  if (d->width == 0 || d->height == 0 || d->layers == 0 || d->layers > APGX_MAX_LAYERS)
    return Fail(APGX_E_INTERNAL_INVARIANT, "lattice dimensions are out of range");
  if (d->cost_bytes != 2 && d->cost_bytes != 4)
    return Fail(APGX_E_INTERNAL_INVARIANT, "lattice cost_bytes must be 2 or 4");
  if (d->via_cost > APGX_MAX_VIA_COST)
    return Fail(APGX_E_INTERNAL_INVARIANT, "via cost exceeds APGX_MAX_VIA_COST");
  const unsigned __int128 cells = static_cast<unsigned __int128>(d->width) * d->height * d->layers;
  if (cells >= 0xFFFFFFFFull)
    return Fail(APGX_E_RESOURCE_EXHAUSTED, "lattice cell count exceeds the uint32 index space");
  return APGX_OK;
}

int FinishLattice(const apgx_lattice_desc* desc, const void* host_costs, apgx_lattice* lat) {
  const uint64_t cells = static_cast<uint64_t>(desc->width) * desc->height * desc->layers;
  uint32_t min_cost = 0xFFFFFFFFu, max_cost = 0;
  uint64_t free_cells = 0;
  double cost_sum = 0.0;
  for (uint64_t c = 0; c < cells; ++c) {
    const uint32_t v = desc->cost_bytes == 2 ? static_cast<const uint16_t*>(host_costs)[c]
                                             : static_cast<const uint32_t*>(host_costs)[c];
    if (v == 0) continue;
    if (v > APGX_MAX_CELL_COST)
      return Fail(APGX_E_INTERNAL_INVARIANT, "cell cost exceeds APGX_MAX_CELL_COST");
    ++free_cells;
    cost_sum += v;
    min_cost = std::min(min_cost, v);
    max_cost = std::max(max_cost, v);
  }
  lat->desc = *desc;
  lat->free_cells = free_cells;
  lat->mean_cost = free_cells ? cost_sum / static_cast<double>(free_cells) : 1.0;
  const uint8_t* bytes = static_cast<const uint8_t*>(host_costs);
  lat->host_costs.assign(bytes, bytes + cells * desc->cost_bytes);
  lat->min_edge = min_cost;
  lat->min_cell_cost = free_cells ? min_cost : 0u;
  lat->max_cell_cost = max_cost;
  if (desc->via_cost != 0 && desc->layers > 1) lat->min_edge = std::min(min_cost, desc->via_cost);
  return APGX_OK;
}

}  // namespace
}  // namespace apgx

extern "C" int apgx_lattice_create(const apgx_lattice_desc* desc, const void* host_costs,
                                   apgx_lattice** out) {
  // (apgx_lattice_release is declared in apgx.h and tolerates partial objects.)
  using namespace apgx;
  if (out == nullptr || host_costs == nullptr)
    return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_lattice_create: null argument");
  *out = nullptr;
  if (int s = ValidateDesc(desc)) return s;
  auto* lat = new (std::nothrow) apgx_lattice();
  if (lat == nullptr) return Fail(APGX_E_RESOURCE_EXHAUSTED, "apgx_lattice_create: host allocation");
  if (int s = FinishLattice(desc, host_costs, lat)) {
    delete lat;
    return s;
  }
  const size_t bytes = static_cast<size_t>(desc->width) * desc->height * desc->layers * desc->cost_bytes;
  hipError_t e = hipGetDevice(&lat->device);
  int lds = 0, cus = 0;
  if (e == hipSuccess) e = hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, lat->device);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, lat->device);
  if (e == hipSuccess) e = hipMalloc(&lat->d_cost, bytes);
  if (e == hipSuccess) e = hipMemcpy(lat->d_cost, host_costs, bytes, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&lat->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&lat->events[0]);
  if (e == hipSuccess) e = hipEventCreate(&lat->events[1]);
  if (e != hipSuccess) {
    apgx_lattice_release(lat);
    return HipFail(e, "lattice upload");
  }
  lat->lds_limit = static_cast<size_t>(lds);
  lat->cu_count = static_cast<uint32_t>(cus);
  *out = lat;
  ClearError();
  return APGX_OK;
}

extern "C" int apgx_lattice_create_from_device(const apgx_lattice_desc* desc,
                                               const void* device_costs, apgx_lattice** out) {
  using namespace apgx;
  if (out == nullptr || device_costs == nullptr)
    return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_lattice_create_from_device: null argument");
  if (int s = ValidateDesc(desc)) return s;
  const size_t bytes = static_cast<size_t>(desc->width) * desc->height * desc->layers * desc->cost_bytes;
  std::vector<uint8_t> host(bytes);
  APGX_HIP(hipMemcpy(host.data(), device_costs, bytes, hipMemcpyDeviceToHost), "lattice validation copy");
  return apgx_lattice_create(desc, host.data(), out);
}

extern "C" void apgx_lattice_release(apgx_lattice* lat) {
  if (lat == nullptr) return;
  if (lat->d_cost) (void)hipFree(lat->d_cost);
  if (lat->cached_labels) (void)hipFree(lat->cached_labels);
  if (lat->cached_block) (void)hipFree(lat->cached_block);
  if (lat->d_lm_dist) (void)hipFree(lat->d_lm_dist);
  if (lat->d_lm_colmax) (void)hipFree(lat->d_lm_colmax);
  for (hipEvent_t ev : lat->events)
    if (ev) (void)hipEventDestroy(ev);
  if (lat->stream) (void)hipStreamDestroy(lat->stream);
  delete lat;
}

extern "C" const void* apgx_lattice_device_costs(const apgx_lattice* lat) {
  return lat ? lat->d_cost : nullptr;
}

namespace apgx {
namespace {

// Register-resident 64 x 64 tiles: one wave per layer, up to eight layers,
// cell costs up to kRegMaxCost, or kRegWideMaxCost in allocator searches.
bool RegTileUsable(uint32_t layers, uint32_t max_cost, size_t lds_limit, bool via_pays_cell) {
  const uint32_t limit = via_pays_cell ? kRegWideMaxCost : kRegMaxCost;
  return layers <= 8 && max_cost <= limit && RegTileLdsBytes(layers) <= lds_limit;
}

TileShape DefaultShape(const apgx_lattice_desc& d, uint32_t max_cost, size_t lds_limit, bool via_pays_cell) {
  if (RegTileUsable(d.layers, max_cost, lds_limit, via_pays_cell)) return TileShape{64, 64};
  const TileShape candidates[] = {{64, 64}, {64, 32}, {32, 32}, {32, 16}};
  // Prefer the largest tile that fits the register-resident cell budget and
  // leaves room for two workgroups per CU.
  for (const TileShape& s : candidates) {
    if (static_cast<uint64_t>(s.tx) * s.ty * d.layers > kMaxTileCells) continue;
    if (TileLdsBytes(s, d.layers) * 2 > lds_limit && !(s.tx == 32 && s.ty == 16)) continue;
    return s;
  }
  return TileShape{32, 16};
}

// Occupancy of a kernel at a launch shape, queried once per shape.
int OccupancyOf(const void* kernel, uint32_t block, size_t lds, int* out) {
  struct Key {
    const void* k;
    uint32_t b;
    size_t l;
    int occ;
  };
  static std::mutex mu;
  static std::vector<Key> cache;
  std::lock_guard<std::mutex> lock(mu);
  for (const Key& c : cache)
    if (c.k == kernel && c.b == block && c.l == lds) {
      *out = c.occ;
      return APGX_OK;
    }
  int occ = 1;
  APGX_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kernel, block, lds), "occupancy query");
  cache.push_back(Key{kernel, block, lds, occ});
  *out = occ;
  return APGX_OK;
}

int Launch(const void* kernel, dim3 grid, dim3 block, size_t lds, hipStream_t stream,
           void** args, const char* op) {
  APGX_HIP(hipLaunchKernel(kernel, grid, block, args, lds, stream), op);
  return APGX_OK;
}

}  // namespace
}  // namespace apgx

namespace apgx {
namespace {

// Runs SaturationCheckKernel over every at-risk net that ended disconnected.
// A saturated relaxation makes the verdict unrepresentable here (the distance
// is 2^29 - 1 or more): the search fails with APGX_E_UNSUPPORTED rather than
// report a reachable pin as disconnected.  (The reference's u64 labels reject
// such sums with checked adds, gpu/cuda/planar_router.cu:332-340.)
int CheckLabelRange(const SearchArgs& a, uint32_t cost_bytes, TileShape shape, const std::vector<apgx_net_query>& nets,
                    const std::vector<NetState>& state, const std::vector<uint8_t>& at_risk, hipStream_t stream) {
  std::vector<uint32_t> list;
  uint64_t largest = 0;
  for (uint32_t i = 0; i < state.size(); ++i) {
    if (!at_risk[i] || state[i].status != APGX_NET_DISCONNECTED) continue;
    list.push_back(i);
  }
  if (list.empty()) return APGX_OK;
  for (uint32_t i : list) {
    apgx_window w = nets[i].window;
    if (w.width == 0 && w.height == 0) w = apgx_window{0, 0, a.W, a.H};
    largest = std::max<uint64_t>(largest, static_cast<uint64_t>(w.width) * w.height * a.L);
  }
  uint32_t* d_list = nullptr;
  uint32_t* d_flags = nullptr;
  const size_t n = list.size();
  APGX_HIP(hipMalloc(&d_list, 8 * n), "label range check allocation");
  d_flags = d_list + n;
  std::vector<uint32_t> flags(n, 0);
  hipError_t e = hipMemcpyAsync(d_list, list.data(), 4 * n, hipMemcpyHostToDevice, stream);
  if (e == hipSuccess) e = hipMemsetAsync(d_flags, 0, 4 * n, stream);
  const uint32_t bx = static_cast<uint32_t>(std::min<uint64_t>((largest + 255) / 256, 2048));
  for (size_t off = 0; e == hipSuccess && off < n; off += 65535) {
    const uint32_t rows = static_cast<uint32_t>(std::min<size_t>(n - off, 65535));
    hipLaunchKernelGGL(SaturationCheckKernel, dim3(bx, rows), dim3(256), 0, stream, a, cost_bytes,
                       static_cast<uint32_t>(shape.tx), static_cast<uint32_t>(shape.ty), d_list + off,
                       d_flags + off);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(flags.data(), d_flags, 4 * n, hipMemcpyDeviceToHost, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  (void)hipFree(d_list);
  if (e != hipSuccess) return HipFail(e, "label range check");
  for (size_t k = 0; k < n; ++k) {
    const std::string id = std::to_string(nets[list[k]].net_id);
    if (flags[k] & 1u)
      return Fail(APGX_E_UNSUPPORTED, "apgx_route_nets: net " + id +
                                          ": a distance reaches the 29-bit label range (2^29 - 1)");
    if (flags[k] & 2u)
      return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_route_nets: net " + id +
                                                 ": disconnected labels are not a fixed point");
  }
  return APGX_OK;
}

}  // namespace
}  // namespace apgx

extern "C" int apgx_route_nets(apgx_lattice* lat, const apgx_net_query* nets, uint32_t net_count,
                               const uint32_t* pins, uint64_t pin_count,
                               const apgx_route_options* options, apgx_search** out) {
  return apgx::RouteNets(lat, nullptr, nets, net_count, pins, pin_count, options, out);
}

int apgx::RouteNets(apgx_lattice* lat, const CostOverride* ov, const apgx_net_query* nets,
                    uint32_t net_count, const uint32_t* pins, uint64_t pin_count,
                    const apgx_route_options* options, apgx_search** out, bool explore) {
  using namespace apgx;
  const auto wall_start = std::chrono::steady_clock::now();
  if (lat == nullptr || out == nullptr || (net_count != 0 && (nets == nullptr || pins == nullptr)))
    return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_route_nets: null argument");
  *out = nullptr;
  int device = 0;
  APGX_HIP(hipGetDevice(&device), "hipGetDevice");
  if (device != lat->device)
    return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_route_nets: current device differs from the upload device");
  apgx_route_options opt{};
  if (options) opt = *options;
  const uint32_t max_rounds = opt.max_rounds ? opt.max_rounds : 100000u;
  const apgx_lattice_desc& d = lat->desc;
  const size_t lds_limit = lat->lds_limit;
  const uint32_t max_cost = ov ? ov->max_cost : lat->max_cell_cost;
  const uint32_t cost_bytes = ov ? 4u : lat->desc.cost_bytes;
  const bool pays = ov != nullptr && ov->via_pays_cell;
  TileShape shape = DefaultShape(lat->desc, max_cost, lds_limit, pays);
  if (opt.tile_w || opt.tile_h) shape = TileShape{static_cast<int>(opt.tile_w), static_cast<int>(opt.tile_h)};
  Kernels k{};
  if (!KernelsForShape(shape, &k))
    return Fail(APGX_E_UNSUPPORTED, "apgx_route_nets: unsupported tile shape");
  const bool reg_tile = shape.tx == 64 && shape.ty == 64 && RegTileUsable(d.layers, max_cost, lds_limit, pays);
  uint32_t block = kThreads;
  size_t lds = TileLdsBytes(shape, d.layers);
  if (reg_tile) {
    k.round = RegKernelFor(d.layers, pays, max_cost);
    block = 64 * d.layers;
    lds = RegTileLdsBytes(d.layers);
  } else if (static_cast<uint64_t>(shape.tx) * shape.ty * d.layers > kMaxTileCells) {
    return Fail(APGX_E_UNSUPPORTED, "apgx_route_nets: tile holds more cells than a workgroup keeps");
  }
  if (lds > lds_limit)
    return Fail(APGX_E_UNSUPPORTED, "apgx_route_nets: tile does not fit in LDS");

  auto* s = new (std::nothrow) apgx_search();
  if (s == nullptr) return Fail(APGX_E_RESOURCE_EXHAUSTED, "apgx_route_nets: host allocation");
  s->lattice = lat;
  s->net_count = net_count;
  s->nets.assign(nets, nets + net_count);
  s->shape = shape;

  // Host preflight: windows, pins, offsets.  Invalid nets are reported, not searched.
  std::vector<NetStatic> ns(net_count);
  std::vector<NetState> st(net_count);
  std::vector<uint32_t> tile_net;
  uint64_t label_words = 0, tiles = 0, conns = 0, window_cells_total = 0;
  const uint64_t plane = static_cast<uint64_t>(d.width) * d.height;
  // Largest edge weight: a simple path through a window of c cells costs at
  // most (c - 1) times it.  Nets whose bound reaches kInfD can saturate a
  // label, and a "disconnected" verdict of theirs is checked after the search.
  const uint64_t via_weight = (d.layers > 1 && d.via_cost)
                                  ? d.via_cost + ((ov && ov->via_pays_cell) ? static_cast<uint64_t>(max_cost) : 0u)
                                  : 0u;
  const uint64_t max_weight = std::max<uint64_t>(max_cost, via_weight);
  std::vector<uint8_t> at_risk(net_count, 0);
  for (uint32_t i = 0; i < net_count; ++i) {
    const apgx_net_query& q = nets[i];
    NetStatic& n = ns[i];
    std::memset(&n, 0, sizeof(n));
    std::memset(&st[i], 0, sizeof(st[i]));
    st[i].status = kSearching;
    apgx_window win = q.window;
    if (win.width == 0 && win.height == 0) win = apgx_window{0, 0, d.width, d.height};
    bool ok = win.width > 0 && win.height > 0 && win.x0 + static_cast<uint64_t>(win.width) <= d.width &&
              win.y0 + static_cast<uint64_t>(win.height) <= d.height && q.pin_count >= 2 &&
              q.pin_count <= APGX_MAX_NET_PINS &&
              static_cast<uint64_t>(q.pin_offset) + q.pin_count <= pin_count;
    if (ok) {
      for (uint32_t p = 0; p < q.pin_count && ok; ++p) {
        const uint32_t c = pins[q.pin_offset + p];
        if (c >= plane * d.layers) { ok = false; break; }
        const uint32_t x = static_cast<uint32_t>(c % d.width), y = static_cast<uint32_t>((c / d.width) % d.height);
        ok = x >= win.x0 && x < win.x0 + win.width && y >= win.y0 && y < win.y0 + win.height;
        for (uint32_t p2 = 0; p2 < p && ok; ++p2) ok = pins[q.pin_offset + p2] != c;
      }
    }
    n.x0 = win.x0; n.y0 = win.y0; n.w = win.width; n.h = win.height;
    n.pin_offset = q.pin_offset; n.pin_count = q.pin_count;
    n.conn_offset = static_cast<uint32_t>(conns);
    n.cost_offset = (ov && ov->net_offset) ? ov->net_offset[i] : 0u;
    if (!ok) {
      st[i].status = APGX_NET_INVALID;
      continue;
    }
    n.tiles_x = (n.w + shape.tx - 1) / shape.tx;
    n.tiles_y = (n.h + shape.ty - 1) / shape.ty;
    n.tile_offset = static_cast<uint32_t>(tiles);
    n.label_offset = label_words;
    const uint64_t nt = static_cast<uint64_t>(n.tiles_x) * n.tiles_y;
    tile_net.insert(tile_net.end(), nt, i);
    tiles += nt;
    label_words += static_cast<uint64_t>(n.w) * n.h * d.layers;
    window_cells_total += static_cast<uint64_t>(n.w) * n.h * d.layers * (q.pin_count - 1);
    const unsigned __int128 path_bound =
        static_cast<unsigned __int128>(static_cast<uint64_t>(n.w) * n.h * d.layers - 1) * max_weight;
    at_risk[i] = path_bound >= kInfD ? 1u : 0u;
    conns += q.pin_count - 1;
  }
  if (tiles >= 0xFFFFFFF0ull) {
    delete s;
    return Fail(APGX_E_RESOURCE_EXHAUSTED, "apgx_route_nets: tile count exceeds the uint32 index space");
  }
  s->total_connections = conns;
  s->conn_offset.resize(net_count);
  for (uint32_t i = 0; i < net_count; ++i) s->conn_offset[i] = ns[i].conn_offset;
  const uint64_t cell_capacity = std::max<uint64_t>(1u << 20, std::min<uint64_t>(window_cells_total, 1ull << 28));
  // Device layout: one block for everything except the label pool.
  const uint64_t counter_words = 2ull * (static_cast<uint64_t>(max_rounds) + 2);
  struct Part { size_t bytes; size_t offset; };
  auto align = [](size_t v) { return (v + 255) & ~static_cast<size_t>(255); };
  Part p_nets{sizeof(NetStatic) * std::max<uint32_t>(net_count, 1), 0};
  Part p_state{sizeof(NetState) * std::max<uint32_t>(net_count, 1), 0};
  Part p_tile_net{4 * std::max<uint64_t>(tiles, 1), 0};
  Part p_stamp{4 * std::max<uint64_t>(tiles, 1), 0};
  Part p_init{4 * std::max<uint64_t>(tiles, 1), 0};
  Part p_wl0{4 * std::max<uint64_t>(tiles, 1), 0};
  Part p_wl1{4 * std::max<uint64_t>(tiles, 1), 0};
  Part p_counters{4 * counter_words, 0};
  Part p_stats{8 * kStatSlots, 0};
  Part p_pins{4 * std::max<uint64_t>(pin_count, 1), 0};
  Part p_conns{sizeof(DevConnection) * std::max<uint64_t>(conns, 1), 0};
  Part p_cursor{8, 0};
  Part p_cells{4 * cell_capacity, 0};
  Part p_tile_key{8 * std::max<uint64_t>(tiles, 1), 0};
  Part p_net_key{8 * std::max<uint32_t>(net_count, 1), 0};
  Part p_tile_fkey{8 * std::max<uint64_t>(tiles, 1), 0};
  Part p_net_fkey{8 * std::max<uint32_t>(net_count, 1), 0};
  Part p_round{8, 0};
  const bool landmarks = opt.landmarks != 0 && !opt.no_goal_bound && !explore;
  Part p_lm_pin{landmarks ? 4 * std::max<uint64_t>(pin_count, 1) * APGX_MAX_LANDMARKS : 4, 0};
  Part* parts[] = {&p_nets, &p_state, &p_tile_net, &p_stamp, &p_init, &p_wl0, &p_wl1,
                   &p_counters, &p_stats, &p_pins, &p_conns, &p_cursor, &p_cells,
                   &p_tile_key, &p_net_key, &p_round, &p_lm_pin, &p_tile_fkey, &p_net_fkey};
  size_t total = 0;
  for (Part* p : parts) { p->offset = total; total += align(p->bytes); }
  const size_t label_bytes = static_cast<size_t>(std::max<uint64_t>(label_words, 1)) * 4;
  if (opt.max_device_bytes && total + label_bytes > opt.max_device_bytes) {
    delete s;
    return Fail(APGX_E_RESOURCE_EXHAUSTED, "apgx_route_nets: workspace exceeds max_device_bytes");
  }
  {
    std::lock_guard<std::mutex> lock(lat->cache_mutex);
    if (lat->cached_block && lat->cached_block_bytes >= total) {
      s->d_block = lat->cached_block;
      s->block_bytes = lat->cached_block_bytes;
      lat->cached_block = nullptr;
      lat->cached_block_bytes = 0;
    }
    if (lat->cached_labels && lat->cached_label_bytes >= label_bytes) {
      s->d_labels = lat->cached_labels;
      s->label_bytes = lat->cached_label_bytes;
      lat->cached_labels = nullptr;
      lat->cached_label_bytes = 0;
    }
  }
  hipError_t e = hipSuccess;
  if (s->d_block == nullptr) {
    e = hipMalloc(&s->d_block, total);
    if (e != hipSuccess) { delete s; return HipFail(e, "search workspace allocation"); }
    s->block_bytes = total;
  }
  if (s->d_labels == nullptr) {
    e = hipMalloc(&s->d_labels, label_bytes);
    if (e != hipSuccess) { (void)hipFree(s->d_block); s->d_block = nullptr; delete s; return HipFail(e, "label pool allocation"); }
    s->label_bytes = label_bytes;
  }
  auto at = [&](const Part& p) { return static_cast<char*>(s->d_block) + p.offset; };
  SearchArgs a{};
  a.cost = ov ? static_cast<const void*>(ov->plane) : lat->d_cost;
  a.W = d.width; a.H = d.height; a.L = d.layers; a.via = d.layers > 1 ? d.via_cost : 0;
  a.nets = reinterpret_cast<NetStatic*>(at(p_nets));
  a.state = reinterpret_cast<NetState*>(at(p_state));
  a.net_count = net_count;
  a.tile_net = reinterpret_cast<uint32_t*>(at(p_tile_net));
  a.pins = reinterpret_cast<uint32_t*>(at(p_pins));
  a.labels = static_cast<uint32_t*>(s->d_labels);
  a.tile_stamp = reinterpret_cast<uint32_t*>(at(p_stamp));
  a.tile_init = reinterpret_cast<uint32_t*>(at(p_init));
  a.wl[0] = reinterpret_cast<uint32_t*>(at(p_wl0));
  a.wl[1] = reinterpret_cast<uint32_t*>(at(p_wl1));
  a.counters = reinterpret_cast<uint32_t*>(at(p_counters));
  a.stats = reinterpret_cast<unsigned long long*>(at(p_stats));
  a.conns = reinterpret_cast<DevConnection*>(at(p_conns));
  a.cells = reinterpret_cast<uint32_t*>(at(p_cells));
  a.cell_cursor = reinterpret_cast<unsigned long long*>(at(p_cursor));
  a.cell_capacity = cell_capacity;
  // Default visit length: register tiles stop after kRegDefaultSweeps sweeps
  // and re-queue (a tile's halos usually improve again before a longer visit
  // would converge); with landmarks the visited tiles lie along the routes,
  // where longer visits pay; LDS tiles run to convergence.
  a.max_sweeps = opt.max_sweeps ? opt.max_sweeps
                                 : (reg_tile ? (landmarks ? kRegLandmarkSweeps : kRegDefaultSweeps) : 64u);
  a.tile_key = reinterpret_cast<uint32_t*>(at(p_tile_key));
  a.net_minkey = reinterpret_cast<uint32_t*>(at(p_net_key));
  a.tile_fkey = reinterpret_cast<uint32_t*>(at(p_tile_fkey));
  a.net_minf = reinterpret_cast<uint32_t*>(at(p_net_fkey));
  a.round_ptr = reinterpret_cast<uint32_t*>(at(p_round));
  a.tile_count = static_cast<uint32_t>(tiles);
  a.min_edge = std::max(lat->min_edge, 1u);
  a.goal_cost = (opt.no_goal_bound || explore) ? 0u : lat->min_cell_cost;
  a.explore = explore ? 1u : 0u;
  a.via_pays_cell = (ov && ov->via_pays_cell) ? 1u : 0u;
  a.sweep_mode = opt.sweep_schedule == 0 ? kDefaultSweepSchedule : opt.sweep_schedule;
  // Bucket width: about three quarters of a tile width of distance by
  // default; all-ones disables ordering (every queued tile every round).
  if (opt.frontier_delta == 0xFFFFFFFFu) {
    a.delta = kInfD;
  } else if (opt.frontier_delta != 0) {
    a.delta = std::min(opt.frontier_delta, kInfD);
  } else {
    a.delta = static_cast<uint32_t>(std::max(1.0, kDefaultDeltaCells * lat->mean_cost));
  }
  s->d_state = a.state;
  s->d_conns = a.conns;
  s->d_cells = a.cells;
  s->d_cursor = a.cell_cursor;
  s->d_nets = a.nets;
  s->d_tile_init = a.tile_init;

  std::unique_lock<std::mutex> lease(lat->exec_mutex);
  // Landmark fields are replaced only under the lease, so they stay put for
  // the whole search.
  if (landmarks) {
    if (lat->lm_count == 0) {
      lease.unlock();
      apgx_search_release(s);
      return Fail(APGX_E_UNSUPPORTED, "apgx_route_nets: landmarks requested but none are built");
    }
    a.lm_pin = reinterpret_cast<uint32_t*>(at(p_lm_pin));
    a.lm_colmax = lat->d_lm_colmax;
    a.lm_count = lat->lm_count;
    a.lm_columns_x = (d.width + kLmColumn - 1) / kLmColumn;
    a.lm_columns = a.lm_columns_x * ((d.height + kLmColumn - 1) / kLmColumn);
    a.lm_weight = opt.landmark_weight ? std::min(opt.landmark_weight, 16u) : kDefaultLandmarkWeight;
  }
  hipStream_t stream = lat->stream;
  hipEvent_t ev0 = lat->events[0], ev1 = lat->events[1];
  int status = APGX_OK;
  auto check = [&](hipError_t err, const char* op) {
    if (status == APGX_OK && err != hipSuccess) status = HipFail(err, op);
    return status == APGX_OK;
  };
  check(hipMemcpyAsync(const_cast<NetStatic*>(a.nets), ns.data(), sizeof(NetStatic) * net_count, hipMemcpyHostToDevice, stream), "net upload");
  check(hipMemcpyAsync(a.state, st.data(), sizeof(NetState) * net_count, hipMemcpyHostToDevice, stream), "net state upload");
  if (!tile_net.empty())
    check(hipMemcpyAsync(const_cast<uint32_t*>(a.tile_net), tile_net.data(), 4 * tile_net.size(), hipMemcpyHostToDevice, stream), "tile map upload");
  if (pin_count)
    check(hipMemcpyAsync(const_cast<uint32_t*>(a.pins), pins, 4 * pin_count, hipMemcpyHostToDevice, stream), "pin upload");
  check(hipMemsetAsync(a.tile_stamp, 0, 4 * std::max<uint64_t>(tiles, 1), stream), "stamp clear");
  check(hipMemsetAsync(a.tile_init, 0xFF, 4 * std::max<uint64_t>(tiles, 1), stream), "init clear");
  check(hipMemsetAsync(a.counters, 0, 4 * counter_words, stream), "counter clear");
  check(hipMemsetAsync(a.stats, 0, 8 * kStatSlots, stream), "stats clear");
  check(hipMemsetAsync(a.cell_cursor, 0, 8, stream), "cursor clear");
  check(hipMemsetAsync(a.tile_key, 0xFF, 8 * std::max<uint64_t>(tiles, 1), stream), "key clear");
  check(hipMemsetAsync(a.net_minkey, 0xFF, 8 * std::max<uint32_t>(net_count, 1), stream), "key clear");
  check(hipMemsetAsync(a.tile_fkey, 0xFF, 8 * std::max<uint64_t>(tiles, 1), stream), "key clear");
  check(hipMemsetAsync(a.net_minf, 0xFF, 8 * std::max<uint32_t>(net_count, 1), stream), "key clear");
  check(hipMemsetD32Async(reinterpret_cast<hipDeviceptr_t>(a.round_ptr), 1u, 1, stream), "round init");
  check(hipMemsetD32Async(reinterpret_cast<hipDeviceptr_t>(a.round_ptr + 1), 0u, 1, stream), "round init");
  if (status == APGX_OK && landmarks && pin_count) {
    const uint64_t slots = pin_count * a.lm_count;
    LandmarkPinKernel<<<dim3(static_cast<uint32_t>((slots + 255) / 256)), dim3(256), 0, stream>>>(
        a.pins, pin_count, lat->d_lm_dist, plane * d.layers, a.lm_count, const_cast<uint32_t*>(a.lm_pin));
    check(hipGetLastError(), "landmark pin launch");
  }
  if (status == APGX_OK && lds > 65536)
    check(hipFuncSetAttribute(k.round, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds)), "hipFuncSetAttribute");

  // Pins must be free cells (checked on the host copy of the plane).
  if (status == APGX_OK && pin_count) {
    for (uint32_t i = 0; i < net_count; ++i) {
      if (st[i].status != kSearching) continue;
      for (uint32_t p = 0; p < ns[i].pin_count; ++p) {
        if (lat->HostCost(pins[ns[i].pin_offset + p]) == 0) { st[i].status = APGX_NET_INVALID; break; }
      }
    }
    check(hipMemcpyAsync(a.state, st.data(), sizeof(NetState) * net_count, hipMemcpyHostToDevice, stream), "net state upload");
  }

  int occupancy = 1;
  if (status == APGX_OK) status = OccupancyOf(k.round, block, lds, &occupancy);
  occupancy = std::max(occupancy, 1);
  const uint32_t grid = static_cast<uint32_t>(std::max<uint64_t>(1, std::min<uint64_t>(
      static_cast<uint64_t>(lat->cu_count) * occupancy, std::max<uint64_t>(tiles, 1))));
  const uint32_t chunk = opt.rounds_per_sync ? opt.rounds_per_sync : 8;
  uint32_t r = 1, syncs = 0, graph_launches = 0;
  std::vector<hipEvent_t> round_events;
  if (status == APGX_OK && net_count) {
    check(hipEventRecord(ev0, stream), "hipEventRecord");
    {
      void* args[] = {&a};
      if (status == APGX_OK) status = Launch(k.seed, dim3(net_count), dim3(kThreads), 0, stream, args, "seed launch");
    }
    bool done = false;
    hipGraphExec_t graph_exec = nullptr;
    // One chunk of rounds: round kernel then phase kernel, per round.  The
    // kernels take the round from device memory, so the same launches replay.
    auto launch_chunk = [&](uint32_t rounds, bool timed) {
      for (uint32_t c = 0; c < rounds && status == APGX_OK; ++c) {
        uint32_t cb = cost_bytes;
        void* rargs[] = {&a, &cb};
        if (timed) {
          hipEvent_t e0 = nullptr, e1 = nullptr;
          if (!check(hipEventCreate(&e0), "hipEventCreate") || !check(hipEventCreate(&e1), "hipEventCreate")) return;
          round_events.push_back(e0);
          round_events.push_back(e1);
          check(hipEventRecord(e0, stream), "hipEventRecord");
        }
        status = Launch(k.round, dim3(grid), dim3(block), lds, stream, rargs, "round launch");
        if (status != APGX_OK) return;
        if (timed) check(hipEventRecord(round_events.back(), stream), "hipEventRecord");
        void* pargs[] = {&a};
        status = Launch(k.phase, dim3(net_count), dim3(64), 0, stream, pargs, "phase launch");
      }
    };
    while (status == APGX_OK && !done) {
      // Rounds 1..max_rounds only: the counter array holds max_rounds + 2 rounds.
      const uint32_t this_chunk = std::min(chunk, max_rounds + 1 - r);
      if (opt.use_graph && !opt.profile_rounds && this_chunk == chunk) {
        if (graph_exec == nullptr) {
          hipGraph_t graph = nullptr;
          if (!check(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture")) break;
          launch_chunk(chunk, false);
          const bool captured = check(hipStreamEndCapture(stream, &graph), "hipStreamEndCapture");
          if (status != APGX_OK || !captured) {
            if (graph) (void)hipGraphDestroy(graph);
            break;
          }
          const bool ok = check(hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0), "hipGraphInstantiate");
          (void)hipGraphDestroy(graph);
          if (!ok) break;
        }
        if (!check(hipGraphLaunch(graph_exec, stream), "hipGraphLaunch")) break;
        ++graph_launches;
      } else {
        launch_chunk(this_chunk, opt.profile_rounds != 0);
      }
      r += this_chunk;
      uint32_t next_count = 0;
      if (!check(hipMemcpyAsync(&next_count, a.counters + 2ull * r, 4, hipMemcpyDeviceToHost, stream), "status readback")) break;
      if (!check(hipStreamSynchronize(stream), "hipStreamSynchronize")) break;
      ++syncs;
      if (next_count == 0) done = true;
      else if (r > max_rounds) done = true;  // budget: remaining searching nets stop
    }
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (status == APGX_OK) {
      void* fargs[] = {&a};
      status = Launch(k.final_walk, dim3(net_count), dim3(64), 0, stream, fargs, "final walk launch");
    }
    check(hipEventRecord(ev1, stream), "hipEventRecord");
    check(hipStreamSynchronize(stream), "hipStreamSynchronize");
  }
  // Results to host.
  s->state.resize(net_count);
  if (status == APGX_OK && net_count)
    check(hipMemcpy(s->state.data(), a.state, sizeof(NetState) * net_count, hipMemcpyDeviceToHost), "state readback");
  s->conns.resize(conns);
  if (status == APGX_OK && conns)
    check(hipMemcpy(s->conns.data(), a.conns, sizeof(DevConnection) * conns, hipMemcpyDeviceToHost), "connection readback");
  if (status == APGX_OK) status = CheckLabelRange(a, cost_bytes, shape, s->nets, s->state, at_risk, stream);
  // Net cost: the sum of the connection distances, checked in 64 bits (the
  // device's running u32 sum is only a hint).
  for (uint32_t i = 0; status == APGX_OK && i < net_count; ++i) {
    NetState& x = s->state[i];
    if (x.status != APGX_NET_REACHED) continue;
    uint64_t sum = 0;
    for (uint32_t c = 0; c < x.n_connected; ++c) sum += s->conns[ns[i].conn_offset + c].distance;
    if (sum >= APGX_UNREACHABLE) {
      status = Fail(APGX_E_UNSUPPORTED, "apgx_route_nets: net " + std::to_string(s->nets[i].net_id) +
                                            ": the sum of its connection distances exceeds the uint32 net cost");
      break;
    }
    x.cost = static_cast<uint32_t>(sum);
  }
  unsigned long long used = 0, stats[kStatSlots] = {0};
  if (status == APGX_OK) {
    check(hipMemcpy(&used, a.cell_cursor, 8, hipMemcpyDeviceToHost), "cursor readback");
    check(hipMemcpy(stats, a.stats, 8 * kStatSlots, hipMemcpyDeviceToHost), "stats readback");
  }
  if (status == APGX_OK) {
    used = std::min<unsigned long long>(used, cell_capacity);
    s->cells.resize(used);
    if (used) check(hipMemcpy(s->cells.data(), a.cells, 4 * used, hipMemcpyDeviceToHost), "cell readback");
  }
  float ms = 0.f;
  if (status == APGX_OK && net_count) check(hipEventElapsedTime(&ms, ev0, ev1), "hipEventElapsedTime");
  double round_ms = 0.0;
  for (size_t i = 0; i + 1 < round_events.size(); i += 2) {
    float x = 0.f;
    if (status == APGX_OK && check(hipEventElapsedTime(&x, round_events[i], round_events[i + 1]), "hipEventElapsedTime"))
      round_ms += x;
  }
  for (hipEvent_t ev : round_events) (void)hipEventDestroy(ev);
  lease.unlock();
  if (status == APGX_OK && stats[7] != 0)
    status = Fail(APGX_E_RESOURCE_EXHAUSTED, "apgx_route_nets: connection cell buffer exhausted");
  if (status != APGX_OK) {
    apgx_search_release(s);
    return status;
  }
  for (auto& x : s->state)
    if (x.status == kSearching || x.status == kFinalWalk) x.status = APGX_NET_BUDGET_EXHAUSTED;
  apgx_route_stats& rs = s->stats;
  rs.device_ms = ms;
  rs.tile_visits = stats[0];
  rs.cell_visits = stats[1];
  rs.inner_sweeps = stats[2];
  rs.tiles_touched = stats[4];
  rs.relaxations = stats[3];
  rs.layer_sweeps = stats[5];
  rs.full_sweeps = stats[6];
  rs.changing_full_sweeps = stats[8];
  rs.free_cell_visits = stats[9];
  rs.algorithmic_bytes = stats[1] * (static_cast<uint64_t>(cost_bytes) + 8ull);
  rs.workspace_bytes = total + s->label_bytes;
  rs.rounds = r - 1;
  rs.tile_w = static_cast<uint32_t>(shape.tx);
  rs.tile_h = static_cast<uint32_t>(shape.ty);
  rs.host_syncs = syncs;
  rs.graph_launches = graph_launches;
  rs.round_kernel_ms = round_ms;
  rs.wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - wall_start).count();
  *out = s;
  ClearError();
  return APGX_OK;
}

extern "C" int apgx_search_sizes(const apgx_search* s, uint64_t* connection_count, uint64_t* cell_count) {
  using namespace apgx;
  if (s == nullptr || connection_count == nullptr || cell_count == nullptr)
    return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_search_sizes: null argument");
  // Connections of net i sit at its conn_offset (pin_count - 1 slots each).
  uint64_t nc = 0, cells = 0;
  for (uint32_t i = 0; i < s->net_count; ++i) {
    const NetState& st = s->state[i];
    if (st.status == APGX_NET_INVALID) continue;
    for (uint32_t c = 0; c < st.n_connected; ++c) cells += s->conns[s->conn_offset[i] + c].cell_count;
    nc += st.n_connected;
  }
  *connection_count = nc;
  *cell_count = cells;
  ClearError();
  return APGX_OK;
}

extern "C" int apgx_search_readback(apgx_search* s, apgx_net_result* results,
                                    apgx_connection* connections, uint64_t connection_capacity,
                                    uint32_t* cells, uint64_t cell_capacity) {
  using namespace apgx;
  if (s == nullptr || (s->net_count && results == nullptr))
    return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_search_readback: null argument");
  uint64_t nc = 0, ncell = 0;
  for (uint32_t i = 0; i < s->net_count; ++i) {
    const NetState& st = s->state[i];
    apgx_net_result& r = results[i];
    std::memset(&r, 0, sizeof(r));
    r.net_id = s->nets[i].net_id;
    r.status = st.status;
    r.rounds = st.rounds;
    r.connection_offset = static_cast<uint32_t>(nc);
    const bool valid = st.status != APGX_NET_INVALID;
    r.cost = st.status == APGX_NET_REACHED ? st.cost : APGX_UNREACHABLE;
    if (!valid) continue;
    r.connection_count = st.n_connected;
    for (uint32_t c = 0; c < st.n_connected; ++c) {
      const DevConnection& dc = s->conns[s->conn_offset[i] + c];
      if (nc >= connection_capacity || (connections == nullptr))
        return Fail(APGX_E_RESOURCE_EXHAUSTED, "apgx_search_readback: connection capacity too small");
      if (ncell + dc.cell_count > cell_capacity || (cells == nullptr && dc.cell_count))
        return Fail(APGX_E_RESOURCE_EXHAUSTED, "apgx_search_readback: cell capacity too small");
      if (dc.cell_offset + dc.cell_count > s->cells.size())
        return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_search_readback: device path outside the cell buffer");
      apgx_connection& out = connections[nc++];
      out.pin_index = dc.pin_index;
      out.distance = dc.distance;
      out.cell_offset = ncell;
      out.cell_count = dc.cell_count;
      out.reserved = 0;
      std::memcpy(cells + ncell, s->cells.data() + dc.cell_offset, 4ull * dc.cell_count);
      ncell += dc.cell_count;
    }
  }
  s->read = true;
  ClearError();
  return APGX_OK;
}

extern "C" int apgx_search_stats(const apgx_search* s, apgx_route_stats* out) {
  using namespace apgx;
  if (s == nullptr || out == nullptr) return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_search_stats: null argument");
  *out = s->stats;
  ClearError();
  return APGX_OK;
}

extern "C" void apgx_search_release(apgx_search* s) {
  if (s == nullptr) return;
  if (s->d_block) {
    apgx_lattice* lat = s->lattice;
    std::lock_guard<std::mutex> lock(lat->cache_mutex);
    if (lat->cached_block == nullptr || lat->cached_block_bytes < s->block_bytes) {
      if (lat->cached_block) (void)hipFree(lat->cached_block);
      lat->cached_block = s->d_block;
      lat->cached_block_bytes = s->block_bytes;
    } else {
      (void)hipFree(s->d_block);
    }
  }
  if (s->d_labels) {
    apgx_lattice* lat = s->lattice;
    std::lock_guard<std::mutex> lock(lat->cache_mutex);
    if (lat->cached_labels == nullptr) {
      lat->cached_labels = s->d_labels;
      lat->cached_label_bytes = s->label_bytes;
    } else if (lat->cached_label_bytes < s->label_bytes) {
      (void)hipFree(lat->cached_labels);
      lat->cached_labels = s->d_labels;
      lat->cached_label_bytes = s->label_bytes;
    } else {
      (void)hipFree(s->d_labels);
    }
  }
  delete s;
}

extern "C" int apgx_autotune(apgx_lattice* lat, const apgx_net_query* nets, uint32_t net_count,
                             const uint32_t* pins, uint64_t pin_count, uint32_t sample_nets,
                             const apgx_route_options* base, apgx_route_options* tuned,
                             apgx_tune_report* report) {
  using namespace apgx;
  if (lat == nullptr || tuned == nullptr || (net_count && (nets == nullptr || pins == nullptr)))
    return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_autotune: null argument");
  apgx_route_options opt{};
  if (base) opt = *base;
  opt.profile_rounds = 0;
  *tuned = opt;
  const uint32_t want = sample_nets ? sample_nets : 256u;
  const uint32_t sample = std::min(net_count, want);
  std::vector<apgx_net_query> picked(sample);
  for (uint32_t i = 0; i < sample; ++i)
    picked[i] = nets[static_cast<uint64_t>(i) * net_count / std::max(sample, 1u)];
  const size_t lds_limit = lat->lds_limit;
  const bool reg = RegTileUsable(lat->desc.layers, lat->max_cell_cost, lds_limit, false);
  const uint32_t mean = static_cast<uint32_t>(std::max(1.0, lat->mean_cost));
  std::vector<apgx_tune_candidate> cands;
  const uint32_t deltas[] = {16 * mean, 48 * mean, 256 * mean, 0xFFFFFFFFu};
  const TileShape shapes[] = {{64, 64}, {64, 32}, {32, 32}};
  for (const TileShape& t : shapes) {
    if (t.tx == 64 && t.ty == 64 && !reg) continue;
    if (!(t.tx == 64 && t.ty == 64) &&
        (static_cast<uint64_t>(t.tx) * t.ty * lat->desc.layers > kMaxTileCells ||
         TileLdsBytes(t, lat->desc.layers) > lds_limit))
      continue;
    for (uint32_t dl : deltas) {
      cands.push_back(apgx_tune_candidate{static_cast<uint32_t>(t.tx), static_cast<uint32_t>(t.ty), dl, 64, 0.0});
      if (t.tx == 64 && t.ty == 64 && reg) {
        cands.push_back(apgx_tune_candidate{64, 64, dl, 16, 0.0});
        cands.push_back(apgx_tune_candidate{64, 64, dl, kRegDefaultSweeps, 0.0});
        if (opt.landmarks) cands.push_back(apgx_tune_candidate{64, 64, dl, 32, 0.0});
      }
    }
  }
  if (cands.size() > APGX_TUNE_MAX_CANDIDATES) cands.resize(APGX_TUNE_MAX_CANDIDATES);
  uint32_t best = 0;
  for (uint32_t i = 0; i < cands.size(); ++i) {
    apgx_route_options o = opt;
    o.tile_w = cands[i].tile_w;
    o.tile_h = cands[i].tile_h;
    o.frontier_delta = cands[i].frontier_delta;
    o.max_sweeps = cands[i].max_sweeps;
    // Best of two runs: a single timing of a short sample is noisy enough to
    // swap neighbouring candidates.
    cands[i].device_ms = 0.0;
    for (int rep = 0; rep < 2; ++rep) {
      apgx_search* s = nullptr;
      const int rc = apgx_route_nets(lat, picked.data(), sample, pins, pin_count, &o, &s);
      if (rc != APGX_OK) return rc;
      const double ms = s->stats.device_ms;
      apgx_search_release(s);
      cands[i].device_ms = rep == 0 ? ms : std::min(cands[i].device_ms, ms);
    }
    if (cands[i].device_ms < cands[best].device_ms) best = i;
  }
  if (!cands.empty()) {
    tuned->tile_w = cands[best].tile_w;
    tuned->tile_h = cands[best].tile_h;
    tuned->frontier_delta = cands[best].frontier_delta;
    tuned->max_sweeps = cands[best].max_sweeps;
  }
  if (report) {
    std::memset(report, 0, sizeof(*report));
    report->sample_nets = sample;
    report->candidate_count = static_cast<uint32_t>(cands.size());
    report->best = best;
    for (uint32_t i = 0; i < cands.size(); ++i) report->candidates[i] = cands[i];
  }
  ClearError();
  return APGX_OK;
}

// ---------------------------------------------------------------------------
// Landmark fields for the ALT goal bound
// ---------------------------------------------------------------------------

namespace apgx {
namespace {

// Free cell nearest (x, y) by Chebyshev rings, layer 0 first within a ring;
// kNever when the lattice has none.
uint32_t NearestFree(const apgx_lattice* lat, uint32_t x, uint32_t y) {
  const apgx_lattice_desc& d = lat->desc;
  const int64_t W = d.width, H = d.height, cx = x, cy = y;
  const int64_t reach = std::max(W, H);
  for (int64_t r = 0; r <= reach; ++r) {
    for (uint32_t l = 0; l < d.layers; ++l) {
      auto probe = [&](int64_t px, int64_t py) -> uint32_t {
        if (px < 0 || py < 0 || px >= W || py >= H) return kNever;
        const uint64_t c = (static_cast<uint64_t>(l) * H + py) * W + px;
        return lat->HostCost(c) ? static_cast<uint32_t>(c) : kNever;
      };
      for (int64_t dx = -r; dx <= r; ++dx) {
        if (uint32_t c = probe(cx + dx, cy - r); c != kNever) return c;
        if (uint32_t c = probe(cx + dx, cy + r); c != kNever) return c;
      }
      for (int64_t dy = -r + 1; dy <= r - 1; ++dy) {
        if (uint32_t c = probe(cx - r, cy + dy); c != kNever) return c;
        if (uint32_t c = probe(cx + r, cy + dy); c != kNever) return c;
      }
    }
  }
  return kNever;
}

}  // namespace
}  // namespace apgx

extern "C" int apgx_lattice_build_landmarks(apgx_lattice* lat, uint32_t count, double* build_ms) {
  using namespace apgx;
  if (build_ms) *build_ms = 0.0;
  if (lat == nullptr) return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_lattice_build_landmarks: null lattice");
  if (count > APGX_MAX_LANDMARKS)
    return Fail(APGX_E_UNSUPPORTED, "apgx_lattice_build_landmarks: count exceeds APGX_MAX_LANDMARKS");
  int device = 0;
  APGX_HIP(hipGetDevice(&device), "hipGetDevice");
  if (device != lat->device)
    return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_lattice_build_landmarks: current device differs from the upload device");
  auto replace = [&](uint32_t n, uint32_t* dist, uint32_t* colmax) {
    std::lock_guard<std::mutex> lease(lat->exec_mutex);
    if (lat->d_lm_dist) (void)hipFree(lat->d_lm_dist);
    if (lat->d_lm_colmax) (void)hipFree(lat->d_lm_colmax);
    lat->d_lm_dist = dist;
    lat->d_lm_colmax = colmax;
    lat->lm_count = n;
  };
  if (count == 0) {
    replace(0, nullptr, nullptr);
    ClearError();
    return APGX_OK;
  }
  if (lat->free_cells < 2)
    return Fail(APGX_E_UNSUPPORTED, "apgx_lattice_build_landmarks: the lattice has fewer than two free cells");
  const apgx_lattice_desc& d = lat->desc;
  const uint32_t W = d.width, H = d.height;
  const uint32_t spots[16][2] = {{0, 0}, {W - 1, H - 1}, {W - 1, 0}, {0, H - 1},
                                 {W / 2, 0}, {W / 2, H - 1}, {0, H / 2}, {W - 1, H / 2},
                                 {W / 4, 0}, {3 * (W / 4), H - 1}, {W - 1, H / 4}, {0, 3 * (H / 4)},
                                 {3 * (W / 4), 0}, {W / 4, H - 1}, {W - 1, 3 * (H / 4)}, {0, H / 4}};
  std::vector<uint32_t> marks;
  for (const auto& sp : spots) {
    if (marks.size() == count) break;
    const uint32_t c = NearestFree(lat, sp[0], sp[1]);
    if (c != kNever && std::find(marks.begin(), marks.end(), c) == marks.end()) marks.push_back(c);
  }
  // Second pin of each exploring net: any other free cell (never settled).
  const uint64_t cells = static_cast<uint64_t>(W) * H * d.layers;
  uint32_t free0 = kNever, free1 = kNever;
  for (uint64_t c = 0; c < cells && free1 == kNever; ++c) {
    if (lat->HostCost(c) == 0) continue;
    if (free0 == kNever) free0 = static_cast<uint32_t>(c);
    else free1 = static_cast<uint32_t>(c);
  }
  const uint32_t n = static_cast<uint32_t>(marks.size());
  std::vector<apgx_net_query> q(n);
  std::vector<uint32_t> pins(2 * n);
  for (uint32_t k = 0; k < n; ++k) {
    q[k] = apgx_net_query{k, 2 * k, 2, apgx_window{0, 0, 0, 0}};
    pins[2 * k] = marks[k];
    pins[2 * k + 1] = marks[k] != free0 ? free0 : free1;
  }
  apgx_route_options opt{};
  opt.no_goal_bound = 1;
  apgx_search* s = nullptr;
  if (int rc = RouteNets(lat, nullptr, q.data(), n, pins.data(), pins.size(), &opt, &s, true); rc != APGX_OK)
    return rc;
  for (uint32_t k = 0; k < n; ++k) {
    if (s->state[k].status != APGX_NET_DISCONNECTED) {
      apgx_search_release(s);
      return Fail(APGX_E_RESOURCE_EXHAUSTED, "apgx_lattice_build_landmarks: an exploring search did not finish");
    }
  }
  const uint32_t columns_x = (W + kLmColumn - 1) / kLmColumn;
  const uint32_t columns = columns_x * ((H + kLmColumn - 1) / kLmColumn);
  uint32_t* dist = nullptr;
  uint32_t* colmax = nullptr;
  hipError_t e = hipMalloc(&dist, 4 * cells * n);
  if (e == hipSuccess) e = hipMalloc(&colmax, 4ull * columns * n * kLmParts);
  if (e != hipSuccess) {
    if (dist) (void)hipFree(dist);
    apgx_search_release(s);
    return HipFail(e, "landmark field allocation");
  }
  float extract_ms = 0.f;
  {
    std::lock_guard<std::mutex> lease(lat->exec_mutex);
    hipStream_t stream = lat->stream;
    e = hipEventRecord(lat->events[0], stream);
    if (e == hipSuccess) {
      LandmarkExtractKernel<<<dim3(columns, n), dim3(256), 0, stream>>>(
          static_cast<const uint32_t*>(s->d_labels), s->d_tile_init, s->d_nets, W, H, d.layers,
          static_cast<uint32_t>(s->shape.tx), static_cast<uint32_t>(s->shape.ty), columns_x, dist, colmax);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(lat->events[1], stream);
    if (e == hipSuccess) e = hipStreamSynchronize(stream);
    if (e == hipSuccess) e = hipEventElapsedTime(&extract_ms, lat->events[0], lat->events[1]);
  }
  const double search_ms = s->stats.device_ms;
  apgx_search_release(s);
  if (e != hipSuccess) {
    (void)hipFree(dist);
    (void)hipFree(colmax);
    return HipFail(e, "landmark extraction");
  }
  replace(n, dist, colmax);
  if (build_ms) *build_ms = search_ms + extract_ms;
  ClearError();
  return APGX_OK;
}

extern "C" uint32_t apgx_lattice_landmark_count(const apgx_lattice* lat) {
  return lat ? lat->lm_count : 0u;
}

extern "C" int apgx_lattice_landmark_field(apgx_lattice* lat, uint32_t index, uint32_t* out, uint64_t cell_count) {
  using namespace apgx;
  if (lat == nullptr || out == nullptr) return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_lattice_landmark_field: null argument");
  const uint64_t cells = static_cast<uint64_t>(lat->desc.width) * lat->desc.height * lat->desc.layers;
  std::lock_guard<std::mutex> lease(lat->exec_mutex);
  if (index >= lat->lm_count) return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_lattice_landmark_field: no such field");
  if (cell_count < cells) return Fail(APGX_E_INTERNAL_INVARIANT, "apgx_lattice_landmark_field: output too small");
  APGX_HIP(hipMemcpy(out, lat->d_lm_dist + index * cells, 4 * cells, hipMemcpyDeviceToHost), "landmark field readback");
  ClearError();
  return APGX_OK;
}
