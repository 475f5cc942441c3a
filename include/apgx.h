/*
 * apgx.h — C-ABI of the MI355X (gfx950) batched lattice search library.
 *
 * Everything that crosses this boundary is a plain C type: fixed-size POD
 * records, pointers and sizes.  No C++ or torch types appear in signatures.
 * Every entry point returns an int status:
 *
 *   0 APGX_OK
 *   1 APGX_E_UNSUPPORTED         BackendErrorCode::kUnsupported
 *   2 APGX_E_RESOURCE_EXHAUSTED  BackendErrorCode::kResourceExhausted
 *   3 APGX_E_BACKEND_FAILURE     BackendErrorCode::kBackendFailure
 *   4 APGX_E_INTERNAL_INVARIANT  BackendErrorCode::kInternalInvariant
 *
 * The four error codes keep the meaning of the reference's BackendErrorCode
 * (include/apgar/gpu/planar_router.h:475-485): malformed requests and foreign
 * handles are internal invariants, allocation failure and byte budgets are
 * resource exhaustion, other HIP runtime failures are backend failures whose
 * detail reads "<operation> failed: <hip message>" (gpu/cuda/planar_router.cu:67-78).
 * apgx_last_error() returns the calling thread's last detail string.
 *
 * Two record families cross the boundary:
 *
 *  - V1: byte-identical restatements of the reference's DeviceCompiledBoardV1 and
 *    DeviceCandidateBatch records (include/apgar/gpu/planar_router.h:170-423),
 *    so a host that already builds them can hand them over unchanged.
 *  - V2 ("lattice"): dense multi-layer planes with per-cell costs, vias, search
 *    windows and multi-pin nets, which V1 cannot express (V1 is planar, has
 *    uniform step costs, no windows and two-terminal queries only).
 */
#ifndef APGX_H_
#define APGX_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define APGX_OK 0
#define APGX_E_UNSUPPORTED 1
#define APGX_E_RESOURCE_EXHAUSTED 2
#define APGX_E_BACKEND_FAILURE 3
#define APGX_E_INTERNAL_INVARIANT 4

/* ------------------------------------------------------------------------ */
/* Library / device                                                         */
/* ------------------------------------------------------------------------ */

typedef struct apgx_device_info {
  char backend[32];         /* "hip" */
  char device_name[256];
  char gcn_arch_name[64];   /* e.g. "gfx950:sramecc+:xnack-" */
  uint32_t compute_units;
  uint32_t runtime_version;
  uint32_t driver_version;
  uint32_t reserved;
  uint64_t global_memory_bytes;
} apgx_device_info;

/* Replaces IPlanarRouteBackend::QueryMetadata (include/apgar/gpu/planar_router.h:550). */
int apgx_device_info_query(apgx_device_info* out);

/* Candidate device class of the current device, the HIP counterpart of
 * CudaCandidateDeviceClass (src/candidates/gpu_candidate_adapter.cc:27-35):
 * "hip-<gfx arch>", e.g. "hip-gfx950".  Fails (APGX_E_UNSUPPORTED) on
 * incomplete metadata. */
int apgx_device_class(char* out, uint64_t capacity);

/* Thread-local detail of the last failing call on this thread ("" if none). */
const char* apgx_last_error(void);

/* Library ABI version (major << 16 | minor). */
uint32_t apgx_abi_version(void);

/* Blocks until all work the library queued on the current device is done. */
int apgx_device_synchronize(void);

/* ------------------------------------------------------------------------ */
/* V2 dense lattice                                                          */
/* ------------------------------------------------------------------------ */
/*
 * Cells are indexed c = (layer * height + y) * width + x.  cost[c] == 0 marks a
 * blocked cell; free cells have cost in [1, APGX_MAX_CELL_COST].  Edges:
 *   in-layer 4-neighbour move u -> v     weight cost[v]  (cost of entering v)
 *   via (x,y,l) <-> (x,y,l+1)            weight via_cost (both cells free)
 * Blocked cells have no edges.  A search window [x0,x0+w) x [y0,y0+h) applies
 * to every layer; cells outside it do not exist for that net.
 *
 * Canonical path rule (the dense restatement of the reference's lowest
 * predecessor rule, gpu/cuda/planar_router.cu:1084-1165): walking back from a
 * reached cell v, the predecessor is the in-window free neighbour u with the
 * lowest cell index such that dist[u] + w(u->v) == dist[v].
 *
 * Multi-pin nets: the tree starts as {pin 0}; each connection step runs an
 * exact multi-source search from every tree cell, connects the unconnected pin
 * with the smallest distance (ties: lowest pin index) along its canonical path
 * back to the tree, and adds that path to the tree.  A two-pin net is one
 * source/target query.
 */
#define APGX_LATTICE_SCHEMA_VERSION 2u
#define APGX_MAX_CELL_COST 1048575u      /* 2^20 - 1 */
#define APGX_MAX_VIA_COST 1048575u
#define APGX_MAX_LAYERS 16u
#define APGX_MAX_NET_PINS 32u
#define APGX_UNREACHABLE 0xFFFFFFFFu

typedef struct apgx_lattice_desc {
  uint32_t schema_version;  /* APGX_LATTICE_SCHEMA_VERSION */
  uint32_t width;
  uint32_t height;
  uint32_t layers;
  uint32_t cost_bytes;      /* 2 (uint16 plane) or 4 (uint32 plane) */
  uint32_t via_cost;        /* 0 disables vias */
  uint32_t reserved[2];
} apgx_lattice_desc;

typedef struct apgx_lattice apgx_lattice;

/* DeviceDenseBoardV2 (schemas/device_dense_board/v2.md): the sealed wire form
 * of a dense board, a 128-byte header followed by the cost plane
 * (cost_bytes per cell, cell order above).  The producer fills the
 * association and shape fields and calls apgx_v2_board_seal, which derives the
 * cost summary, the persistent bytes and the FNV-1a device view fingerprint;
 * the consumer's apgx_v2_board_validate / apgx_lattice_create_v2 re-derive all
 * of them and reject any difference, as DeviceCompiledBoardV1 uploads are
 * re-checked (gpu/cuda/planar_router.cu:1256-1451).  V1 records are unchanged. */
typedef struct apgx_v2_board_header {
  uint32_t schema_version;                 /* APGX_LATTICE_SCHEMA_VERSION */
  uint32_t compiler_version;
  uint64_t source_board_content_hash;
  uint64_t compiler_profile_fingerprint;
  uint64_t rule_bucket_identity;
  uint64_t device_view_fingerprint;        /* derived */
  uint64_t estimated_persistent_device_bytes;  /* derived: 128 + plane bytes */
  int64_t lattice_origin_x;                /* exact coordinate of cell (0, 0) */
  int64_t lattice_origin_y;
  int64_t lattice_step;                    /* > 0 */
  uint32_t width;
  uint32_t height;
  uint32_t layers;
  uint32_t cost_bytes;                     /* 2 or 4 */
  uint32_t via_cost;
  uint32_t min_cell_cost;                  /* derived: smallest free cost, 0 without free cells */
  uint32_t max_cell_cost;                  /* derived */
  uint32_t reserved0;                      /* 0 */
  uint64_t free_cells;                     /* derived */
  uint8_t reserved[16];                    /* 0 */
} apgx_v2_board_header;

uint64_t apgx_v2_board_fingerprint(const apgx_v2_board_header* header, const void* plane);
int apgx_v2_board_seal(apgx_v2_board_header* header, const void* plane, uint64_t plane_bytes);
int apgx_v2_board_validate(const apgx_v2_board_header* header, const void* plane, uint64_t plane_bytes);
/* Validates a sealed board, then uploads it as apgx_lattice_create does. */
int apgx_lattice_create_v2(const apgx_v2_board_header* header, const void* plane, uint64_t plane_bytes,
                           apgx_lattice** out);

/* Validates and uploads one immutable cost plane (host memory, cost_bytes per
 * cell).  The handle owns the device copy until apgx_lattice_release. */
int apgx_lattice_create(const apgx_lattice_desc* desc, const void* host_cost_plane,
                        apgx_lattice** out);
/* Same, from a device-resident plane that the library copies (no PCIe). */
int apgx_lattice_create_from_device(const apgx_lattice_desc* desc, const void* device_cost_plane,
                                    apgx_lattice** out);
void apgx_lattice_release(apgx_lattice* lattice);
/* Device address of the lattice's cost plane (read-only for callers). */
const void* apgx_lattice_device_costs(const apgx_lattice* lattice);

/* Landmark distance fields for the ALT goal bound (no counterpart in the
 * reference, whose searches are single-query A*: an optional accelerator of
 * apgx_route_nets that changes no result).  Builds `count` (1 ..
 * APGX_MAX_LANDMARKS) fields d(landmark, cell) on the lattice's own cost
 * plane with vias costing via_cost (the cheapest metric any search on this
 * lattice uses: override planes of the allocator cost at least as much), by
 * exploring searches of this library, plus the per-column (64 x 64 cells, all
 * layers) maxima of the finite distances.  Landmarks are the free cells
 * nearest the board's corners, then its edge midpoints, then its edge
 * quarter points, on layer 0 first.
 * Replaces earlier fields; count 0 frees them.  *build_ms (optional) receives
 * the device time of the build.  Device memory: 4 bytes per cell per field. */
#define APGX_MAX_LANDMARKS 16u
int apgx_lattice_build_landmarks(apgx_lattice* lattice, uint32_t count, double* build_ms);
/* Number of landmark fields built (0: none). */
uint32_t apgx_lattice_landmark_count(const apgx_lattice* lattice);
/* Copies field `index` (cells in lattice order; 0x1FFFFFFF: blocked or
 * unreachable) into host memory of `cell_count` entries. */
int apgx_lattice_landmark_field(apgx_lattice* lattice, uint32_t index, uint32_t* out, uint64_t cell_count);

typedef struct apgx_window {
  uint32_t x0;
  uint32_t y0;
  uint32_t width;   /* 0 with height 0 selects the whole lattice */
  uint32_t height;
} apgx_window;

typedef struct apgx_net_query {
  uint64_t net_id;
  uint32_t pin_offset;      /* into the pin cell array */
  uint32_t pin_count;       /* 2 .. APGX_MAX_NET_PINS */
  apgx_window window;
} apgx_net_query;

#define APGX_NET_REACHED 0u
#define APGX_NET_DISCONNECTED 1u
#define APGX_NET_BUDGET_EXHAUSTED 2u
#define APGX_NET_INVALID 3u

typedef struct apgx_net_result {
  uint64_t net_id;
  uint32_t status;            /* APGX_NET_* */
  uint32_t cost;              /* sum of connection distances; APGX_UNREACHABLE unless reached */
  uint32_t connection_count;  /* pin_count - 1 when reached */
  uint32_t connection_offset; /* into the connection array (result order) */
  uint32_t rounds;            /* sweep rounds in which this net had work */
  uint32_t reserved;
} apgx_net_result;

typedef struct apgx_connection {
  uint32_t pin_index;     /* pin joined by this connection */
  uint32_t distance;      /* its exact multi-source distance from the tree */
  uint64_t cell_offset;   /* into the cell array */
  uint32_t cell_count;    /* cells from the pin back to the tree cell it joins, inclusive */
  uint32_t reserved;
} apgx_connection;

typedef struct apgx_route_options {
  uint32_t max_rounds;        /* 0: 100,000 (PlanarRoutePolicy::maximum_rounds) */
  uint32_t tile_w;            /* 0: autotuned (also tile_h) */
  uint32_t tile_h;
  uint32_t rounds_per_sync;   /* rounds launched between host convergence checks; 0: auto */
  uint64_t max_device_bytes;  /* 0: unlimited (workspace budget) */
  uint32_t use_graph;         /* 1: launch round chunks as a captured hipGraph */
  uint32_t profile_rounds;    /* 1: time every round kernel with hipEvents (round_kernel_ms) */
  uint32_t frontier_delta;    /* distance bucket width of tile scheduling; 0: 48 x mean cell
                                 cost; 0xFFFFFFFF: unordered (every queued tile every round) */
  uint32_t max_sweeps;        /* sweeps per tile visit before re-queueing; 0: 10 on register
                                 tiles (64 with landmarks), 64 on LDS tiles */
  uint32_t no_goal_bound;     /* 1: order tiles by distance only; 0: distance plus the
                                 consistent lower bound to the nearest unconnected pin */
  uint32_t sweep_schedule;    /* register tiles: k = one full sweep (via exchange and
                                 stability test) every k sweeps, the rest in-layer only;
                                 1 every sweep full; 0 default (2) */
  uint32_t landmarks;         /* 1: raise the goal bound with the ALT bound of the
                                 lattice's landmark fields (apgx_lattice_build_landmarks;
                                 APGX_E_UNSUPPORTED when none are built); 0: off; ignored
                                 with no_goal_bound.
                                 Distances and paths are the same either way. */
  uint32_t landmark_weight;   /* with landmarks: tiles are ordered by distance plus
                                 landmark_weight / 16 of the bound (1 .. 16; 0: 16), pins
                                 settle on distance plus the whole bound */
  uint32_t reserved;
} apgx_route_options;

typedef struct apgx_route_stats {
  double device_ms;              /* hipEvent envelope: first search kernel to last */
  double wall_ms;                /* host wall time of apgx_route_nets */
  uint64_t tile_visits;          /* tiles processed over all rounds */
  uint64_t cell_visits;          /* tile_visits * cells per tile (window clipped) */
  uint64_t algorithmic_bytes;    /* cell_visits * (cost bytes + 4 label read + 4 label write) */
  uint64_t relaxations;          /* directed edges relaxed into free cells by the sweeps executed:
                                    4 in-layer per free cell and sweep, plus one per adjacent
                                    layer on full sweeps */
  uint64_t inner_sweeps;         /* sweep iterations summed over tile visits (register tiles: including
                                    the in-layer sweeps that quiet layers skip) */
  uint64_t workspace_bytes;      /* device bytes owned by the search */
  uint32_t rounds;               /* global rounds executed */
  uint32_t phases;               /* connection phases executed */
  uint32_t tile_w;
  uint32_t tile_h;
  uint32_t host_syncs;
  uint32_t graph_launches;
  double round_kernel_ms;        /* summed hipEvent time of the round kernels */
  uint64_t tiles_touched;        /* distinct tiles visited (tile_visits counts revisits) */
  uint64_t layer_sweeps;         /* sweeps executed, summed over each visited tile's layers */
  uint64_t full_sweeps;          /* sweeps with the via exchange and the stability test */
  uint64_t changing_full_sweeps; /* full sweeps whose stability test found a changed label */
  uint64_t free_cell_visits;     /* free cells of the visited tiles, summed over visits */
} apgx_route_stats;

typedef struct apgx_search apgx_search;

/* Routes a batch of nets (host arrays) on one lattice and keeps the results on
 * the device until readback.  Replaces the execution half of
 * IPlanarRouteBackend::ExecuteCandidateBatch for V2 lattices. */
int apgx_route_nets(apgx_lattice* lattice, const apgx_net_query* nets, uint32_t net_count,
                    const uint32_t* pins, uint64_t pin_count, const apgx_route_options* options,
                    apgx_search** out);
/* Sizes needed by apgx_search_readback. */
int apgx_search_sizes(const apgx_search* search, uint64_t* connection_count,
                      uint64_t* cell_count);
/* Copies results in net order; connections and cells are packed in net order. */
int apgx_search_readback(apgx_search* search, apgx_net_result* results,
                         apgx_connection* connections, uint64_t connection_capacity,
                         uint32_t* cells, uint64_t cell_capacity);
int apgx_search_stats(const apgx_search* search, apgx_route_stats* out);
void apgx_search_release(apgx_search* search);

/* Autotuner.  Times a fixed set of candidate tile shapes and tile schedules
 * (frontier_delta, max_sweeps) on an evenly strided sample of the nets and
 * writes the fastest candidate into *tuned; every other field is copied from
 * *base (NULL: defaults).  The set: tiles 64x64 (register-resident, when the
 * lattice allows), 64x32 and 32x32 (LDS); delta 16, 48 and 256 x mean cell
 * cost and unordered; max_sweeps 64, 16 and 10 on 64x64 tiles (and 32 when
 * base->landmarks is set).  Each candidate is timed twice, the faster run
 * kept.  Results do not depend on the choice (labels are the unique fixed
 * point); only speed does.
 * No reference counterpart: the reference has one fixed launch shape. */
#define APGX_TUNE_MAX_CANDIDATES 24u
typedef struct apgx_tune_candidate {
  uint32_t tile_w;
  uint32_t tile_h;
  uint32_t frontier_delta;
  uint32_t max_sweeps;
  double device_ms;            /* search time of the sample (faster of two runs) */
} apgx_tune_candidate;

typedef struct apgx_tune_report {
  uint32_t sample_nets;
  uint32_t candidate_count;
  uint32_t best;               /* index of the chosen candidate */
  uint32_t reserved;
  apgx_tune_candidate candidates[APGX_TUNE_MAX_CANDIDATES];
} apgx_tune_report;

int apgx_autotune(apgx_lattice* lattice, const apgx_net_query* nets, uint32_t net_count,
                  const uint32_t* pins, uint64_t pin_count, uint32_t sample_nets,
                  const apgx_route_options* base, apgx_route_options* tuned,
                  apgx_tune_report* report);

/* ------------------------------------------------------------------------ */
/* Negotiated-congestion allocator                                          */
/* ------------------------------------------------------------------------ */
/* Canonical Sequential Rip-Up-and-Reroute v1 on the dense lattice with one
 * resource per cell.  Replaces RouteCpuSequentialNegotiated
 * (src/allocator/cpu_sequential_negotiated_routing.cc:993-1329) for this
 * lattice model; the policy scalars are CanonicalSequentialRipUpAndReroutePolicyV1
 * (include/apgar/allocator/cpu_sequential_negotiated_routing.h:38-56) and the
 * value bound is CpuSequentialNegotiatedRoutingLimits::maximum_congestion_cost_value
 * (:73).  Per pass p the present factor is initial + p * present_increment;
 * nets are ripped up and rerouted in order, batch_nets at a time against one
 * snapshot of cell cost
 *     base + history + present_factor * max(0, usage + 1 - capacity)
 * (usage without the batch's own routes); a disconnected net keeps its
 * incumbent.  At each pass boundary: feasible (every net routed, zero
 * overuse), else stalled (every route unchanged), else pass limit, else
 * history += historical_increment * max(0, usage - capacity).  A congestion
 * cost above maximum_congestion_cost_value, or a cell cost above
 * APGX_MAX_CELL_COST, stops the run with APGX_E_RESOURCE_EXHAUSTED (the
 * reference's kBoundExhausted).  batch_nets = 1 is the reference's order
 * exactly, and so is batch_nets = 0: nets are grouped into levels of pairwise
 * disjoint windows (a net's level is one above the highest level of an
 * earlier net whose window overlaps it), so a level routed at once equals its
 * nets routed one after another, with less latency.  All arithmetic is integer
 * and runs on the GPU.  A pass whose present factor is zero prices no cell by
 * its usage, so its snapshot costs do not depend on any route: its nets are
 * routed as one batch in every mode, with the same routes as in any order.
 *
 * Speculative groups (batch_nets = 0 and group_nets > 0): the
 * exact order is kept, but group_nets consecutive nets are searched in one
 * launch, each on its own snapshot of its window taken after the rip-ups of
 * the group's earlier nets and before their installs.  Nets then commit in
 * order.  Between that snapshot and the exact one (every earlier net
 * committed) costs can only rise, and only at cells an earlier group member
 * installed; a search result does not change when costs rise at cells off
 * its tree, so a net whose tree avoids every such cell commits its result,
 * and a net whose tree meets one is searched again on the exact snapshot.
 * Routes, planes, passes and termination equal the one-net-at-a-time order,
 * and so do the bound checks: they are evaluated per net, at each net's own
 * rip-up, over the whole plane (first net of a pass) and then over the cells
 * the previous net installed, the only cells whose cost can have risen. */
#define APGX_NEGOTIATION_FEASIBLE 0u
#define APGX_NEGOTIATION_STALLED 1u
#define APGX_NEGOTIATION_PASS_LIMIT 2u

/* Progress of a negotiation, reported after every completed pass. */
typedef void (*apgx_negotiation_pass_fn)(void* context, uint32_t pass, uint64_t overuse_units,
                                         uint32_t routed_nets, double wall_ms);

typedef struct apgx_negotiation_policy {
  uint32_t maximum_passes;                 /* default 8 */
  uint32_t batch_nets;                     /* default 1; 0: window-disjoint levels (exact order) */
  uint64_t initial_present_cost;           /* default 0 */
  uint64_t present_cost_increment;         /* default 100 */
  uint64_t historical_cost_increment;      /* default 10 */
  uint64_t maximum_congestion_cost_value;  /* default 10,000,000 */
  apgx_negotiation_pass_fn on_pass;        /* optional, called on the calling thread (default NULL) */
  void* on_pass_context;
  uint32_t group_nets;                     /* batch_nets = 0: nets per speculative group; 0 (default):
                                              window-disjoint levels */
  uint32_t reserved;
} apgx_negotiation_policy;

typedef struct apgx_negotiation_summary {
  uint32_t termination;                    /* APGX_NEGOTIATION_* */
  uint32_t completed_passes;
  uint32_t routed_nets;
  uint32_t batches;                        /* snapshots routed */
  uint64_t total_overuse_units;            /* at the last completed pass */
  uint64_t tile_visits;                    /* summed over every search */
  double search_device_ms;                 /* summed search envelopes */
  double wall_ms;                          /* host wall time of apgx_negotiate */
  uint64_t search_rounds;                  /* search rounds summed over every batch */
  uint64_t group_routed_nets;              /* net searches run in speculative groups */
  uint64_t redone_nets;                    /* of those, searched again on the exact snapshot */
} apgx_negotiation_summary;

typedef struct apgx_negotiation apgx_negotiation;

/* Fills the reference defaults. */
void apgx_negotiation_policy_defaults(apgx_negotiation_policy* policy);
/* capacity: one byte per cell (NULL: capacity 1 everywhere); initial_history:
 * one u32 per cell seeding the history plane (NULL: zero, the reference's
 * pass-zero state); search_options: passed to every search (NULL: defaults).
 * Searches of at most 8 nets, which leave the GPU nearly idle, take a bucket
 * 16 times the default width and check for their end every 16 rounds where
 * frontier_delta and rounds_per_sync are 0; neither changes a route. */
int apgx_negotiate(apgx_lattice* lattice, const uint8_t* capacity, const apgx_net_query* nets,
                   uint32_t net_count, const uint32_t* pins, uint64_t pin_count,
                   const apgx_negotiation_policy* policy, const uint32_t* initial_history,
                   const apgx_route_options* search_options, apgx_negotiation** out);
/* overuse_per_pass (may be NULL) receives min(capacity, completed_passes) values. */
int apgx_negotiation_get_summary(const apgx_negotiation* negotiation, apgx_negotiation_summary* out,
                                 uint64_t* overuse_per_pass, uint32_t capacity);
/* Final history and usage planes (one u32 per cell each; either may be NULL). */
int apgx_negotiation_planes(const apgx_negotiation* negotiation, uint32_t* history,
                            uint32_t* usage);
/* Routes: route_len[n] tree cells per net (0: never routed), cells packed in
 * net order: pin 0, then each connection path from its pin up to (not
 * including) the tree cell it reaches.  *cell_count gets the total. */
int apgx_negotiation_routes(const apgx_negotiation* negotiation, uint32_t* route_len,
                            uint32_t* cells, uint64_t cell_capacity, uint64_t* cell_count);
void apgx_negotiation_release(apgx_negotiation* negotiation);

/* ------------------------------------------------------------------------ */
/* Candidate-allocator price epoch                                          */
/* ------------------------------------------------------------------------ */
/* Replaces the price update of PlanNegotiatedRegeneration
 * (src/allocator/negotiated_regeneration_plan.cc:697-790) and NextPresentFactor
 * (:430-442) on cell resources: resource order is cell order, capacity is per
 * cell (NULL: 1), usage per cell.  Policy and limit defaults are
 * NegotiatedPriceUpdatePolicyV1 and NegotiatedRegenerationPlanLimits
 * (include/apgar/allocator/negotiated_regeneration_plan.h:28-35, 130-140).
 *   factor = initial_present_factor + epoch_index * present_factor_increment
 *   overused cell: present = factor * overuse, history += increment * overuse
 *   every entry (history != 0 or overused): total = present + history
 * Checks, first failure wins, in the reference's order: the factor's
 * arithmetic; hot resources (overused cells) <= maximum_hot_resources; per
 * overused cell in cell order the entry bound (prior entries + new entries
 * so far < maximum_price_entries) and the multiply / history / total
 * overflows; per entry in cell order the value bound on present, history and
 * total, then the running aggregate of totals (overflow, then bound).  A
 * failing epoch reports its outcome and cell and leaves the planes untouched;
 * the call still returns APGX_OK.  Planes are host memory, u32 per cell
 * (maximum_price_value must fit in u32); the update runs on the device. */
#define APGX_PRICE_OK 0u
#define APGX_PRICE_PRESENT_FACTOR_OVERFLOW 1u
#define APGX_PRICE_HOT_RESOURCE_BOUND 2u
#define APGX_PRICE_ENTRY_BOUND 3u
#define APGX_PRICE_MULTIPLY_OVERFLOW 4u
#define APGX_PRICE_HISTORY_OVERFLOW 5u
#define APGX_PRICE_TOTAL_OVERFLOW 6u
#define APGX_PRICE_VALUE_BOUND 7u
#define APGX_PRICE_AGGREGATE_OVERFLOW 8u
#define APGX_PRICE_AGGREGATE_BOUND 9u

typedef struct apgx_price_policy {
  uint64_t initial_present_factor;      /* 100 */
  uint64_t present_factor_increment;    /* 100 */
  uint64_t historical_price_increment;  /* 10 */
} apgx_price_policy;

typedef struct apgx_price_limits {
  uint64_t maximum_price_entries;       /* 1,000,000 */
  uint64_t maximum_hot_resources;       /* 1,000,000 */
  uint64_t maximum_price_value;         /* 10,000,000 */
  uint64_t maximum_aggregate_price;     /* 1,000,000,000,000 */
} apgx_price_limits;

typedef struct apgx_price_epoch {
  uint32_t outcome;                     /* APGX_PRICE_* */
  uint32_t reserved;
  uint64_t outcome_cell;                /* failing cell, or UINT64_MAX */
  uint64_t epoch_index;
  uint64_t present_factor;
  uint64_t hot_resources;               /* overused cells */
  uint64_t price_entries;               /* cells with a price after the update */
  uint64_t aggregate_price;             /* sum of totals (when the epoch passes) */
} apgx_price_epoch;

void apgx_price_defaults(apgx_price_policy* policy, apgx_price_limits* limits);

/* Price snapshots (NegotiatedPriceSnapshot, include/apgar/allocator/
 * negotiated_regeneration_plan.h:49-60) on cell resources: the nonzero cells of
 * the present and history planes in cell order, with the board associations,
 * the price policy identity, the snapshot chain, the epoch and its present
 * factor.  Capture compacts the planes on the device (ballot and prefix-scan
 * compaction); the identity is SnapshotIdentity's field order
 * (src/allocator/negotiated_regeneration_plan.cc:108-127) with a cell resource
 * (layer, x, y) in place of an edge resource and the lattice shape among the
 * associations; validation restates ValidatePriorSnapshot (:211-396) check for
 * check, with the reference's invariant identifiers; restore writes a
 * snapshot back into dense planes so the next epoch resumes from it. */
typedef struct apgx_price_associations {   /* ResourceLatticeAssociations + lattice shape */
  uint64_t board_content_hash;
  uint64_t compiler_profile_fingerprint;
  uint32_t geometry_compiler_version;
  uint32_t width;
  uint32_t height;
  uint32_t layers;
} apgx_price_associations;

typedef struct apgx_price_entry {          /* NegotiatedResourcePrice on one cell */
  uint64_t cell;
  uint64_t present_price;
  uint64_t historical_price;
  uint64_t total_price;
} apgx_price_entry;

typedef struct apgx_price_snapshot {
  apgx_price_associations associations;
  uint32_t capacity_units;
  uint32_t reserved;
  uint64_t policy_identity;
  uint64_t prior_snapshot_identity;        /* 0 exactly at epoch 0 */
  uint64_t epoch_index;
  uint64_t present_factor;
  uint64_t price_count;
  uint64_t snapshot_identity;
} apgx_price_snapshot;

typedef struct apgx_price_check {
  uint32_t valid;
  uint32_t reserved;
  uint64_t entry_index;                    /* failing entry, or UINT64_MAX */
  char invariant[64];                      /* e.g. "allocator.negotiated_plan.prior_price_order.v1" */
} apgx_price_check;

/* PricePolicyIdentity (src/allocator/negotiated_regeneration_plan.cc:98-106). */
uint64_t apgx_price_policy_identity(const apgx_price_policy* policy);
uint64_t apgx_price_snapshot_identity(const apgx_price_snapshot* snapshot, const apgx_price_entry* entries);
/* present / history: host u32 planes of width*height*layers cells.  entries
 * NULL: only out->price_count is set (size query). */
int apgx_price_snapshot_capture(const apgx_price_associations* associations, uint32_t capacity_units,
                                const apgx_price_policy* policy, uint64_t epoch_index,
                                uint64_t prior_snapshot_identity, const uint32_t* present, const uint32_t* history,
                                apgx_price_entry* entries, uint64_t entry_capacity, apgx_price_snapshot* out);
/* maximum_overuse_per_epoch: the most overuse one epoch can put on a cell
 * (the reference derives it from its candidate-selection bounds; for cells:
 * nets routed in the epoch minus the cell capacity). */
int apgx_price_snapshot_validate(const apgx_price_snapshot* prior, const apgx_price_entry* entries,
                                 const apgx_price_associations* expected, uint32_t capacity_units,
                                 const apgx_price_policy* policy, const apgx_price_limits* limits,
                                 uint64_t maximum_epoch_index, uint64_t maximum_overuse_per_epoch,
                                 apgx_price_check* out);
int apgx_price_snapshot_restore(const apgx_price_snapshot* snapshot, const apgx_price_entry* entries,
                                uint32_t* present, uint32_t* history);
/* history: in/out; present and total (may be NULL): out. */
int apgx_price_update(uint64_t epoch_index, const apgx_price_policy* policy,
                      const apgx_price_limits* limits, uint64_t cells, const uint32_t* usage,
                      const uint8_t* capacity, uint32_t* history, uint32_t* present, uint32_t* total,
                      apgx_price_epoch* out);

/* No counterpart in the reference: the price epoch of apgx_price_update for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same checks,
 * outcomes and kernels, but usage, capacity (may be NULL), history (in/out),
 * present and total (may be NULL) are DEVICE planes of `cells` entries on
 * the current device, and the call allocates nothing: its block partials and
 * outcome record live in a fixed 4 MiB scratch array of the library, one per
 * GPU.  At most 67,108,864 cells (APGX_E_RESOURCE_EXHAUSTED beyond).  Calls
 * that overlap on one GPU share that array and race; the call returns when
 * the epoch has finished. */
#define APGX_PRICE_NO_ALLOC_MAX_CELLS 67108864ull
int apgx_price_update_no_alloc(uint64_t epoch_index, const apgx_price_policy* policy,
                               const apgx_price_limits* limits, uint64_t cells, const uint32_t* usage,
                               const uint8_t* capacity, uint32_t* history, uint32_t* present, uint32_t* total,
                               apgx_price_epoch* out);

/* ------------------------------------------------------------------------ */
/* Learned history-cost prior                                               */
/* ------------------------------------------------------------------------ */
/* No reference counterpart: the reference has no learned component.  A small
 * convolutional encoder-decoder maps per-layer maps of a board, one value per
 * lattice tile (pin density and blocked fraction), to per-layer history-cost
 * maps; the decoded integer plane only seeds apgx_negotiate's history
 * (initial_history), so exact searches and the overuse check still judge
 * every route and a zero plane is the fallback.
 *
 * Network: 4 levels (grid, /2, /4, /8) of two 3x3 convs (32 channels, ReLU),
 * 2x2 average pooling down, nearest 2x upsampling plus skip on the way up,
 * and a 3x3 head; 15 convs.  Convs run on bf16 MFMA with fp32 accumulation
 * (bf16 weights and activations at the matrix-core inputs, fp32 elsewhere).
 * Parameters, flat fp32, per conv i = 0..14 in order: weights
 * w[cout][3*3][cin] then bias[cout]; cin is 16 for conv 0 (inputs 2L..15 are
 * zero planes) and 32 otherwise, cout is 16 for conv 14 (outputs L..15 are
 * unused) and 32 otherwise.
 *
 * Planar layouts: inputs [sample][2L][grid_h][grid_w] with plane 2l the pin
 * density of layer l (pins in the tile / 64) and 2l+1 the blocked fraction
 * (lattice cells outside the grid count as blocked); targets and outputs
 * [sample][L][grid_h][grid_w] = log1p(mean tile history / unit). */
typedef struct apgx_prior apgx_prior;

typedef struct apgx_prior_config {
  uint32_t lattice_layers;  /* L, 1..8 */
  uint32_t grid_h;          /* tiles per column, multiple of 8 */
  uint32_t grid_w;          /* tiles per row, multiple of 8 */
  uint32_t reserved;
  uint64_t seed;            /* He-normal weight initialisation */
} apgx_prior_config;

typedef struct apgx_prior_train_options {
  uint32_t epochs;          /* default 200 */
  uint32_t batch;           /* default 32 */
  float learning_rate;      /* Adam, default 2e-3 */
  float beta1;              /* 0.9 */
  float beta2;              /* 0.999 */
  float epsilon;            /* 1e-8 */
  uint32_t augment;         /* 1: random flip / transpose per sample and step (square grids) */
  uint32_t use_graph;       /* 1: each step replays one captured hipGraph */
  uint64_t seed;            /* shuffling and augmentation */
} apgx_prior_train_options;

typedef struct apgx_prior_train_report {
  uint64_t steps;
  float first_epoch_loss;   /* mean step loss of the first and last epoch */
  float last_epoch_loss;
  double device_ms;         /* hipEvent envelope of all steps */
  double wall_ms;
  uint32_t graph;           /* 1 if steps replayed a hipGraph */
  uint32_t reserved;
} apgx_prior_train_report;

typedef struct apgx_prior_decode {
  uint32_t tile;            /* lattice cells per grid cell side (64) */
  uint32_t max_value;       /* clamp of the decoded history */
  float unit;               /* history unit of the targets (the historical increment) */
  float gain;               /* history = gain * unit * expm1(max(output, 0)), rounded */
} apgx_prior_decode;

int apgx_prior_create(const apgx_prior_config* config, apgx_prior** out);
void apgx_prior_release(apgx_prior* prior);
uint64_t apgx_prior_param_count(const apgx_prior* prior);
int apgx_prior_get_params(const apgx_prior* prior, float* params, uint64_t count);
int apgx_prior_set_params(apgx_prior* prior, const float* params, uint64_t count);
/* Outputs for n planar input samples. */
int apgx_prior_forward(apgx_prior* prior, const float* inputs, uint32_t n, float* outputs);
/* Mean squared error over the L used outputs of n samples, and its gradient
 * with respect to every parameter (flat layout above). */
int apgx_prior_gradients(apgx_prior* prior, const float* inputs, const float* targets, uint32_t n,
                         float* loss, float* grads);
void apgx_prior_train_defaults(apgx_prior_train_options* options);
/* Adam on the whole data set (uploaded once), `epochs` shuffled passes. */
int apgx_prior_train(apgx_prior* prior, const float* inputs, const float* targets, uint32_t samples,
                     const apgx_prior_train_options* options, apgx_prior_train_report* report);
/* Input planes of one board (2L x ceil(H/tile) x ceil(W/tile)), from the
 * lattice's blocked cells and the given pin cells. */
int apgx_prior_features(apgx_lattice* lattice, const uint32_t* pins, uint64_t pin_count, uint32_t tile,
                        float* features);
/* Training targets from a history plane (L x H x W u32). */
int apgx_prior_pool_history(const uint32_t* history, uint32_t width, uint32_t height, uint32_t layers,
                            uint32_t tile, float unit, float* targets);
/* Features, network and decoding on the device: the integer history plane
 * (L x H x W u32) that seeds apgx_negotiate for this board. */
int apgx_prior_seed_history(apgx_prior* prior, apgx_lattice* lattice, const uint32_t* pins,
                            uint64_t pin_count, const apgx_prior_decode* decode, uint32_t* history);

/* ------------------------------------------------------------------------ */
/* V1: reference DeviceCompiledBoardV1 / DeviceCandidateBatch records        */
/* ------------------------------------------------------------------------ */
/* Byte-identical to include/apgar/gpu/planar_router.h:170-423 (sizes checked
 * with static asserts in the library).  State index = node * 9 + incoming
 * heading (0..7 = E, NE, N, NW, W, SW, S, SE; 8 = none, start only);
 * UINT64_MAX labels are unreachable. */
#define APGX_V1_SCHEMA_VERSION 1u
#define APGX_V1_INVALID_INDEX 0xFFFFFFFFu
#define APGX_V1_GENERATOR_BUCKETED_FRONTIER 0u
#define APGX_V1_GENERATOR_HEADING_AWARE_SWEEP 1u
#define APGX_V1_REACHED 0u
#define APGX_V1_DISCONNECTED 1u
#define APGX_V1_BUDGET_EXHAUSTED 2u
#define APGX_V1_CANCELLED 3u
#define APGX_V1_BANNED_ADJUSTMENT 0xFFFFFFFFFFFFFFFFull
/* CandidateBatchReadbackKind (include/apgar/gpu/planar_router.h): the sweep
 * returns compact paths, the bucketed frontier its full query-major workspace. */
#define APGX_V1_READBACK_COMPACT_PATHS 0u
#define APGX_V1_READBACK_FULL_WORKSPACE 1u

typedef struct apgx_v1_header {            /* DeviceCompiledHeaderV1, 112 bytes */
  uint32_t schema_version;
  uint32_t compiler_version;
  uint64_t source_board_content_hash;
  uint64_t compiler_profile_fingerprint;
  uint64_t rule_bucket_identity;
  uint64_t device_view_fingerprint;
  uint64_t represented_nodes;
  uint64_t represented_states;
  uint64_t estimated_persistent_device_bytes;
  int64_t lattice_origin_x;
  int64_t lattice_origin_y;
  int64_t lattice_step;
  uint32_t orthogonal_step;
  uint32_t diagonal_step;
  uint32_t bend;
  uint8_t heading_mask;                    /* 1 horizontal | 2 vertical | 4 diagonal45 */
  uint8_t reserved[7];
} apgx_v1_header;

typedef struct apgx_v1_layer_range {       /* DeviceLayerRangeV1, 48 bytes */
  uint32_t layer;
  uint32_t node_offset;
  uint32_t node_count;
  uint32_t pad;
  int64_t minimum_lattice_x;
  int64_t maximum_lattice_x;
  int64_t minimum_lattice_y;
  int64_t maximum_lattice_y;
} apgx_v1_layer_range;

typedef struct apgx_v1_node {              /* DeviceNodeV1, 56 bytes */
  int64_t lattice_x;
  int64_t lattice_y;
  uint32_t layer;
  uint8_t legal_edges;
  uint8_t reserved[3];
  uint32_t neighbors[8];
} apgx_v1_node;

typedef struct apgx_v1_run {               /* DeviceRunV1, 12 bytes */
  uint32_t node_offset;
  uint32_t node_count;
  uint8_t direction;
  uint8_t reserved[3];
} apgx_v1_run;

typedef struct apgx_v1_policy_edge {       /* DeviceCandidatePolicyEdgeV1, 16 bytes */
  uint32_t physical_edge_index;            /* canonical source node * 8 + direction (< 4) */
  uint32_t reserved;
  uint64_t adjustment;                     /* APGX_V1_BANNED_ADJUSTMENT bans the edge */
} apgx_v1_policy_edge;

typedef struct apgx_v1_query {             /* DeviceCandidateBatchQueryV1, 120 bytes */
  uint32_t schema_version;
  uint32_t input_ordinal;
  uint64_t batch_id;
  uint64_t query_id;
  uint64_t workspace_owner;
  uint64_t policy_identity;
  uint64_t routing_profile_fingerprint;
  uint32_t start_node;
  uint32_t goal_node;
  uint32_t policy_offset;
  uint32_t policy_count;
  uint64_t workspace_offset;
  uint64_t workspace_state_count;
  uint64_t frontier_state_capacity;
  uint64_t orthogonal_step_surcharge;
  uint64_t diagonal_step_surcharge;
  uint64_t bend_surcharge;
  uint32_t maximum_rounds;
  uint8_t generator;
  uint8_t reserved[3];
} apgx_v1_query;

typedef struct apgx_v1_result {            /* DeviceCandidateBatchResultV1, 120 bytes */
  uint32_t schema_version;
  uint32_t compiler_version;
  uint32_t input_ordinal;
  uint32_t start_node;
  uint32_t goal_node;
  uint32_t goal_state;
  uint32_t rounds;
  uint8_t generator;
  uint8_t completion;                      /* APGX_V1_REACHED .. APGX_V1_CANCELLED */
  uint8_t reserved[2];
  uint64_t batch_id;
  uint64_t query_id;
  uint64_t workspace_owner;
  uint64_t policy_identity;
  uint64_t routing_profile_fingerprint;
  uint64_t source_board_content_hash;
  uint64_t compiler_profile_fingerprint;
  uint64_t rule_bucket_identity;
  uint64_t device_view_fingerprint;
  uint64_t examined_work;
  uint64_t heading_turn_relaxations;
} apgx_v1_result;

typedef struct apgx_v1_compact_path {      /* DeviceCandidateCompactPathV1, 32 bytes */
  uint32_t schema_version;
  uint32_t state_count;                    /* reverse (goal..start) states; 0 unless reached */
  uint64_t query_id;
  uint64_t workspace_owner;
  uint64_t state_offset;
} apgx_v1_compact_path;

typedef struct apgx_v1_batch_request {     /* BackendCandidateBatchExecutionRequest scalars */
  uint32_t schema_version;
  uint32_t generator;
  uint64_t batch_id;
  uint32_t maximum_rounds;
  uint32_t reserved;
  uint64_t maximum_workspace_states_per_query;
  uint64_t maximum_frontier_states_per_query;
  uint64_t maximum_device_bytes;
  uint64_t maximum_host_bytes;
} apgx_v1_batch_request;

typedef struct apgx_v1_batch_telemetry {   /* CandidateBatchTelemetry */
  uint64_t persistent_device_bytes;
  uint64_t batch_device_bytes;
  uint64_t workspace_capacity_device_bytes;
  uint64_t peak_device_bytes;
  uint64_t batch_host_bytes;
  uint64_t device_to_host_readback_bytes;
  uint64_t kernel_launch_count;
  uint64_t blocking_status_readback_count;
  uint32_t dispatched_rounds;
  uint32_t finalization_launch_count;
  uint32_t chunk_rounds;
  uint32_t reserved;
  double kernel_milliseconds;
} apgx_v1_batch_telemetry;

/* Host-side inputs of BuildDeviceCompiledBoardV1: the compiled profile scalars
 * and every represented node with its legal-direction mask, in any order. */
typedef struct apgx_v1_compiled_profile {
  uint32_t compiler_version;
  uint32_t heading_mask;
  uint64_t source_board_content_hash;
  uint64_t compiler_profile_fingerprint;
  uint64_t rule_bucket_identity;
  int64_t lattice_origin_x;
  int64_t lattice_origin_y;
  int64_t lattice_step;
  uint32_t tile_width_nodes;
  uint32_t tile_height_nodes;
  uint32_t orthogonal_step;
  uint32_t diagonal_step;
  uint32_t bend;
  uint32_t reserved;
} apgx_v1_compiled_profile;

typedef struct apgx_v1_compiled_node {
  int64_t lattice_x;
  int64_t lattice_y;
  uint32_t layer;
  uint8_t legal_edges;
  uint8_t reserved[3];
} apgx_v1_compiled_node;

typedef struct apgx_v1_board apgx_v1_board;        /* flattened view on the host */
typedef struct apgx_v1_view apgx_v1_view;          /* uploaded immutable view */
typedef struct apgx_v1_pending apgx_v1_pending;    /* executed batch awaiting readback */

/* Replaces BuildDeviceCompiledBoardV1 (src/gpu/device_board.cc:136-325):
 * tile-order flattening, neighbours, layer ranges, maximal directed runs,
 * persistent bytes and the FNV-1a view fingerprint. */
int apgx_v1_build_board(const apgx_v1_compiled_profile* profile, const apgx_v1_compiled_node* nodes,
                        uint64_t node_count, apgx_v1_board** out);
int apgx_v1_board_sizes(const apgx_v1_board* board, uint64_t* layer_count, uint64_t* node_count,
                        uint64_t* run_count, uint64_t* run_node_count);
int apgx_v1_board_copy(const apgx_v1_board* board, apgx_v1_header* header,
                       apgx_v1_layer_range* layers, apgx_v1_node* nodes, apgx_v1_run* runs,
                       uint32_t* run_nodes);
void apgx_v1_board_release(apgx_v1_board* board);

/* Replaces ComputeDeviceCompiledBoardFingerprintV1 (src/gpu/device_board.cc:82-134). */
uint64_t apgx_v1_view_fingerprint(const apgx_v1_header* header, const apgx_v1_layer_range* layers,
                                  uint64_t layer_count, const apgx_v1_node* nodes,
                                  uint64_t node_count, const apgx_v1_run* runs, uint64_t run_count,
                                  const uint32_t* run_nodes, uint64_t run_node_count);

/* Replaces IPlanarRouteBackend::UploadCompiledView (include/apgar/gpu/planar_router.h:551):
 * validates every structural invariant and the fingerprint, then uploads. */
int apgx_v1_view_upload(const apgx_v1_header* header, const apgx_v1_layer_range* layers,
                        uint64_t layer_count, const apgx_v1_node* nodes, uint64_t node_count,
                        const apgx_v1_run* runs, uint64_t run_count, const uint32_t* run_nodes,
                        uint64_t run_node_count, apgx_v1_view** out);
void apgx_v1_view_release(apgx_v1_view* view);

/* Replaces IPlanarRouteBackend::ExecuteCandidateBatch (include/apgar/gpu/planar_router.h:556-557)
 * for both generators: the heading-aware sweep (8 rounds per chunk) and the
 * bucketed frontier (CandidateFrontierChunk semantics, 32 rounds per chunk;
 * queries carry frontier_state_capacity = represented states).  `cancel`
 * (may be NULL) is sampled between round chunks; a nonzero value cancels the
 * unfinished queries.  Launches: init, one per chunk, one finalization when
 * queries are still running after the last chunk, then 3 compact-path kernels
 * (sweep) or 1 predecessor-selection kernel (frontier).  The sweep is read
 * back in compact-path form, the frontier as its full workspace
 * (apgx_v1_batch_readback_workspace), as the reference reads each back. */
int apgx_v1_batch_execute(apgx_v1_view* view, const apgx_v1_batch_request* request,
                          const apgx_v1_query* queries, uint32_t query_count,
                          const apgx_v1_policy_edge* policy_edges, uint64_t policy_edge_count,
                          const volatile uint32_t* cancel, apgx_v1_pending** out);
/* Replaces IPlanarRouteBackend::ReadbackCandidateBatch (include/apgar/gpu/planar_router.h:558-559),
 * compact-path form: results and compact headers in query order, reverse state
 * lists packed into `states`. */
int apgx_v1_batch_readback(apgx_v1_pending* pending, apgx_v1_result* results,
                           apgx_v1_compact_path* paths, uint32_t* states, uint64_t state_capacity,
                           uint64_t* states_used, apgx_v1_batch_telemetry* telemetry);
/* Per-query envelope of a readback: UntrustedCandidateBatchQueryResult
 * (include/apgar/gpu/planar_router.h:502-508) less the result header and the
 * compact path, i.e. the workspace slice and the query telemetry
 * (CandidateQueryTelemetry) the backend reports beside the header. */
typedef struct apgx_v1_query_envelope {
  uint64_t workspace_offset;
  uint64_t workspace_state_count;
  uint64_t examined_work;
  uint64_t heading_turn_relaxations;
  uint32_t rounds;
  uint32_t reserved;
} apgx_v1_query_envelope;

/* Full-workspace half of ReadbackCandidateBatch (gpu/cuda/planar_router.cu:
 * 2700-2760): the readback kind, the per-query envelopes (both kinds), and for
 * the frontier every state's label, lowest tight predecessor, state owner and
 * predecessor owner, query-major (`total` = queries x represented states; 0 for
 * the sweep).  Any array pointer may be NULL; non-NULL arrays need `capacity`
 * >= total. */
int apgx_v1_batch_readback_workspace(apgx_v1_pending* pending, uint32_t* readback_kind,
                                     apgx_v1_query_envelope* per_query, uint64_t* labels,
                                     uint32_t* predecessors, uint64_t* state_owners,
                                     uint64_t* predecessor_owners, uint64_t capacity, uint64_t* total);
/* Replay/test-only readback decorator (FaultInjectingCandidateBatchBackend,
 * src/gpu/fault_injecting_backend.cc:113-224): corrupts query 0 of an executed
 * batch's host copy the way the reference's decorator does, so the next
 * readback returns the corrupted records.  Production execution has no fault
 * controls; this exists for replays and trust-boundary tests. */
#define APGX_V1_FAULT_NONE 0u
#define APGX_V1_FAULT_WORKSPACE_BOUNDS 1u
#define APGX_V1_FAULT_WORKSPACE_OWNER 2u
#define APGX_V1_FAULT_QUERY_TELEMETRY 3u
#define APGX_V1_FAULT_MEMORY_ACCOUNTING 4u
#define APGX_V1_FAULT_BATCH_TELEMETRY 5u
#define APGX_V1_FAULT_QUERY_IDENTITY 6u
#define APGX_V1_FAULT_FALSE_DISCONNECTED 7u
#define APGX_V1_FAULT_COMPACT_PATH_BOUNDS 8u
#define APGX_V1_FAULT_COMPACT_PATH_ENDPOINT 9u
#define APGX_V1_FAULT_COMPACT_PATH_CYCLE 10u
#define APGX_V1_FAULT_COMPACT_PATH_HEADING 11u
#define APGX_V1_FAULT_COMPACT_PATH_EDGE 12u
int apgx_v1_pending_inject_fault(apgx_v1_pending* pending, uint32_t fault);
void apgx_v1_pending_release(apgx_v1_pending* pending);

/* Host trust boundary: re-checks an untrusted readback before any result is
 * used, as RouteCandidateBatchWithPreparedPreflight does (src/gpu/
 * planar_router.cc:2195-2380).  Order: generator; batch query identity (no
 * result names a foreign query, :2195-2218); batch telemetry (readback kind
 * and flat bounds, compact contiguity, byte accounting, peak / chunk / launch /
 * readback counts: ValidateCandidateBatchTelemetry, :1841-1955); then per query
 * the envelope (header echo, workspace bounds, ownership stamps or compact
 * header, round / work telemetry: ValidateCandidateBatchEnvelope, :1732-1839;
 * results out of query order are matched by query id, and a query whose
 * result is missing or duplicated fails gpu.batch.query.identity.v1) and the
 * result itself.  Compact paths (sweep): endpoints, cycles, headings,
 * edges, banned resources and cost (:969-1131).  Full workspace (frontier):
 * every predecessor exact for reached / disconnected and a lower bound
 * otherwise (ValidateAllPredecessors, :491-612), the goal state as the least
 * finite goal label and the predecessor walk to the start (:829-930), finite
 * goal labels under "disconnected" (:614-627).  Every "disconnected" is
 * confirmed by an exact CPU search (:629-667).
 * `valid` = 0 names the first failing query (APGX_V1_INVALID_INDEX: the whole
 * batch), the failure class (APGX_V1_CHECK_*) and the reference's invariant
 * identifier where it defines one (e.g. "gpu.batch.compact_path.cycle.v1";
 * empty otherwise; `detail` then carries the reference's message).
 * path_costs (may be NULL) receives each reached query's validated cost
 * (UINT64_MAX otherwise).  The call returns APGX_OK unless an argument is null. */
#define APGX_V1_CHECK_OK 0u
#define APGX_V1_CHECK_INTERNAL_INVARIANT 1u  /* PlanarGpuFailureCode::kInternalInvariant */
#define APGX_V1_CHECK_VALIDATION_FAILED 2u   /* PlanarGpuFailureCode::kValidationFailed */
typedef struct apgx_v1_validation {
  uint32_t valid;
  uint32_t query_index;                    /* APGX_V1_INVALID_INDEX: batch-level */
  uint32_t failure;                        /* APGX_V1_CHECK_* */
  uint32_t reserved;
  char invariant[64];
  char detail[128];
} apgx_v1_validation;

/* The workspace half of a readback, as apgx_v1_batch_readback_workspace
 * returns it; counts are the lengths of the flat arrays (0 and NULL for the
 * compact form). */
typedef struct apgx_v1_workspace_view {
  uint32_t readback_kind;                  /* APGX_V1_READBACK_* */
  uint32_t reserved;
  const apgx_v1_query_envelope* queries;  /* query_count envelopes */
  const uint64_t* labels;
  uint64_t label_count;
  const uint32_t* predecessors;
  uint64_t predecessor_count;
  const uint64_t* state_owners;
  uint64_t state_owner_count;
  const uint64_t* predecessor_owners;
  uint64_t predecessor_owner_count;
} apgx_v1_workspace_view;

int apgx_v1_validate_batch(const apgx_v1_header* header, const apgx_v1_node* nodes, uint64_t node_count,
                           const apgx_v1_batch_request* request, const apgx_v1_query* queries,
                           uint32_t query_count, const apgx_v1_policy_edge* policy_edges,
                           uint64_t policy_edge_count, const apgx_v1_result* results,
                           const apgx_v1_compact_path* paths, const uint32_t* states,
                           uint64_t state_count, const apgx_v1_workspace_view* workspace,
                           const apgx_v1_batch_telemetry* telemetry, uint64_t* path_costs,
                           apgx_v1_validation* out);

/* The same checks with a verdict per query, as the reference's host router
 * returns one result per admitted query (src/gpu/planar_router.cc:2185-2436):
 * a corrupted query fails while its peers keep theirs.  `batch` gets the
 * batch-level verdict: the result set must name the admitted queries (a
 * foreign query id fails the batch), then the readback kind, flat bounds,
 * compact-path storage and memory / launch telemetry; on a batch failure
 * every verdicts[i] is that failure.  Otherwise batch->valid = 1 and
 * verdicts[i] (query_index i) is query i's own: valid, or its first failing
 * check as for apgx_v1_validate_batch.  Results are matched to queries by
 * query id when they are not in query order; a query whose result is missing
 * or came back twice fails with gpu.batch.query.identity.v1.  Queries are
 * checked by min(8, ceil(query_count / 64)) host threads over contiguous
 * partitions (the reference's CandidateCompactValidationWorkerCountV1,
 * include/apgar/gpu/planar_router.h:439-450); the output does not depend on
 * their scheduling.  path_costs as for apgx_v1_validate_batch, for every
 * valid reached query.  Returns APGX_OK unless an argument is null. */
int apgx_v1_validate_batch_queries(const apgx_v1_header* header, const apgx_v1_node* nodes, uint64_t node_count,
                                   const apgx_v1_batch_request* request, const apgx_v1_query* queries,
                                   uint32_t query_count, const apgx_v1_policy_edge* policy_edges,
                                   uint64_t policy_edge_count, const apgx_v1_result* results,
                                   const apgx_v1_compact_path* paths, const uint32_t* states,
                                   uint64_t state_count, const apgx_v1_workspace_view* workspace,
                                   const apgx_v1_batch_telemetry* telemetry, uint64_t* path_costs,
                                   apgx_v1_validation* verdicts, apgx_v1_validation* batch);

/* This backend's execution shape and deterministic byte accounting, for a
 * host that validates telemetry (the reference's DescribePlanarGenerator,
 * include/apgar/gpu/planar_router.h:71-99, and ExpectedCandidateBatchBytes /
 * EstimateCandidateBatchHostBytesV1, src/gpu/planar_router.cc:1701-1730,
 * 738-794, model the CUDA backend's launches and buffers; behind this
 * library a host takes these figures instead).  Launches per batch =
 * fixed_launches + ceil(dispatched_rounds / chunk_rounds) + finalization;
 * the sweep fuses the reference's per-round kernels into one launch per
 * chunk, so launches_once_per_chunk is 1 for both generators. */
typedef struct apgx_v1_generator_shape_info {
  uint32_t chunk_rounds;
  uint32_t fixed_launches;
  uint32_t fixed_examined_work_per_round;
  uint32_t examined_work_per_node_per_round;
  uint32_t heading_turn_relaxations_per_examined_work;
  uint32_t readback_kind;                  /* APGX_V1_READBACK_* */
  uint32_t launches_once_per_chunk;
  uint32_t reserved;
} apgx_v1_generator_shape_info;

int apgx_v1_generator_shape(uint32_t generator, apgx_v1_generator_shape_info* out);
/* batch_device_bytes, batch_host_bytes and device_to_host_readback_bytes of a
 * batch; compact_path_states: packed path states read back (sweep; the
 * pre-execution host budget uses query_count * represented states). */
int apgx_v1_expected_batch_bytes(uint32_t generator, uint64_t query_count, uint64_t policy_edge_count,
                                 uint64_t node_count, uint64_t compact_path_states, uint64_t* device_bytes,
                                 uint64_t* host_bytes, uint64_t* readback_bytes);

/* ------------------------------------------------------------------------ */
/* V1 single route                                                          */
/* ------------------------------------------------------------------------ */
/* One start / goal node pair on an uploaded view under the profile's step and
 * bend costs (no candidate policy): replaces IPlanarRouteBackend::ExecuteRoute
 * and ReadbackRoute (include/apgar/gpu/planar_router.h:549-551; the CUDA
 * backend's gpu/cuda/planar_router.cu:2049-2181 and 2794-2853), which
 * RouteWithPreparedPlanarGpuBackend (src/gpu/planar_router.cc:1206-1372)
 * calls for default-policy routes.  Two generators, each with the reference's
 * round semantics, so rounds and examined work equal the reference's:
 *  - bucketed frontier: delta-stepping over heading states (FrontierRelax /
 *    RunFrontier, :363-412, :1552-1640).  Each round relaxes every active
 *    state whose label lies in the current bucket (width min(orthogonal,
 *    diagonal step)) and defers the others; a round with nothing left to do
 *    converges.  It explores the whole reachable state space, it does not stop
 *    at the goal.
 *  - heading-aware sweep: departures then segmented run scans each round
 *    (ComputeDepartureLabels / SweepRuns / RunSweep, :414-484, :1642-1722),
 *    converged when a round lowers no label.
 * Then every state's lowest tight predecessor (SelectStablePredecessors,
 * :486-524).  The whole round loop runs in one launch (one workgroup), with
 * cancellation read once per round from a host-mapped word that the executing
 * thread mirrors from `cancel`; the call returns when the route has finished.
 * Errors as the reference's: foreign or null handles, endpoints outside the
 * view, equal endpoints, a zero round budget or an unknown generator are
 * internal invariants; a current device other than the upload device is a
 * backend failure; a device budget below the deterministic route bytes is
 * resource exhaustion.  The pending route holds the view's execution lease
 * until it is released (a second route on the same view waits). */
#define APGX_V1_ROUTE_RESULT_SCHEMA_VERSION 1u

typedef struct apgx_v1_route_request {     /* BackendExecutionRequest (planar_router.h:326-334) */
  uint32_t generator;                      /* APGX_V1_GENERATOR_* */
  uint32_t start_node;
  uint32_t goal_node;
  uint32_t maximum_rounds;                 /* > 0 */
  uint64_t routing_profile_fingerprint;
  uint64_t maximum_device_bytes;
} apgx_v1_route_request;

typedef struct apgx_v1_route_header {      /* DeviceRouteResultHeaderV1, 64 bytes (planar_router.h:225-241) */
  uint32_t schema_version;
  uint32_t compiler_version;
  uint32_t start_node;
  uint32_t goal_node;
  uint64_t source_board_content_hash;
  uint64_t compiler_profile_fingerprint;
  uint64_t routing_profile_fingerprint;
  uint64_t rule_bucket_identity;
  uint64_t device_view_fingerprint;
  uint8_t generator;
  uint8_t reserved[7];
} apgx_v1_route_header;

typedef struct apgx_v1_kernel_telemetry {  /* KernelTelemetry (planar_router.h:114-124) */
  uint64_t persistent_device_bytes;
  uint64_t batch_device_bytes;
  uint64_t peak_device_bytes;
  uint64_t examined_work;
  uint64_t heading_turn_relaxations;
  uint32_t rounds;
  uint32_t reserved;
  double kernel_milliseconds;
} apgx_v1_kernel_telemetry;

/* UntrustedKernelResult (planar_router.h:301-318) less its two arrays. */
typedef struct apgx_v1_route_result {
  apgx_v1_route_header header;             /* as the device wrote it */
  uint32_t completion;                     /* APGX_V1_REACHED .. APGX_V1_CANCELLED */
  uint32_t goal_state;                     /* reached: the lowest-index minimal goal state; else invalid */
  uint64_t state_count;                    /* entries of the label and predecessor arrays */
  apgx_v1_kernel_telemetry telemetry;
} apgx_v1_route_result;

typedef struct apgx_v1_route_pending apgx_v1_route_pending;

int apgx_v1_route_execute(apgx_v1_view* view, const apgx_v1_route_request* request,
                          const volatile uint32_t* cancel, apgx_v1_route_pending** out);
/* Result, then every state's label (UINT64_MAX: unreachable) and predecessor
 * (APGX_V1_INVALID_INDEX: none) when the arrays are non-NULL; they need
 * `capacity` >= the view's represented states. */
int apgx_v1_route_readback(apgx_v1_route_pending* pending, apgx_v1_route_result* result, uint64_t* labels,
                           uint32_t* predecessors, uint64_t capacity);
/* Replay/test-only readback decorator (FaultInjectingPlanarRouteBackend,
 * src/gpu/fault_injecting_backend.cc:47-79): with the goal-predecessor fault a
 * reached readback names its goal state as its own predecessor. */
#define APGX_V1_ROUTE_FAULT_NONE 0u
#define APGX_V1_ROUTE_FAULT_GOAL_PREDECESSOR_SELF_CYCLE 1u
int apgx_v1_route_pending_inject_fault(apgx_v1_route_pending* pending, uint32_t fault);
void apgx_v1_route_pending_release(apgx_v1_route_pending* pending);

/* Deterministic device bytes of one route (ValidateKernelTelemetry's formula,
 * src/gpu/planar_router.cc:411-454): the 64-byte result header, a u64 label
 * and a u32 predecessor per state, then two u32 round marks per state and two
 * u64 words (frontier) or eight u64 departures per node and two u64 words
 * (sweep). */
int apgx_v1_expected_route_bytes(uint32_t generator, uint64_t node_count, uint64_t* device_bytes);

/* Host trust boundary of a single route, in RouteWithPreparedPlanarGpuBackend's
 * order (src/gpu/planar_router.cc:1306-1364): kernel telemetry (:411-454), raw
 * associations (:382-409), completion telemetry (:456-484), every predecessor
 * (ValidateAllPredecessors, exact for reached and disconnected, a lower bound
 * otherwise), then for a reached route the goal state (least finite goal
 * label, lowest index on a tie) and the predecessor walk from it to the start
 * (:829-930), for a disconnected one finite goal labels (:613-627) and an exact
 * CPU search confirming the verdict (:629-667).  `out` as for
 * apgx_v1_validate_batch (query_index 0); path_cost (may be NULL) gets the
 * validated cost of a reached route (UINT64_MAX otherwise).  Returns APGX_OK
 * unless an argument is null. */
int apgx_v1_validate_route(const apgx_v1_header* header, const apgx_v1_node* nodes, uint64_t node_count,
                           const apgx_v1_route_request* request, const apgx_v1_route_result* result,
                           const uint64_t* labels, const uint32_t* predecessors, uint64_t* path_cost,
                           apgx_v1_validation* out);

/* ------------------------------------------------------------------------ */
/* Candidate-batch invariant replays (gpu_batch_replay v1)                   */
/* ------------------------------------------------------------------------ */

/* A parsed replay artifact (schemas/gpu_batch_replay/v1.md; field meaning and
 * order of tools/gpu_batch_replay.cc:129-226).  The policy is the v1 replay
 * policy: base_scalar_cost with literal empty ban and penalty sets. */
#define APGX_REPLAY_FAULT_UNAUTHENTICATED_PRODUCER 13u  /* beside APGX_V1_FAULT_* */
#define APGX_REPLAY_EXPECT_INTERNAL_INVARIANT 1u
#define APGX_REPLAY_EXPECT_UNSUPPORTED 2u
typedef struct apgx_batch_replay {
  char fixture[256];                       /* repository-relative fixture path */
  uint64_t fixture_checksum;               /* FNV-1a of the fixture bytes */
  uint32_t generator;                      /* APGX_V1_GENERATOR_* */
  uint32_t fault;                          /* APGX_V1_FAULT_* or APGX_REPLAY_FAULT_* */
  uint32_t layer;
  uint32_t maximum_rounds;
  uint64_t maximum_workspace_states_per_query;
  uint64_t maximum_frontier_states_per_query;
  uint64_t maximum_device_bytes;
  uint64_t maximum_host_bytes;
  uint64_t batch_id;
  uint64_t query_id;
  uint32_t input_ordinal;
  uint32_t policy_schema_version;
  uint64_t deterministic_seed;
  uint32_t candidate_ordinal;
  uint32_t expected_compiler_version;
  uint64_t orthogonal_step_surcharge;
  uint64_t diagonal_step_surcharge;
  uint64_t bend_surcharge;
  uint64_t expected_policy_identity;
  uint64_t expected_board_content_hash;
  uint64_t expected_compiler_profile_fingerprint;
  uint64_t expected_rule_bucket_identity;
  uint64_t expected_routing_profile_fingerprint;
  uint64_t expected_device_view_fingerprint;
  uint32_t expected_failure;               /* APGX_REPLAY_EXPECT_* */
  uint32_t reserved;
  char expected_invariant[64];
} apgx_batch_replay;

/* Parses one canonical artifact: UTF-8, no NUL / CR, LF line ends, the FNV-1a
 * trailer, every field once in canonical order, canonical unsigned decimals
 * within their widths, and the generator / fault / expected-outcome contract.
 * Returns APGX_OK, or APGX_E_UNSUPPORTED with the reference's message
 * ("replay payload checksum mismatch", "replay format, schema, generator,
 * policy, fault, or expected outcome is unsupported", ...) in `error`. */
int apgx_batch_replay_parse(const char* text, uint64_t length, apgx_batch_replay* out, char* error,
                            uint64_t error_capacity);

/* FingerprintCandidateGenerationPolicy (src/routing/candidate_policy.cc:139-160)
 * of the artifact's policy. */
uint64_t apgx_batch_replay_policy_identity(const apgx_batch_replay* artifact);

/* Replay fixture: the output of the reference's board compiler for one corpus
 * case -- compiled profile, request endpoints and represented nodes -- which
 * apgx_v1_build_board turns into the DeviceCompiledBoardV1.  Layout: this
 * header, then node_count apgx_v1_compiled_node records, little endian. */
#define APGX_REPLAY_FIXTURE_MAGIC "APGX-V1-FIXTURE"
typedef struct apgx_replay_fixture_header {
  char magic[16];                          /* APGX_REPLAY_FIXTURE_MAGIC, NUL padded */
  uint32_t schema_version;                 /* 1 */
  uint32_t layer;                          /* request layer */
  int64_t start_x, start_y;                /* request endpoints, lattice coordinates */
  int64_t goal_x, goal_y;
  uint64_t routing_profile_fingerprint;    /* FingerprintRoutingProfile of the routed net */
  apgx_v1_compiled_profile profile;
  uint64_t node_count;
} apgx_replay_fixture_header;

#define APGX_REPLAY_STAGE_PARSE 1u
#define APGX_REPLAY_STAGE_FIXTURE 2u
#define APGX_REPLAY_STAGE_ASSOCIATION 3u
#define APGX_REPLAY_STAGE_EXECUTE 4u
#define APGX_REPLAY_STAGE_REPRODUCE 5u
#define APGX_REPLAY_DRY_RUN 1u                 /* stop after the association gate (no GPU work) */
typedef struct apgx_replay_outcome {
  int32_t exit_code;     /* the tool's: 0 reproduced (or dry run passed), 1 not reproduced, 2 artifact,
                            fixture or association error */
  uint32_t stage;        /* APGX_REPLAY_STAGE_* reached */
  char invariant[64];    /* invariant the trust boundary reported */
  char message[256];
} apgx_replay_outcome;

/* Runs one artifact against its fixture bytes (Replay, tools/gpu_batch_replay.cc:
 * 228-391): parse, fixture checksum, device-view compilation, the association
 * gate (board hash, compiler profile and version, rule bucket, routing profile,
 * device view, policy identity) before any GPU work, then one query executed
 * with the forced generator and budgets, the readback corrupted by
 * apgx_v1_pending_inject_fault, validated by apgx_v1_validate_batch, and the
 * reported failure class and invariant compared with the expected ones.  The
 * unauthenticated-producer contract needs candidate construction, which this
 * library does not contain: it stops at the gate with exit code 2.  Returns
 * APGX_OK unless an argument is null; the verdict is in `out`. */
int apgx_batch_replay_run(const char* text, uint64_t text_length, const uint8_t* fixture, uint64_t fixture_length,
                          uint32_t flags, apgx_replay_outcome* out);

/* ------------------------------------------------------------------------ */
/* Single-route invariant replays (gpu_replay v1)                            */
/* ------------------------------------------------------------------------ */
/* GpuReplayArtifactV1 (tools/gpu_replay_parser.h:16-28): fixture, forced
 * generator (bucketed frontier), the goal-predecessor fault, layer, round
 * budget and the four expected associations; the only v1 contract is
 * internal_invariant / gpu.predecessor.self_reference.v1. */
typedef struct apgx_route_replay {
  char fixture[256];
  uint64_t fixture_checksum;
  uint32_t generator;                      /* APGX_V1_GENERATOR_BUCKETED_FRONTIER */
  uint32_t fault;                          /* APGX_V1_ROUTE_FAULT_GOAL_PREDECESSOR_SELF_CYCLE */
  uint32_t layer;
  uint32_t maximum_rounds;
  uint64_t expected_board_content_hash;
  uint64_t expected_profile_fingerprint;
  uint64_t expected_routing_profile_fingerprint;
  uint64_t expected_device_view_fingerprint;
  char expected_invariant[64];
} apgx_route_replay;

/* ParseGpuReplayArtifactV1 (tools/gpu_replay_parser.cc:11-56): the canonical
 * envelope of the batch artifacts with this format's 14 keys; APGX_E_UNSUPPORTED
 * with the reference's message in `error` otherwise. */
int apgx_route_replay_parse(const char* text, uint64_t length, apgx_route_replay* out, char* error,
                            uint64_t error_capacity);
/* RunGpuReplayArtifactV1 (tools/gpu_replay_runner.cc:40-124): parse, fixture
 * checksum and layer, device-view compilation, the association gate (board,
 * compiler profile, routing profile, device view; any mismatch is exit code 2
 * before GPU work), then the route on the GPU with the fault decorator,
 * validated by apgx_v1_validate_route; reproduced (exit 0) when it fails as an
 * internal invariant with the expected identifier, else exit 1.  The fixture
 * is the batch replays' (apgx_replay_fixture_header + nodes). */
int apgx_route_replay_run(const char* text, uint64_t text_length, const uint8_t* fixture,
                          uint64_t fixture_length, uint32_t flags, apgx_replay_outcome* out);

/* ------------------------------------------------------------------------ */
/* Synthetic lattice generators (deterministic xoshiro256** streams)         */
/* ------------------------------------------------------------------------ */

typedef struct apgx_lattice_gen_params {
  uint32_t width;
  uint32_t height;
  uint32_t layers;
  uint32_t cost_min;            /* uniform integer cost in [cost_min, cost_max] */
  uint32_t cost_max;
  uint32_t blocked_ppm;         /* i.i.d. blocked probability, parts per million */
  uint32_t keepout_ppm;         /* per-layer rectangular keepout coverage target */
  uint32_t keepout_min_side;
  uint32_t keepout_max_side;
  uint32_t reserved;
  uint64_t seed;
} apgx_lattice_gen_params;

/* Writes width*height*layers uint32 costs (0 = blocked). */
int apgx_generate_lattice(const apgx_lattice_gen_params* params, uint32_t* cost_out);

/* Random single-pair queries: pins_out gets 2*count cells (source, target),
 * both free and distinct, uniform over all free cells. */
int apgx_generate_pairs(const uint32_t* cost, uint32_t width, uint32_t height, uint32_t layers,
                        uint32_t count, uint64_t seed, uint32_t* pins_out);

/* Random multi-pin nets with log-uniform square windows of side
 * [min_side, max_side]; pin counts uniform in [min_pins, max_pins]; pins are
 * distinct free cells inside the window.  Writes count net records (pin
 * offsets into pins_out) and returns the pin total in *pin_total.
 * pins_out must hold count*max_pins cells. */
int apgx_generate_nets(const uint32_t* cost, uint32_t width, uint32_t height, uint32_t layers,
                       uint32_t count, uint32_t min_pins, uint32_t max_pins, uint32_t min_side,
                       uint32_t max_side, uint64_t seed, apgx_net_query* nets_out,
                       uint32_t* pins_out, uint64_t* pin_total);

/* Allocator boards: the lattice model above (costs U[cost_min, cost_max],
 * rectangular keepouts), `bga_fields` non-overlapping fields of bga_pads x
 * bga_pads pads at bga_pitch (pads on layer 0, forced free), and `nets` nets
 * of [min_pins, max_pins] distinct pins.  Pin 0 of each net is an unused pad
 * of a random field; every further pin is, with probability pad_pin_ppm, an
 * unused pad (same field or another), otherwise a free cell on a random layer
 * within scatter_radius (Chebyshev) of that field's centre.  No cell is a pin
 * of two nets.  The window of a net is its pin bounding box grown by
 * window_margin and clipped.  cost_out holds width*height*layers u32;
 * pins_out nets*max_pins cells. */
typedef struct apgx_board_gen_params {
  uint32_t width, height, layers;
  uint32_t cost_min, cost_max;
  uint32_t keepout_ppm, keepout_min_side, keepout_max_side;
  uint32_t nets, min_pins, max_pins;
  uint32_t bga_fields, bga_pads, bga_pitch;
  uint32_t pad_pin_ppm;
  uint32_t scatter_radius;
  uint32_t window_margin;
  uint32_t pin_clearance;       /* r > 0: no scattered pin within Chebyshev r of another pin on
                                   its layer or on the cell above a pad; those cells are forced free */
  uint64_t seed;
} apgx_board_gen_params;

int apgx_generate_board(const apgx_board_gen_params* params, uint32_t* cost_out,
                        apgx_net_query* nets_out, uint32_t* pins_out, uint64_t* pin_total);

#ifdef __cplusplus
}
#endif

#endif /* APGX_H_ */
