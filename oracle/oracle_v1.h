/*
 * oracle_v1.h — CPU oracle for the reference's planar heading-state model.
 * TEST INFRASTRUCTURE ONLY (tests/, smoke check); never linked by the product.
 *
 * Pinning: orc_v1_astar reproduces the reference's published known answers
 * (tests/routing/cpu_astar_test.cc:264-311 cost 37 and its exact path;
 * :528-557 costs 100/126/126; :559-598 cost 137 in both directions) and the
 * corpus optimal costs and CPU expansion counts of
 * benchmarks/results/phase2_planar_bakeoff_a2b73bd.json (140/149/454/1148/50/370,
 * disconnected; expanded 20/14/16/62/10/10/9).  orc_v1_canonical agrees with
 * orc_v1_astar on every cost and its paths satisfy the lowest-predecessor
 * condition checked independently in tests.
 */
#ifndef APGX_ORACLE_V1_H_
#define APGX_ORACLE_V1_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_v1_node {
  int64_t x, y;
  uint32_t layer;
  uint8_t mask;     /* legal directions, bit d = direction d */
  uint8_t pad[3];
} orc_v1_node;

typedef struct orc_v1_profile {
  uint32_t orth, diag, bend;
  uint32_t heading_mask;  /* 1 horizontal, 2 vertical, 4 diagonal45 */
  uint32_t tile_w, tile_h;
} orc_v1_profile;

typedef struct orc_v1_policy {
  uint64_t orth_s, diag_s, bend_s;
  const int64_t* ban_keys;   /* n_bans canonical (x, y, dir<4) triples */
  uint32_t n_bans;
  uint32_t n_pens;
  const int64_t* pen_keys;   /* n_pens canonical triples; duplicates add */
  const uint64_t* pen_vals;
} orc_v1_policy;

/* RouteWithCpuAStar restatement (src/routing/cpu_astar.cc:323-638).  Returns
 * 0 reached, 1 disconnected, -1 error; path_xy gets goal..start lattice points. */
int orc_v1_astar(const orc_v1_node* nodes, uint64_t n, const orc_v1_profile* pr,
                 const orc_v1_policy* po, uint32_t layer, int64_t sx, int64_t sy, int64_t gx,
                 int64_t gy, uint64_t* cost_out, int64_t* path_xy, uint64_t path_cap,
                 uint64_t* path_len, uint64_t* expanded_out);

/* Flattened node order of BuildDeviceCompiledBoardV1: order[i] = input index
 * of flattened node i. */
int orc_v1_flatten_order(const orc_v1_node* nodes, uint64_t n, uint32_t tile_w, uint32_t tile_h,
                         uint32_t* order);

/* Exact heading-state labels + lowest-predecessor reverse state path. */
int orc_v1_canonical(const orc_v1_node* nodes, uint64_t n, const orc_v1_profile* pr,
                     const orc_v1_policy* po, uint32_t layer, int64_t sx, int64_t sy, int64_t gx,
                     int64_t gy, uint64_t* cost_out, uint32_t* states_out, uint64_t cap,
                     uint64_t* len_out);

/* Batch frontier generator restatement (CandidateFrontierChunk): rounds,
 * examined work, goal state and the lowest-predecessor reverse path on its
 * labels, all in flattened numbering.  Returns 0 reached, 1 disconnected,
 * 2 budget exhausted, -1 error. */
int orc_v1_frontier(const orc_v1_node* nodes, uint64_t n, const orc_v1_profile* pr,
                    const orc_v1_policy* po, uint32_t layer, int64_t sx, int64_t sy, int64_t gx,
                    int64_t gy, uint32_t max_rounds, uint64_t* cost_out, uint32_t* goal_state_out,
                    uint32_t* rounds_out, uint64_t* examined_out, uint32_t* states_out, uint64_t cap,
                    uint64_t* len_out);

/* The frontier's full workspace after it stops (reached, disconnected or out
 * of rounds): every state's label and lowest tight predecessor
 * (UINT32_MAX: none), flattened numbering, n * 9 entries each. */
int orc_v1_frontier_workspace(const orc_v1_node* nodes, uint64_t n, const orc_v1_profile* pr,
                              const orc_v1_policy* po, uint32_t layer, int64_t sx, int64_t sy, int64_t gx,
                              int64_t gy, uint32_t max_rounds, uint32_t* goal_state_out, uint32_t* rounds_out,
                              uint64_t* labels_out, uint32_t* preds_out);

/* Single-route generators (generator 0 bucketed frontier, 1 heading-aware
 * sweep) of the reference's ExecuteRoute on the profile's costs, then every
 * state's lowest tight predecessor, flattened numbering.  completion: 0
 * reached, 1 disconnected, 2 budget exhausted.  Pinned by the reference's
 * published single-route rounds and examined work for both generators on the
 * corpus (benchmarks/results/phase2_planar_bakeoff_a2b73bd.json), labels equal
 * to an independent Dijkstra over the heading states at convergence, and a
 * hand-worked corridor (rounds, examined work, turn relaxations, labels,
 * predecessors). */
int orc_v1_route(const orc_v1_node* nodes, uint64_t n, const orc_v1_profile* pr, uint32_t generator, uint32_t layer,
                 int64_t sx, int64_t sy, int64_t gx, int64_t gy, uint32_t max_rounds, uint32_t* completion_out,
                 uint32_t* rounds_out, uint64_t* examined_out, uint64_t* turns_out, uint32_t* goal_state_out,
                 uint64_t* labels_out, uint32_t* preds_out);

#ifdef __cplusplus
}
#endif

#endif
