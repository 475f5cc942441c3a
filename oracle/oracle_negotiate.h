/*
 * oracle_negotiate.h — CPU oracle of the negotiated-congestion allocator.
 * TEST INFRASTRUCTURE ONLY (see negotiate.c for the restated algorithm and
 * its pinning).
 */
#ifndef APGX_ORACLE_NEGOTIATE_H_
#define APGX_ORACLE_NEGOTIATE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_NEGO_FEASIBLE 0u
#define ORC_NEGO_STALLED 1u
#define ORC_NEGO_PASS_LIMIT 2u
#define ORC_NEGO_BOUND 1

typedef struct orc_nego_policy {
  uint32_t maximum_passes;
  uint32_t batch_nets;                    /* nets per snapshot; 1: CS-RR-v1 order */
  uint64_t initial_present_cost;
  uint64_t present_cost_increment;
  uint64_t historical_cost_increment;
  uint64_t maximum_congestion_cost_value;
  uint64_t maximum_cell_cost;             /* base + congestion bound of the search */
} orc_nego_policy;

typedef struct orc_nego_net {
  uint32_t pin_offset, pin_count;
  uint32_t wx0, wy0, ww, wh;              /* ww == 0: whole lattice */
} orc_nego_net;

typedef struct orc_nego_result {
  int32_t status;                         /* 0 ok, ORC_NEGO_BOUND: a value bound was exceeded */
  uint32_t termination;                   /* ORC_NEGO_* */
  uint32_t completed_passes;
  uint32_t routed_nets;
  uint64_t total_overuse_units;           /* at the last completed pass */
} orc_nego_result;

/* Returns ORC_OK (status in res) or ORC_ERR (allocation / internal). */
int orc_negotiate(uint32_t W, uint32_t H, uint32_t L, const uint32_t* base, uint32_t via,
                  const uint8_t* capacity, const orc_nego_net* nets, uint32_t net_count,
                  const uint32_t* pins, const orc_nego_policy* pol, const uint32_t* initial_history,
                  uint32_t* history_out, uint32_t* occupancy_out, uint64_t* overuse_per_pass,
                  uint32_t* route_len, uint32_t* route_cells, uint64_t route_cap,
                  orc_nego_result* res);

#ifdef __cplusplus
}
#endif

#endif /* APGX_ORACLE_NEGOTIATE_H_ */
