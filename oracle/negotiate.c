/*
 * negotiate.c — CPU oracle of the negotiated-congestion allocator.  TEST
 * INFRASTRUCTURE ONLY: loaded by tests/ and the benchmark's CPU-baseline leg,
 * never by the product library.
 *
 * Restates Canonical Sequential Rip-Up-and-Reroute v1
 * (src/allocator/cpu_sequential_negotiated_routing.cc:993-1329, policy
 * include/apgar/allocator/cpu_sequential_negotiated_routing.h:38-56) on the
 * dense lattice with one resource per cell:
 *
 *   pass p: present factor pf = initial + p * present_increment (:306-326)
 *   per net, in net order, in batches of `batch_nets` nets:
 *     rip up the batch's incumbents (:1024-1028),
 *     snapshot congestion cost per free cell
 *       cong = history + pf * max(0, usage + 1 - capacity)  (:446-480),
 *       cong above maximum_congestion_cost_value is a bound error (:473-480),
 *     route every net of the batch on cell cost base + cong (multi-pin growth
 *     from pin 0, canonical paths), install the new route or restore the
 *     incumbent when the net is disconnected (:1191-1254);
 *   accounting: overuse = max(0, usage - capacity) per cell (resource_accounting.cc:109-237);
 *   termination in the order feasible (every net routed, zero overuse),
 *     stalled (every route equal to the previous pass's), pass limit (:1284-1300);
 *   otherwise history += historical_increment * overuse, bound-checked (:1306-1328).
 *
 * batch_nets == 1 is exactly the sequential order of CS-RR-v1; larger batches
 * route the nets of a batch against one snapshot (each sees the others'
 * previous routes ripped up).  A route is the net's tree cells: pin 0, then
 * each connection path from its pin up to (not including) the tree cell it
 * reaches.
 *
 * Pinned by: the reference's zero-capacity known answer (congestion cost 12 =
 * 5 + 7 * max(0, 0 + 1 - 0) with bound 12 accepted and 11 rejected,
 * tests/allocator/cpu_sequential_negotiated_routing_test.cc:787-831), hand
 * examples of two nets crossing a one-cell bottleneck, history = sum over
 * non-terminal passes of increment x overuse, and the final overuse recomputed
 * from the routes.
 */
#include <stdlib.h>
#include <string.h>

#include "oracle.h"
#include "oracle_negotiate.h"

static int route_equal(const uint32_t* a, uint32_t na, const uint32_t* b, uint32_t nb) {
  return na == nb && (na == 0 || memcmp(a, b, (size_t)na * 4) == 0);
}

int orc_negotiate(uint32_t W, uint32_t H, uint32_t L, const uint32_t* base, uint32_t via,
                  const uint8_t* capacity, const orc_nego_net* nets, uint32_t net_count,
                  const uint32_t* pins, const orc_nego_policy* pol, const uint32_t* initial_history,
                  uint32_t* history_out, uint32_t* occupancy_out, uint64_t* overuse_per_pass,
                  uint32_t* route_len, uint32_t* route_cells, uint64_t route_cap,
                  orc_nego_result* res) {
  const uint64_t cells = (uint64_t)W * H * L;
  const uint32_t batch = pol->batch_nets ? pol->batch_nets : 1;
  memset(res, 0, sizeof(*res));
  uint64_t* hist = (uint64_t*)calloc(cells, 8);
  uint64_t* occ = (uint64_t*)calloc(cells, 8);
  uint32_t* snap = (uint32_t*)malloc(cells * 4);
  /* incumbent and previous-pass routes: per net an owned array */
  uint32_t** inc = (uint32_t**)calloc(net_count, sizeof(uint32_t*));
  uint32_t* inc_n = (uint32_t*)calloc(net_count, 4);
  uint32_t** prev = (uint32_t**)calloc(net_count, sizeof(uint32_t*));
  uint32_t* prev_n = (uint32_t*)calloc(net_count, 4);
  uint32_t* has_prev = (uint32_t*)calloc(net_count, 4);
  uint32_t* cpin = (uint32_t*)malloc(32 * 4);
  uint32_t* cdist = (uint32_t*)malloc(32 * 4);
  uint32_t* clen = (uint32_t*)malloc(32 * 4);
  uint64_t path_cap = cells * 4 + 64;
  uint32_t* path = (uint32_t*)malloc(path_cap * 4);
  int rc = ORC_OK;
  if (!hist || !occ || !snap || !inc || !inc_n || !prev || !prev_n || !has_prev || !path) {
    rc = ORC_ERR;
    goto done;
  }
  if (initial_history)
    for (uint64_t c = 0; c < cells; ++c) hist[c] = initial_history[c];
  int prev_valid = 0;
  for (uint32_t pass = 0; pass < pol->maximum_passes; ++pass) {
    const uint64_t pf = pol->initial_present_cost + (uint64_t)pass * pol->present_cost_increment;
    for (uint32_t b0 = 0; b0 < net_count; b0 += batch) {
      const uint32_t b1 = b0 + batch < net_count ? b0 + batch : net_count;
      for (uint32_t n = b0; n < b1; ++n)
        for (uint32_t i = 0; i < inc_n[n]; ++i) occ[inc[n][i]] -= 1;
      for (uint64_t c = 0; c < cells; ++c) {
        if (base[c] == 0) {
          snap[c] = 0;
          continue;
        }
        const uint64_t cap = capacity ? capacity[c] : 1;
        const uint64_t predicted = occ[c] + 1;
        const uint64_t cong = hist[c] + pf * (predicted > cap ? predicted - cap : 0);
        if (cong > pol->maximum_congestion_cost_value || base[c] + cong > pol->maximum_cell_cost) {
          res->status = ORC_NEGO_BOUND;
          goto done;
        }
        snap[c] = (uint32_t)(base[c] + cong);
      }
      for (uint32_t n = b0; n < b1; ++n) {
        const orc_nego_net* q = &nets[n];
        orc_lattice lat = {W, H, L, snap, via, q->wx0, q->wy0, q->ww, q->wh, 1u};
        uint64_t used = 0;
        const int r = orc_dense_route_net(&lat, pins + q->pin_offset, q->pin_count, cpin, cdist, clen,
                                          path, path_cap, &used);
        if (r < 0) {
          rc = ORC_ERR;
          goto done;
        }
        if (r == 0) {
          /* tree cells: pin 0, then each path minus its last (tree) cell */
          uint32_t m = 1;
          for (uint32_t k = 0; k + 1 < q->pin_count; ++k) m += clen[k] - 1;
          uint32_t* t = (uint32_t*)malloc((size_t)m * 4);
          if (!t) {
            rc = ORC_ERR;
            goto done;
          }
          uint32_t w = 0;
          t[w++] = pins[q->pin_offset];
          uint64_t off = 0;
          for (uint32_t k = 0; k + 1 < q->pin_count; ++k) {
            for (uint32_t i = 0; i + 1 < clen[k]; ++i) t[w++] = path[off + i];
            off += clen[k];
          }
          free(inc[n]);
          inc[n] = t;
          inc_n[n] = m;
        }
        /* disconnected: the incumbent (if any) is restored below */
      }
      for (uint32_t n = b0; n < b1; ++n)
        for (uint32_t i = 0; i < inc_n[n]; ++i) occ[inc[n][i]] += 1;
    }
    /* accounting at the pass boundary */
    uint64_t overuse = 0;
    for (uint64_t c = 0; c < cells; ++c) {
      const uint64_t cap = capacity ? capacity[c] : 1;
      if (occ[c] > cap) overuse += occ[c] - cap;
    }
    uint32_t routed = 0;
    for (uint32_t n = 0; n < net_count; ++n) routed += inc[n] != NULL;
    if (overuse_per_pass) overuse_per_pass[pass] = overuse;
    res->completed_passes = pass + 1;
    res->total_overuse_units = overuse;
    res->routed_nets = routed;
    if (routed == net_count && overuse == 0) {
      res->termination = ORC_NEGO_FEASIBLE;
      break;
    }
    int same = prev_valid;
    for (uint32_t n = 0; n < net_count && same; ++n)
      same = has_prev[n] == (inc[n] != NULL) && route_equal(prev[n], prev_n[n], inc[n], inc_n[n]);
    if (same) {
      res->termination = ORC_NEGO_STALLED;
      break;
    }
    if (pass + 1 == pol->maximum_passes) {
      res->termination = ORC_NEGO_PASS_LIMIT;
      break;
    }
    for (uint32_t n = 0; n < net_count; ++n) {
      free(prev[n]);
      prev[n] = NULL;
      prev_n[n] = inc_n[n];
      has_prev[n] = inc[n] != NULL;
      if (inc[n]) {
        prev[n] = (uint32_t*)malloc((size_t)inc_n[n] * 4 + 4);
        memcpy(prev[n], inc[n], (size_t)inc_n[n] * 4);
      }
    }
    prev_valid = 1;
    for (uint64_t c = 0; c < cells; ++c) {
      const uint64_t cap = capacity ? capacity[c] : 1;
      if (occ[c] > cap) {
        hist[c] += pol->historical_cost_increment * (occ[c] - cap);
        if (hist[c] > pol->maximum_congestion_cost_value) {
          res->status = ORC_NEGO_BOUND;
          goto done;
        }
      }
    }
  }
  if (history_out)
    for (uint64_t c = 0; c < cells; ++c) history_out[c] = (uint32_t)hist[c];
  if (occupancy_out)
    for (uint64_t c = 0; c < cells; ++c) occupancy_out[c] = (uint32_t)occ[c];
  if (route_len) {
    uint64_t w = 0;
    for (uint32_t n = 0; n < net_count; ++n) {
      route_len[n] = inc_n[n];
      if (route_cells) {
        if (w + inc_n[n] > route_cap) {
          rc = ORC_ERR;
          goto done;
        }
        if (inc_n[n]) memcpy(route_cells + w, inc[n], (size_t)inc_n[n] * 4);
      }
      w += inc_n[n];
    }
  }
done:
  if (inc)
    for (uint32_t n = 0; n < net_count; ++n) free(inc[n]);
  if (prev)
    for (uint32_t n = 0; n < net_count; ++n) free(prev[n]);
  free(inc);
  free(inc_n);
  free(prev);
  free(prev_n);
  free(has_prev);
  free(hist);
  free(occ);
  free(snap);
  free(path);
  free(cpin);
  free(cdist);
  free(clen);
  return rc;
}
