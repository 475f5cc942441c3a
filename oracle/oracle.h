/*
 * oracle.h — CPU oracle for the apgx library.  TEST INFRASTRUCTURE ONLY.
 *
 * This is a plain-C restatement of the reference algorithms on the search
 * path.  Only the test suite, the smoke check and the CPU-baseline leg of the
 * benchmark may load it, and only as the checker or as the timed CPU baseline.
 * The product (apgar_amd, libapgx.so) never links, loads or calls it.
 *
 * Pinning (what each function is checked against) is listed per function; a
 * function without such a check would say "parity unpinned" here.
 */
#ifndef APGX_ORACLE_H_
#define APGX_ORACLE_H_

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_UNREACHABLE 0xFFFFFFFFu
#define ORC_OK 0
#define ORC_ERR -1

/* ---------------------------- dense lattice ------------------------------ */
/* Cells c = (l*H + y)*W + x; cost[c] == 0 blocked; in-layer move u->v costs
 * cost[v]; via (x,y,l)<->(x,y,l+1) costs via (0 = no vias).  Window
 * [wx0,wx0+ww) x [wy0,wy0+wh) on every layer (ww == 0 selects everything). */
typedef struct orc_lattice {
  uint32_t W, H, L;
  const uint32_t* cost;
  uint32_t via;
  uint32_t wx0, wy0, ww, wh;
  uint32_t via_pays_cell; /* 1: a via move into v weighs via + cost[v] (allocator searches) */
} orc_lattice;

/* Exact multi-source Dijkstra (binary heap keyed (dist, cell)); dist has W*H*L
 * entries, ORC_UNREACHABLE outside the reached set.  Pinned by brute-force
 * Bellman-Ford on tiny lattices, closed forms (uniform-cost Manhattan
 * distances, single corridors) and agreement with orc_dense_astar. */
int orc_dense_dijkstra(const orc_lattice* lat, const uint32_t* sources, uint32_t nsrc,
                       uint32_t* dist);

/* Canonical reverse path from `target` back to a dist == 0 cell: at each cell
 * the predecessor is the lowest-index in-window free neighbour u with
 * dist[u] + w(u->v) == dist[v] (dense restatement of
 * gpu/cuda/planar_router.cu:1084-1165).  Writes target..source inclusive and
 * returns the count, or -1 on a broken chain / capacity overflow.  Pinned by
 * exhaustive enumeration of all shortest paths on tiny lattices. */
int64_t orc_dense_canonical_path(const orc_lattice* lat, const uint32_t* dist, uint32_t target,
                                 uint32_t* out, uint64_t cap);

/* A* restatement of routing::RouteWithCpuAStar (src/routing/cpu_astar.cc:323-638)
 * on the dense lattice: queue key (f, h, g, y, x, layer, sequence) as
 * QueueGreater (:67-74), lazy deletion (:404-407), strict-improvement
 * relaxation (:488-491) in a fixed neighbour order, goal on first pop
 * (:408-411).  Heuristic: (|dx|+|dy|)*min_free_cost + |dl|*via (consistent).
 * Returns the distance (ORC_UNREACHABLE if disconnected); the A*-order path
 * (goal..start) if path != NULL.  Pinned by Dijkstra distances. */
uint32_t orc_dense_astar(const orc_lattice* lat, uint32_t source, uint32_t target,
                         uint32_t* path, uint64_t cap, uint64_t* path_len, uint64_t* expanded);

/* Multi-pin net: tree = {pin0}; repeatedly multi-source Dijkstra from the tree,
 * connect the nearest unconnected pin (ties: lowest pin index) along its
 * canonical path, add the path.  conn_pin/conn_dist/conn_len get npins-1
 * entries, cells gets the concatenated connection paths (pin..tree cell).
 * Returns 0 reached, 1 disconnected, -1 error.  Pinned by the two-pin case
 * (equals Dijkstra + canonical path), hand-worked trees and ties, and an
 * independent brute-force restatement on tiny lattices (Bellman-Ford from the
 * tree over the explicit edge list, predecessor walk by lowest index). */
int orc_dense_route_net(const orc_lattice* lat, const uint32_t* pins, uint32_t npins,
                        uint32_t* conn_pin, uint32_t* conn_dist, uint32_t* conn_len,
                        uint32_t* cells, uint64_t cell_cap, uint64_t* cells_used);

#ifdef __cplusplus
}
#endif

#endif /* APGX_ORACLE_H_ */
