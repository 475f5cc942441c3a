/*
 * v1.c — CPU oracle for the reference's planar heading-state model (V1
 * DeviceCompiledBoard semantics).  TEST INFRASTRUCTURE ONLY: loaded by tests/
 * and the smoke check, never by the product library.
 *
 * Model (include/apgar/geometry_compiler/compiled_board.h:23-78,
 * src/routing/candidate_policy.cc:419-438):
 *   directions 0..7 = E, NE, N, NW, W, SW, S, SE; deltas as DeltaFor;
 *   state = (node, incoming heading 0..7 or 8 = none);
 *   w(u->v, d | h) = (diag ? diag_step + diag_surcharge : orth_step + orth_surcharge)
 *                    + [h not in {8, d}] * (bend + bend_surcharge)
 *                    + penalty(canonical undirected edge);
 *   banned edges are absent; sums reaching UINT64_MAX are rejected.
 */
#include <stdlib.h>
#include <string.h>

#include "oracle_v1.h"

static const int DX[8] = {1, 1, 0, -1, -1, -1, 0, 1};
static const int DY[8] = {0, 1, 1, 1, 0, -1, -1, -1};
#define INF64 UINT64_MAX

/* --------------------------------------------------------- node lookup -- */
typedef struct {
  int64_t x, y;
  uint32_t layer;
  uint32_t idx;
} keyed_t;

static int key_cmp(const void* a, const void* b) {
  const keyed_t* p = (const keyed_t*)a;
  const keyed_t* q = (const keyed_t*)b;
  if (p->layer != q->layer) return p->layer < q->layer ? -1 : 1;
  if (p->x != q->x) return p->x < q->x ? -1 : 1;
  if (p->y != q->y) return p->y < q->y ? -1 : 1;
  return 0;
}

typedef struct {
  keyed_t* keys;
  uint64_t n;
} lookup_t;

static int lookup_init(lookup_t* lk, const orc_v1_node* nodes, uint64_t n) {
  lk->keys = (keyed_t*)malloc(sizeof(keyed_t) * (n ? n : 1));
  if (!lk->keys) return -1;
  lk->n = n;
  for (uint64_t i = 0; i < n; ++i) {
    lk->keys[i].x = nodes[i].x;
    lk->keys[i].y = nodes[i].y;
    lk->keys[i].layer = nodes[i].layer;
    lk->keys[i].idx = (uint32_t)i;
  }
  qsort(lk->keys, n, sizeof(keyed_t), key_cmp);
  return 0;
}

static int64_t lookup(const lookup_t* lk, uint32_t layer, int64_t x, int64_t y) {
  keyed_t k = {x, y, layer, 0};
  uint64_t lo = 0, hi = lk->n;
  while (lo < hi) {
    uint64_t mid = lo + (hi - lo) / 2;
    int c = key_cmp(&lk->keys[mid], &k);
    if (c == 0) return lk->keys[mid].idx;
    if (c < 0) lo = mid + 1; else hi = mid;
  }
  return -1;
}

/* ------------------------------------------------------- policy lookup -- */
/* canonical physical edge (CanonicalPhysicalEdgeResource, candidate_policy.cc:72-116) */
static void canonical_edge(int64_t x, int64_t y, int dir, int64_t* cx, int64_t* cy, int* cd) {
  *cx = x; *cy = y; *cd = dir;
  switch (dir) {
    case 4: *cx = x - 1; *cd = 0; break;
    case 5: *cx = x - 1; *cy = y - 1; *cd = 1; break;
    case 6: *cy = y - 1; *cd = 2; break;
    case 7: *cx = x + 1; *cy = y - 1; *cd = 3; break;
    default: break;
  }
}

static int policy_find(const int64_t* keys, uint32_t n, int64_t x, int64_t y, int d) {
  for (uint32_t i = 0; i < n; ++i)
    if (keys[3 * i] == x && keys[3 * i + 1] == y && keys[3 * i + 2] == d) return (int)i;
  return -1;
}

/* Returns 0 and the transition cost, or -1 if the edge is banned/overflows. */
static int transition(const orc_v1_profile* pr, const orc_v1_policy* po, int64_t x, int64_t y,
                      int dir, int incoming, uint64_t* out) {
  int64_t cx, cy;
  int cd;
  canonical_edge(x, y, dir, &cx, &cy, &cd);
  if (po && policy_find(po->ban_keys, po->n_bans, cx, cy, cd) >= 0) return -1;
  unsigned __int128 c = (dir & 1) ? (unsigned __int128)pr->diag + (po ? po->diag_s : 0)
                                  : (unsigned __int128)pr->orth + (po ? po->orth_s : 0);
  if (incoming != 8 && incoming != dir) c += (unsigned __int128)pr->bend + (po ? po->bend_s : 0);
  if (po) {
    /* duplicate penalty entries merge by summation (NormalizeCandidateGenerationPolicy) */
    for (uint32_t i = 0; i < po->n_pens; ++i)
      if (po->pen_keys[3 * i] == cx && po->pen_keys[3 * i + 1] == cy && po->pen_keys[3 * i + 2] == cd)
        c += po->pen_vals[i];
  }
  if (c >= (unsigned __int128)INF64) return -1;
  *out = (uint64_t)c;
  return 0;
}

static int checked_add(uint64_t a, uint64_t b, uint64_t* out) {
  const uint64_t maxf = INF64 - 1;
  if (a > maxf || b > maxf - a) return -1;
  *out = a + b;
  return 0;
}

static int heading_allowed(const orc_v1_profile* pr, int dir) {
  int h = (DY[dir] == 0) ? 1 : (DX[dir] == 0 ? 2 : 4);
  return (pr->heading_mask & h) != 0;
}

/* ------------------------------------------------------------- heap -- */
typedef struct {
  uint64_t k[7];
} item_t;
typedef struct {
  item_t* a;
  uint64_t n, cap;
} pq_t;

/* Signed coordinates in an unsigned key: flipping the sign bit keeps int64
 * order, as the reference's std::tie over int64 y and x does
 * (src/routing/cpu_astar.cc:67-74). */
static uint64_t coord_key(int64_t v) { return (uint64_t)v ^ 0x8000000000000000ull; }
static int64_t key_coord(uint64_t k) { return (int64_t)(k ^ 0x8000000000000000ull); }

static int item_less(const item_t* p, const item_t* q) {
  for (int i = 0; i < 7; ++i)
    if (p->k[i] != q->k[i]) return p->k[i] < q->k[i];
  return 0;
}
static int pq_push(pq_t* h, item_t it) {
  if (h->n == h->cap) {
    uint64_t nc = h->cap ? h->cap * 2 : 64;
    item_t* na = (item_t*)realloc(h->a, sizeof(item_t) * nc);
    if (!na) return -1;
    h->a = na;
    h->cap = nc;
  }
  uint64_t i = h->n++;
  h->a[i] = it;
  while (i) {
    uint64_t p = (i - 1) / 2;
    if (!item_less(&h->a[i], &h->a[p])) break;
    item_t t = h->a[i]; h->a[i] = h->a[p]; h->a[p] = t;
    i = p;
  }
  return 0;
}
static item_t pq_pop(pq_t* h) {
  item_t top = h->a[0];
  h->a[0] = h->a[--h->n];
  uint64_t i = 0;
  for (;;) {
    uint64_t l = 2 * i + 1, r = l + 1, m = i;
    if (l < h->n && item_less(&h->a[l], &h->a[m])) m = l;
    if (r < h->n && item_less(&h->a[r], &h->a[m])) m = r;
    if (m == i) break;
    item_t t = h->a[i]; h->a[i] = h->a[m]; h->a[m] = t;
    i = m;
  }
  return top;
}

/* Heuristic (cpu_astar.cc:159-212) */
static uint64_t heuristic(const orc_v1_profile* pr, const orc_v1_policy* po, int64_t x, int64_t y,
                          int64_t gx, int64_t gy) {
  uint64_t dx = x > gx ? (uint64_t)(x - gx) : (uint64_t)(gx - x);
  uint64_t dy = y > gy ? (uint64_t)(y - gy) : (uint64_t)(gy - y);
  int hz = pr->heading_mask & 1, vt = pr->heading_mask & 2, dg = pr->heading_mask & 4;
  uint64_t orth = (uint64_t)pr->orth + (po ? po->orth_s : 0);
  uint64_t diag = (uint64_t)pr->diag + (po ? po->diag_s : 0);
  unsigned __int128 e = 0;
  uint64_t mx = dx > dy ? dx : dy, mn = dx < dy ? dx : dy;
  if (hz && vt) {
    if (dg) {
      if (diag < orth) {
        e = (unsigned __int128)mx * diag;
      } else {
        uint64_t eff = (unsigned __int128)diag < (unsigned __int128)orth * 2 ? diag : orth * 2;
        e = (unsigned __int128)mn * eff + (unsigned __int128)(mx - mn) * orth;
      }
    } else {
      e = ((unsigned __int128)dx + dy) * orth;
    }
  } else if (hz && !vt && !dg) {
    e = (unsigned __int128)dx * orth;
  } else if (vt && !hz && !dg) {
    e = (unsigned __int128)dy * orth;
  } else if (dg) {
    uint64_t m = diag;
    if ((hz || vt) && orth < m) m = orth;
    e = (unsigned __int128)mx * m;
  } else {
    return 0;
  }
  if (e > (unsigned __int128)INF64) return 0;
  return (uint64_t)e;
}

int orc_v1_astar(const orc_v1_node* nodes, uint64_t n, const orc_v1_profile* pr,
                 const orc_v1_policy* po, uint32_t layer, int64_t sx, int64_t sy, int64_t gx,
                 int64_t gy, uint64_t* cost_out, int64_t* path_xy, uint64_t path_cap,
                 uint64_t* path_len, uint64_t* expanded_out) {
  lookup_t lk;
  if (lookup_init(&lk, nodes, n)) return -1;
  int64_t s = lookup(&lk, layer, sx, sy), g = lookup(&lk, layer, gx, gy);
  if (s < 0 || g < 0) { free(lk.keys); return -1; }
  uint64_t S = n * 9;
  uint64_t* best = (uint64_t*)malloc(sizeof(uint64_t) * S);
  int64_t* pred = (int64_t*)malloc(sizeof(int64_t) * S);
  if (!best || !pred) { free(best); free(pred); free(lk.keys); return -1; }
  for (uint64_t i = 0; i < S; ++i) { best[i] = INF64; pred[i] = -1; }
  pq_t q = {0, 0, 0};
  uint64_t seq = 0, expanded = 0;
  uint64_t h0 = heuristic(pr, po, sx, sy, gx, gy);
  best[s * 9 + 8] = 0;
  item_t it0 = {{h0, h0, 0, coord_key(sy), coord_key(sx), 8, seq++}};
  pq_push(&q, it0);
  int64_t goal_state = -1;
  int err = 0;
  while (q.n) {
    item_t cur = pq_pop(&q);
    int64_t cx = key_coord(cur.k[4]), cy = key_coord(cur.k[3]);
    int inc = (int)cur.k[5];
    int64_t u = lookup(&lk, layer, cx, cy);
    uint64_t st = (uint64_t)u * 9 + inc;
    if (best[st] != cur.k[2]) continue;               /* lazy deletion */
    if (cx == gx && cy == gy) { goal_state = (int64_t)st; break; }  /* goal on pop */
    ++expanded;
    for (int d = 0; d < 8; ++d) {                      /* stable direction order */
      if (!heading_allowed(pr, d)) continue;
      if (!(nodes[u].mask & (1u << d))) continue;
      int64_t vx = cx + DX[d], vy = cy + DY[d];
      int64_t v = lookup(&lk, layer, vx, vy);
      if (v < 0) { err = 1; break; }
      uint64_t w;
      if (transition(pr, po, cx, cy, d, inc, &w)) continue;   /* banned */
      uint64_t nc;
      if (checked_add(cur.k[2], w, &nc)) { err = 1; break; }
      uint64_t vs = (uint64_t)v * 9 + d;
      if (best[vs] <= nc) continue;                    /* strict improvement */
      best[vs] = nc;
      pred[vs] = (int64_t)st;
      uint64_t hv = heuristic(pr, po, vx, vy, gx, gy);
      uint64_t f;
      if (checked_add(nc, hv, &f)) f = INF64;
      item_t it = {{f, hv, nc, coord_key(vy), coord_key(vx), (uint64_t)d, seq++}};
      if (pq_push(&q, it)) { err = 1; break; }
    }
    if (err) break;
  }
  if (expanded_out) *expanded_out = expanded;
  int rc = 0;
  if (err) rc = -1;
  else if (goal_state < 0) { rc = 1; if (cost_out) *cost_out = INF64; }
  else {
    if (cost_out) *cost_out = best[goal_state];
    uint64_t k = 0;
    for (int64_t c = goal_state; c >= 0 && k < path_cap; c = pred[c]) {
      if (path_xy) {
        path_xy[2 * k] = nodes[c / 9].x;
        path_xy[2 * k + 1] = nodes[c / 9].y;
      }
      ++k;
    }
    if (path_len) *path_len = k;
  }
  free(q.a);
  free(best);
  free(pred);
  free(lk.keys);
  return rc;
}

/* -------------------------------------------- flattened order & labels -- */
static int64_t floordiv(int64_t a, int64_t b) {
  int64_t q = a / b;
  if ((a % b) < 0) --q;
  return q;
}

typedef struct {
  uint32_t layer;
  int64_t tx, ty;
  uint64_t local;
  uint32_t idx;
} tileord_t;

static int tileord_cmp(const void* a, const void* b) {
  const tileord_t* p = (const tileord_t*)a;
  const tileord_t* q = (const tileord_t*)b;
  if (p->layer != q->layer) return p->layer < q->layer ? -1 : 1;
  if (p->tx != q->tx) return p->tx < q->tx ? -1 : 1;
  if (p->ty != q->ty) return p->ty < q->ty ? -1 : 1;
  if (p->local != q->local) return p->local < q->local ? -1 : 1;
  return 0;
}

int orc_v1_flatten_order(const orc_v1_node* nodes, uint64_t n, uint32_t tile_w, uint32_t tile_h,
                         uint32_t* order) {
  /* Sparse-tile order: TileKey (layer, tile x, tile y), then row-major local
   * index inside the tile (compiled_board.cc:240-266, device_board.cc:167-193). */
  tileord_t* t = (tileord_t*)malloc(sizeof(tileord_t) * (n ? n : 1));
  if (!t) return -1;
  for (uint64_t i = 0; i < n; ++i) {
    int64_t tx = floordiv(nodes[i].x, tile_w), ty = floordiv(nodes[i].y, tile_h);
    uint64_t lx = (uint64_t)(nodes[i].x - tx * (int64_t)tile_w);
    uint64_t ly = (uint64_t)(nodes[i].y - ty * (int64_t)tile_h);
    t[i].layer = nodes[i].layer;
    t[i].tx = tx;
    t[i].ty = ty;
    t[i].local = ly * tile_w + lx;
    t[i].idx = (uint32_t)i;
  }
  qsort(t, n, sizeof(tileord_t), tileord_cmp);
  for (uint64_t i = 0; i < n; ++i) order[i] = t[i].idx;   /* flattened position -> input node */
  free(t);
  return 0;
}

/* Exact labels over (flattened node, heading) states by Dijkstra, then the
 * lowest-predecessor-state walk of gpu/cuda/planar_router.cu:1084-1165.
 * states_out gets the reverse state path (goal..start) in flattened numbering. */
int orc_v1_canonical(const orc_v1_node* nodes, uint64_t n, const orc_v1_profile* pr,
                     const orc_v1_policy* po, uint32_t layer, int64_t sx, int64_t sy, int64_t gx,
                     int64_t gy, uint64_t* cost_out, uint32_t* states_out, uint64_t cap,
                     uint64_t* len_out) {
  uint32_t* order = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
  uint32_t* pos = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
  orc_v1_node* flat = (orc_v1_node*)malloc(sizeof(orc_v1_node) * (n ? n : 1));
  if (!order || !pos || !flat || orc_v1_flatten_order(nodes, n, pr->tile_w, pr->tile_h, order)) {
    free(order); free(pos); free(flat); return -1;
  }
  for (uint64_t i = 0; i < n; ++i) { flat[i] = nodes[order[i]]; pos[order[i]] = (uint32_t)i; }
  lookup_t lk;
  if (lookup_init(&lk, flat, n)) { free(order); free(pos); free(flat); return -1; }
  int64_t s = lookup(&lk, layer, sx, sy), g = lookup(&lk, layer, gx, gy);
  uint64_t S = n * 9;
  uint64_t* L = (uint64_t*)malloc(sizeof(uint64_t) * (S ? S : 1));
  int rc = 0;
  if (s < 0 || g < 0 || !L) { rc = -1; goto done; }
  for (uint64_t i = 0; i < S; ++i) L[i] = INF64;
  {
    pq_t q = {0, 0, 0};
    L[s * 9 + 8] = 0;
    item_t it0 = {{0, (uint64_t)s * 9 + 8, 0, 0, 0, 0, 0}};
    pq_push(&q, it0);
    while (q.n) {
      item_t cur = pq_pop(&q);
      uint64_t st = cur.k[1];
      if (L[st] != cur.k[0]) continue;
      uint64_t u = st / 9;
      int inc = (int)(st % 9);
      for (int d = 0; d < 8; ++d) {
        if (!heading_allowed(pr, d) || !(flat[u].mask & (1u << d))) continue;
        int64_t v = lookup(&lk, layer, flat[u].x + DX[d], flat[u].y + DY[d]);
        if (v < 0) { rc = -1; break; }
        uint64_t w, nc;
        if (transition(pr, po, flat[u].x, flat[u].y, d, inc, &w)) continue;
        if (checked_add(cur.k[0], w, &nc)) continue;
        uint64_t vs = (uint64_t)v * 9 + d;
        if (nc < L[vs]) {
          L[vs] = nc;
          item_t it = {{nc, vs, 0, 0, 0, 0, 0}};
          if (pq_push(&q, it)) { rc = -1; break; }
        }
      }
      if (rc) break;
    }
    free(q.a);
  }
  if (rc) goto done;
  {
    /* goal state: lowest state index among the minimal goal headings 0..7 */
    uint64_t best = INF64, bs = 0;
    for (int h = 0; h < 8; ++h) {
      uint64_t st = (uint64_t)g * 9 + h;
      if (L[st] < best) { best = L[st]; bs = st; }
    }
    if (best == INF64) { rc = 1; if (cost_out) *cost_out = INF64; goto done; }
    if (cost_out) *cost_out = best;
    uint64_t k = 0, cur = bs, start_state = (uint64_t)s * 9 + 8;
    for (;;) {
      if (k >= cap || k > S) { rc = -1; break; }
      states_out[k++] = (uint32_t)cur;
      if (cur == start_state) break;
      uint64_t v = cur / 9;
      int inc = (int)(cur % 9);
      if (inc == 8) { rc = -1; break; }
      int opp = (inc + 4) % 8;
      int64_t p = lookup(&lk, layer, flat[v].x + DX[opp], flat[v].y + DY[opp]);
      if (p < 0 || !(flat[p].mask & (1u << inc))) { rc = -1; break; }
      uint64_t chosen = INF64;
      for (int ph = 0; ph < 9; ++ph) {
        if (ph == 8 && p != s) continue;
        uint64_t ps = (uint64_t)p * 9 + ph;
        if (L[ps] == INF64) continue;
        uint64_t w, c;
        if (transition(pr, po, flat[p].x, flat[p].y, inc, ph, &w)) continue;
        if (checked_add(L[ps], w, &c)) continue;
        if (c == L[cur] && ps < chosen) chosen = ps;
      }
      if (chosen == INF64) { rc = -1; break; }
      cur = chosen;
    }
    if (len_out) *len_out = k;
  }
done:
  free(L);
  free(lk.keys);
  free(order);
  free(pos);
  free(flat);
  return rc;
}

/* Restatement of the batch frontier generator (CandidateFrontierChunk,
 * gpu/cuda/planar_router.cu:715-845) in flattened numbering: each round takes
 * the open finite state with the least (f, g, state), f = g + max(|dx|,|dy|) *
 * min(orth + s_o, diag + s_d) (DeviceAdmissibleHeuristic :590-608, winner
 * order :610-619), counts the round, closes it, stops at the goal node, and
 * otherwise relaxes every legal direction (each counted as examined work;
 * labels drop only on strict improvement).  Then the lowest-predecessor walk
 * of :1084-1165 on the resulting labels.  Returns 0 reached, 1 disconnected,
 * 2 round budget exhausted, -1 error. */
static int frontier_impl(const orc_v1_node* nodes, uint64_t n, const orc_v1_profile* pr,
                         const orc_v1_policy* po, uint32_t layer, int64_t sx, int64_t sy, int64_t gx,
                         int64_t gy, uint32_t max_rounds, uint64_t* cost_out, uint32_t* goal_state_out,
                         uint32_t* rounds_out, uint64_t* examined_out, uint32_t* states_out, uint64_t cap,
                         uint64_t* len_out, uint64_t* labels_out, uint32_t* preds_out) {
  uint32_t* order = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
  orc_v1_node* flat = (orc_v1_node*)malloc(sizeof(orc_v1_node) * (n ? n : 1));
  if (!order || !flat || orc_v1_flatten_order(nodes, n, pr->tile_w, pr->tile_h, order)) {
    free(order); free(flat); return -1;
  }
  for (uint64_t i = 0; i < n; ++i) flat[i] = nodes[order[i]];
  lookup_t lk;
  if (lookup_init(&lk, flat, n)) { free(order); free(flat); return -1; }
  int64_t s = lookup(&lk, layer, sx, sy), g = lookup(&lk, layer, gx, gy);
  uint64_t S = n * 9;
  uint64_t* L = (uint64_t*)malloc(sizeof(uint64_t) * (S ? S : 1));
  uint8_t* closed = (uint8_t*)calloc(S ? S : 1, 1);
  int rc = 2;
  uint32_t rounds = 0, goal_state = UINT32_MAX;
  uint64_t examined = 0;
  if (s < 0 || g < 0 || !L || !closed) { rc = -1; goto done; }
  for (uint64_t i = 0; i < S; ++i) L[i] = INF64;
  L[s * 9 + 8] = 0;
  const uint64_t orth = (uint64_t)pr->orth + (po ? po->orth_s : 0);
  const uint64_t diag = (uint64_t)pr->diag + (po ? po->diag_s : 0);
  const uint64_t step = orth < diag ? orth : diag;
  const int64_t goal_x = flat[g].x, goal_y = flat[g].y;
  while (rounds < max_rounds) {
    uint64_t bf = INF64, bg = INF64, bs = UINT64_MAX;
    for (uint64_t st = 0; st < S; ++st) {
      if (closed[st] || L[st] == INF64) continue;
      const int64_t x = flat[st / 9].x, y = flat[st / 9].y;
      const uint64_t dx = x >= goal_x ? (uint64_t)(x - goal_x) : (uint64_t)(goal_x - x);
      const uint64_t dy = y >= goal_y ? (uint64_t)(y - goal_y) : (uint64_t)(goal_y - y);
      const uint64_t steps = dx > dy ? dx : dy;
      const uint64_t h = (steps != 0 && step > (INF64 - 1) / steps) ? INF64 - 1 : steps * step;
      uint64_t f;
      if (checked_add(L[st], h, &f)) f = INF64 - 1;
      if (f < bf || (f == bf && (L[st] < bg || (L[st] == bg && st < bs)))) {
        bf = f; bg = L[st]; bs = st;
      }
    }
    ++rounds;
    if (bs == UINT64_MAX) { rc = 1; break; }
    closed[bs] = 1;
    const uint64_t u = bs / 9;
    const int inc = (int)(bs % 9);
    if ((int64_t)u == g) { goal_state = (uint32_t)bs; rc = 0; break; }
    for (int d = 0; d < 8; ++d) {
      if (!(flat[u].mask & (1u << d))) continue;
      ++examined;
      uint64_t w, c;
      if (transition(pr, po, flat[u].x, flat[u].y, d, inc, &w)) continue;
      if (checked_add(bg, w, &c)) continue;
      int64_t v = lookup(&lk, layer, flat[u].x + DX[d], flat[u].y + DY[d]);
      if (v < 0) { rc = -1; goto done; }
      uint64_t vs = (uint64_t)v * 9 + d;
      if (c < L[vs]) L[vs] = c;
    }
  }
  if (rc == 0) {
    if (cost_out) *cost_out = L[goal_state];
    uint64_t k = 0, cur = goal_state, start_state = (uint64_t)s * 9 + 8;
    for (;;) {
      if (k >= cap || k > S) { rc = -1; break; }
      states_out[k++] = (uint32_t)cur;
      if (cur == start_state) break;
      uint64_t v = cur / 9;
      int in = (int)(cur % 9);
      if (in == 8) { rc = -1; break; }
      int opp = (in + 4) % 8;
      int64_t p = lookup(&lk, layer, flat[v].x + DX[opp], flat[v].y + DY[opp]);
      if (p < 0) { rc = -1; break; }
      uint64_t chosen = INF64;
      for (int ph = 0; ph < 9; ++ph) {
        if (ph == 8 && p != s) continue;
        uint64_t ps = (uint64_t)p * 9 + ph;
        if (L[ps] == INF64) continue;
        uint64_t w, c;
        if (transition(pr, po, flat[p].x, flat[p].y, in, ph, &w)) continue;
        if (checked_add(L[ps], w, &c)) continue;
        if (c == L[cur] && ps < chosen) chosen = ps;
      }
      if (chosen == INF64) { rc = -1; break; }
      cur = chosen;
    }
    if (len_out) *len_out = k;
  }
  if (rc >= 0 && labels_out && preds_out) {
    /* Final labels and the lowest tight predecessor of every state
     * (SelectCandidateBatchPredecessors, gpu/cuda/planar_router.cu:1021-1082). */
    const uint64_t start_state = (uint64_t)s * 9 + 8;
    for (uint64_t st = 0; st < S; ++st) {
      labels_out[st] = L[st];
      preds_out[st] = UINT32_MAX;
      const int in = (int)(st % 9);
      if (st == start_state || in == 8 || L[st] == INF64) continue;
      const uint64_t v = st / 9;
      const int opp = (in + 4) % 8;
      const int64_t p = lookup(&lk, layer, flat[v].x + DX[opp], flat[v].y + DY[opp]);
      if (p < 0 || !(flat[p].mask & (1u << in))) continue;
      uint64_t chosen = INF64;
      for (int ph = 0; ph < 9; ++ph) {
        if (ph == 8 && p != s) continue;
        const uint64_t ps = (uint64_t)p * 9 + ph;
        uint64_t w, c;
        if (L[ps] == INF64 || transition(pr, po, flat[p].x, flat[p].y, in, ph, &w)) continue;
        if (checked_add(L[ps], w, &c)) continue;
        if (c == L[st] && ps < chosen) chosen = ps;
      }
      if (chosen != INF64) preds_out[st] = (uint32_t)chosen;
    }
  }
done:
  if (rounds_out) *rounds_out = rounds;
  if (examined_out) *examined_out = examined;
  if (goal_state_out) *goal_state_out = goal_state;
  free(L);
  free(closed);
  free(lk.keys);
  free(order);
  free(flat);
  return rc;
}

int orc_v1_frontier(const orc_v1_node* nodes, uint64_t n, const orc_v1_profile* pr,
                    const orc_v1_policy* po, uint32_t layer, int64_t sx, int64_t sy, int64_t gx,
                    int64_t gy, uint32_t max_rounds, uint64_t* cost_out, uint32_t* goal_state_out,
                    uint32_t* rounds_out, uint64_t* examined_out, uint32_t* states_out, uint64_t cap,
                    uint64_t* len_out) {
  return frontier_impl(nodes, n, pr, po, layer, sx, sy, gx, gy, max_rounds, cost_out, goal_state_out,
                       rounds_out, examined_out, states_out, cap, len_out, NULL, NULL);
}

int orc_v1_frontier_workspace(const orc_v1_node* nodes, uint64_t n, const orc_v1_profile* pr,
                              const orc_v1_policy* po, uint32_t layer, int64_t sx, int64_t sy, int64_t gx,
                              int64_t gy, uint32_t max_rounds, uint32_t* goal_state_out, uint32_t* rounds_out,
                              uint64_t* labels_out, uint32_t* preds_out) {
  uint64_t len = 0;
  uint32_t* states = (uint32_t*)malloc(sizeof(uint32_t) * (n * 9 + 2));
  if (!states) return -1;
  const int rc = frontier_impl(nodes, n, pr, po, layer, sx, sy, gx, gy, max_rounds, NULL, goal_state_out,
                               rounds_out, NULL, states, n * 9 + 2, &len, labels_out, preds_out);
  free(states);
  return rc;
}

/* Single-route generators of the reference's ExecuteRoute in flattened
 * numbering, with the profile's step and bend costs only (no candidate
 * policy).  Frontier: RunFrontier / FrontierRelax
 * (gpu/cuda/planar_router.cu:363-412, 1552-1640), delta-stepping over heading
 * states with bucket width min(orth, diag); a state is active in a round when
 * the previous round deferred it or lowered its label; active states in the
 * bucket relax every legal direction (each counted as examined work), the
 * others are deferred; converged when nothing is deferred or lowered.  Sweep:
 * RunSweep / ComputeDepartureLabels / SweepRuns (:414-484, :1642-1722),
 * departures then segmented scans along every maximal directed run; examined
 * work is the legal directed edges per round, a turn relaxation is counted for
 * each departure whose final minimum came from another heading; converged
 * when a round lowers no label.  Then SelectStablePredecessors (:486-524).
 * Sequential emulation of a round equals the parallel one: every candidate
 * made inside a bucket lies above it, and every run writes its own labels. */
static uint64_t route_step(const orc_v1_profile* pr, int dir, int incoming) {
  uint64_t c = (dir & 1) ? pr->diag : pr->orth;
  if (incoming != 8 && incoming != dir) c += pr->bend;
  return c;
}

int orc_v1_route(const orc_v1_node* nodes, uint64_t n, const orc_v1_profile* pr, uint32_t generator, uint32_t layer,
                 int64_t sx, int64_t sy, int64_t gx, int64_t gy, uint32_t max_rounds, uint32_t* completion_out,
                 uint32_t* rounds_out, uint64_t* examined_out, uint64_t* turns_out, uint32_t* goal_state_out,
                 uint64_t* labels_out, uint32_t* preds_out) {
  uint32_t* order = (uint32_t*)malloc(sizeof(uint32_t) * (n ? n : 1));
  orc_v1_node* flat = (orc_v1_node*)malloc(sizeof(orc_v1_node) * (n ? n : 1));
  if (!order || !flat || orc_v1_flatten_order(nodes, n, pr->tile_w, pr->tile_h, order)) {
    free(order); free(flat); return -1;
  }
  for (uint64_t i = 0; i < n; ++i) flat[i] = nodes[order[i]];
  lookup_t lk;
  if (lookup_init(&lk, flat, n)) { free(order); free(flat); return -1; }
  const int64_t s = lookup(&lk, layer, sx, sy), g = lookup(&lk, layer, gx, gy);
  const uint64_t S = n * 9;
  uint64_t* L = (uint64_t*)malloc(sizeof(uint64_t) * (S ? S : 1));
  uint64_t* dep = (uint64_t*)malloc(sizeof(uint64_t) * (n ? n * 8 : 1));
  int64_t* nb = (int64_t*)malloc(sizeof(int64_t) * (n ? n * 8 : 1));
  uint8_t* cur = (uint8_t*)calloc(S ? S : 1, 1);
  uint8_t* nxt = (uint8_t*)calloc(S ? S : 1, 1);
  int rc = 0;
  uint32_t rounds = 0, completion = 2;
  uint64_t examined = 0, turns = 0, legal = 0;
  if (s < 0 || g < 0 || s == g || !L || !dep || !nb || !cur || !nxt) { rc = -1; goto done; }
  for (uint64_t u = 0; u < n; ++u)
    for (int d = 0; d < 8; ++d) {
      nb[u * 8 + d] = -1;
      if (!(flat[u].mask & (1u << d))) continue;
      nb[u * 8 + d] = lookup(&lk, layer, flat[u].x + DX[d], flat[u].y + DY[d]);
      if (nb[u * 8 + d] < 0) { rc = -1; goto done; }
      ++legal;
    }
  for (uint64_t i = 0; i < S; ++i) L[i] = INF64;
  const uint64_t start_state = (uint64_t)s * 9 + 8;
  L[start_state] = 0;
  int converged = 0;
  if (generator == 0) {
    const uint64_t delta = pr->orth < pr->diag ? pr->orth : pr->diag;
    if (delta == 0) { rc = -1; goto done; }
    uint64_t bucket_upper = 0;
    cur[start_state] = 1;
    while (rounds < max_rounds) {
      uint64_t next_min = INF64;
      memset(nxt, 0, S);
      for (uint64_t st = 0; st < S; ++st) {
        if (!cur[st] || L[st] == INF64) continue;
        const uint64_t label = L[st];
        if (label > bucket_upper) {
          nxt[st] = 1;
          if (label < next_min) next_min = label;
          continue;
        }
        const uint64_t u = st / 9;
        const int inc = (int)(st % 9);
        for (int d = 0; d < 8; ++d) {
          if (nb[u * 8 + d] < 0) continue;
          ++examined;
          uint64_t c;
          if (checked_add(label, route_step(pr, d, inc), &c)) continue;
          const uint64_t t = (uint64_t)nb[u * 8 + d] * 9 + d;
          if (c < L[t]) {
            L[t] = c;
            nxt[t] = 1;
            if (c < next_min) next_min = c;
          }
        }
      }
      ++rounds;
      if (next_min == INF64) { converged = 1; break; }
      const uint64_t base = (next_min / delta) * delta;
      bucket_upper = base > INF64 - (delta - 1) ? INF64 : base + delta - 1;
      uint8_t* tmp = cur; cur = nxt; nxt = tmp;
    }
  } else {
    while (rounds < max_rounds) {
      for (uint64_t u = 0; u < n; ++u)
        for (int d = 0; d < 8; ++d) {
          uint64_t best = (int64_t)u == s ? 0 : L[u * 9 + d];
          int used_turn = 0;
          for (int inc = 0; inc < 8; ++inc) {
            const uint64_t lab = L[u * 9 + inc];
            uint64_t c;
            if (lab == INF64 || checked_add(lab, inc == d ? 0 : pr->bend, &c)) continue;
            if (c < best) { best = c; used_turn = inc != d; }
          }
          dep[u * 8 + d] = best;
          turns += used_turn ? 1 : 0;
        }
      int changed = 0;
      for (int d = 0; d < 8; ++d) {
        const int opp = (d + 4) % 8;
        for (uint64_t u = 0; u < n; ++u) {
          if (nb[u * 8 + d] < 0) continue;
          const int64_t before = nb[u * 8 + opp] >= 0 ? nb[u * 8 + opp] : lookup(&lk, layer, flat[u].x - DX[d], flat[u].y - DY[d]);
          if (before >= 0 && nb[(uint64_t)before * 8 + d] == (int64_t)u) continue; /* not a run start */
          uint64_t carry = INF64, v = u;
          while (nb[v * 8 + d] >= 0) {
            const uint64_t t = (uint64_t)nb[v * 8 + d];
            if (dep[v * 8 + d] < carry) carry = dep[v * 8 + d];
            uint64_t c;
            if (!checked_add(carry, (d & 1) ? pr->diag : pr->orth, &c)) {
              if (c < L[t * 9 + d]) { L[t * 9 + d] = c; changed = 1; }
              carry = c < dep[t * 8 + d] ? c : dep[t * 8 + d];
            } else {
              carry = dep[t * 8 + d];
            }
            v = t;
          }
        }
      }
      ++rounds;
      examined += legal;
      if (!changed) { converged = 1; break; }
    }
  }
  if (converged) {
    completion = 1;
    for (int h = 0; h < 8; ++h)
      if (L[(uint64_t)g * 9 + h] != INF64) completion = 0;
  }
  if (goal_state_out) {
    *goal_state_out = UINT32_MAX;
    if (completion == 0) {
      uint64_t best = INF64;
      for (int h = 0; h < 8; ++h) {
        const uint64_t st = (uint64_t)g * 9 + h;
        if (L[st] < best) { best = L[st]; *goal_state_out = (uint32_t)st; }
      }
    }
  }
  for (uint64_t st = 0; st < S; ++st) {
    if (labels_out) labels_out[st] = L[st];
    if (!preds_out) continue;
    preds_out[st] = UINT32_MAX;
    const int in = (int)(st % 9);
    if (st == start_state || in == 8 || L[st] == INF64) continue;
    const uint64_t v = st / 9;
    const int opp = (in + 4) % 8;
    const int64_t p = lookup(&lk, layer, flat[v].x + DX[opp], flat[v].y + DY[opp]);
    if (p < 0) continue;
    uint64_t chosen = INF64;
    for (int ph = 0; ph < 9; ++ph) {
      if (ph == 8 && p != s) continue;
      const uint64_t ps = (uint64_t)p * 9 + ph;
      uint64_t c;
      if (L[ps] == INF64 || checked_add(L[ps], route_step(pr, in, ph), &c)) continue;
      if (c == L[st] && ps < chosen) chosen = ps;
    }
    if (chosen != INF64) preds_out[st] = (uint32_t)chosen;
  }
done:
  if (completion_out) *completion_out = completion;
  if (rounds_out) *rounds_out = rounds;
  if (examined_out) *examined_out = examined;
  if (turns_out) *turns_out = turns;
  free(L); free(dep); free(nb); free(cur); free(nxt);
  free(lk.keys);
  free(order);
  free(flat);
  return rc;
}
