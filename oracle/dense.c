/*
 * dense.c — CPU oracle for the dense multi-layer lattice.  TEST INFRASTRUCTURE
 * ONLY: loaded by tests/, the smoke check and the benchmark's CPU-baseline leg;
 * never by the product library.
 */
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

/* ------------------------------------------------------------------ heap -- */
/* Generic binary min-heap over fixed-size keys compared lexicographically as
 * arrays of uint64. */
typedef struct {
  uint64_t* keys; /* n * width */
  uint64_t n, cap;
  uint32_t width;
} heap_t;

static int heap_init(heap_t* h, uint32_t width, uint64_t cap) {
  h->width = width;
  h->n = 0;
  h->cap = cap ? cap : 16;
  h->keys = (uint64_t*)malloc(sizeof(uint64_t) * width * h->cap);
  return h->keys ? 0 : -1;
}
static void heap_free(heap_t* h) { free(h->keys); }

static int key_less(const uint64_t* a, const uint64_t* b, uint32_t w) {
  for (uint32_t i = 0; i < w; ++i) {
    if (a[i] != b[i]) return a[i] < b[i];
  }
  return 0;
}
static void key_swap(uint64_t* a, uint64_t* b, uint32_t w) {
  for (uint32_t i = 0; i < w; ++i) {
    uint64_t t = a[i];
    a[i] = b[i];
    b[i] = t;
  }
}
static int heap_push(heap_t* h, const uint64_t* key) {
  if (h->n == h->cap) {
    uint64_t ncap = h->cap * 2;
    uint64_t* k = (uint64_t*)realloc(h->keys, sizeof(uint64_t) * h->width * ncap);
    if (!k) return -1;
    h->keys = k;
    h->cap = ncap;
  }
  uint32_t w = h->width;
  uint64_t i = h->n++;
  memcpy(h->keys + i * w, key, sizeof(uint64_t) * w);
  while (i > 0) {
    uint64_t p = (i - 1) / 2;
    if (!key_less(h->keys + i * w, h->keys + p * w, w)) break;
    key_swap(h->keys + i * w, h->keys + p * w, w);
    i = p;
  }
  return 0;
}
static void heap_pop(heap_t* h, uint64_t* out) {
  uint32_t w = h->width;
  memcpy(out, h->keys, sizeof(uint64_t) * w);
  h->n--;
  if (h->n == 0) return;
  memcpy(h->keys, h->keys + h->n * w, sizeof(uint64_t) * w);
  uint64_t i = 0;
  for (;;) {
    uint64_t l = 2 * i + 1, r = l + 1, m = i;
    if (l < h->n && key_less(h->keys + l * w, h->keys + m * w, w)) m = l;
    if (r < h->n && key_less(h->keys + r * w, h->keys + m * w, w)) m = r;
    if (m == i) break;
    key_swap(h->keys + i * w, h->keys + m * w, w);
    i = m;
  }
}

/* --------------------------------------------------------- lattice model -- */
static int in_window(const orc_lattice* g, uint32_t x, uint32_t y) {
  if (g->ww == 0 && g->wh == 0) return 1;
  return x >= g->wx0 && x < g->wx0 + g->ww && y >= g->wy0 && y < g->wy0 + g->wh;
}

static int cell_free(const orc_lattice* g, uint32_t c) {
  uint32_t x = c % g->W, y = (c / g->W) % g->H;
  return g->cost[c] != 0 && in_window(g, x, y);
}

/* Weight of a via move into cell v: via, plus v's cost when via_pays_cell. */
static uint32_t via_weight(const orc_lattice* g, uint32_t v) {
  return g->via + (g->via_pays_cell ? g->cost[v] : 0u);
}

/* Neighbours of c in ascending cell-index order: l-1, y-1, x-1, x+1, y+1, l+1.
 * Writes (cell, weight of the move INTO the neighbour) pairs; returns count. */
static int neighbours(const orc_lattice* g, uint32_t c, uint32_t* nb, uint32_t* w_into) {
  uint32_t W = g->W, H = g->H, L = g->L;
  uint32_t x = c % W, y = (c / W) % H, l = c / (W * H);
  int n = 0;
  uint64_t plane = (uint64_t)W * H;
  if (g->via && l > 0) { nb[n] = (uint32_t)(c - plane); w_into[n++] = via_weight(g, (uint32_t)(c - plane)); }
  if (y > 0) { nb[n] = c - W; w_into[n++] = g->cost[c - W]; }
  if (x > 0) { nb[n] = c - 1; w_into[n++] = g->cost[c - 1]; }
  if (x + 1 < W) { nb[n] = c + 1; w_into[n++] = g->cost[c + 1]; }
  if (y + 1 < H) { nb[n] = c + W; w_into[n++] = g->cost[c + W]; }
  if (g->via && l + 1 < L) { nb[n] = (uint32_t)(c + plane); w_into[n++] = via_weight(g, (uint32_t)(c + plane)); }
  int k = 0;
  for (int i = 0; i < n; ++i) {
    if (cell_free(g, nb[i])) {
      nb[k] = nb[i];
      w_into[k] = w_into[i];
      ++k;
    }
  }
  return k;
}

/* weight of the move u -> v where v is a neighbour of u */
static uint32_t move_weight(const orc_lattice* g, uint32_t u, uint32_t v) {
  uint64_t plane = (uint64_t)g->W * g->H;
  uint64_t d = u > v ? (uint64_t)u - v : (uint64_t)v - u;
  if (d == plane) return via_weight(g, v);
  return g->cost[v];
}

int orc_dense_dijkstra(const orc_lattice* g, const uint32_t* sources, uint32_t nsrc,
                       uint32_t* dist) {
  uint64_t N = (uint64_t)g->W * g->H * g->L;
  for (uint64_t i = 0; i < N; ++i) dist[i] = ORC_UNREACHABLE;
  heap_t h;
  if (heap_init(&h, 2, 1024)) return ORC_ERR;
  for (uint32_t s = 0; s < nsrc; ++s) {
    if (sources[s] >= N || !cell_free(g, sources[s])) { heap_free(&h); return ORC_ERR; }
    if (dist[sources[s]] != 0) {
      dist[sources[s]] = 0;
      uint64_t k[2] = {0, sources[s]};
      if (heap_push(&h, k)) { heap_free(&h); return ORC_ERR; }
    }
  }
  uint32_t nb[6], w[6];
  while (h.n) {
    uint64_t k[2];
    heap_pop(&h, k);
    uint32_t u = (uint32_t)k[1];
    if (k[0] != dist[u]) continue;
    int n = neighbours(g, u, nb, w);
    for (int i = 0; i < n; ++i) {
      uint64_t nd = k[0] + w[i];
      if (nd < dist[nb[i]]) {
        dist[nb[i]] = (uint32_t)nd;
        uint64_t kk[2] = {nd, nb[i]};
        if (heap_push(&h, kk)) { heap_free(&h); return ORC_ERR; }
      }
    }
  }
  heap_free(&h);
  return ORC_OK;
}

int64_t orc_dense_canonical_path(const orc_lattice* g, const uint32_t* dist, uint32_t target,
                                 uint32_t* out, uint64_t cap) {
  uint64_t N = (uint64_t)g->W * g->H * g->L;
  if (target >= N || dist[target] == ORC_UNREACHABLE) return -1;
  uint64_t n = 0;
  uint32_t v = target;
  uint32_t nb[6], w[6];
  for (;;) {
    if (n >= cap || n > N) return -1;
    out[n++] = v;
    if (dist[v] == 0) return (int64_t)n;
    int k = neighbours(g, v, nb, w);
    int found = -1;
    for (int i = 0; i < k; ++i) { /* ascending index order: first match is lowest */
      uint32_t u = nb[i];
      if (dist[u] == ORC_UNREACHABLE) continue;
      if ((uint64_t)dist[u] + move_weight(g, u, v) == dist[v]) { found = (int)u; break; }
    }
    if (found < 0) return -1;
    v = (uint32_t)found;
  }
}

uint32_t orc_dense_astar(const orc_lattice* g, uint32_t source, uint32_t target, uint32_t* path,
                         uint64_t cap, uint64_t* path_len, uint64_t* expanded) {
  uint64_t N = (uint64_t)g->W * g->H * g->L;
  if (path_len) *path_len = 0;
  if (expanded) *expanded = 0;
  if (source >= N || target >= N || !cell_free(g, source) || !cell_free(g, target))
    return ORC_UNREACHABLE;
  /* minimum free cell cost in the window: an in-layer move costs at least this */
  uint32_t min_cost = 0xFFFFFFFFu;
  for (uint64_t c = 0; c < N; ++c)
    if (cell_free(g, (uint32_t)c) && g->cost[c] < min_cost) min_cost = g->cost[c];
  uint32_t W = g->W, H = g->H;
  uint32_t tx = target % W, ty = (target / W) % H, tl = target / (W * H);
#define HEUR(c)                                                                         \
  ({                                                                                    \
    uint32_t _x = (c) % W, _y = ((c) / W) % H, _l = (c) / (W * H);                      \
    uint64_t _dx = _x > tx ? _x - tx : tx - _x, _dy = _y > ty ? _y - ty : ty - _y;      \
    uint64_t _dl = _l > tl ? _l - tl : tl - _l;                                         \
    (_dx + _dy) * (uint64_t)min_cost + _dl * (uint64_t)g->via;                          \
  })
  uint64_t* rec = (uint64_t*)malloc(sizeof(uint64_t) * N);   /* best cost */
  uint32_t* pred = (uint32_t*)malloc(sizeof(uint32_t) * N);
  if (!rec || !pred) { free(rec); free(pred); return ORC_UNREACHABLE; }
  for (uint64_t i = 0; i < N; ++i) { rec[i] = UINT64_MAX; pred[i] = ORC_UNREACHABLE; }
  heap_t h;
  if (heap_init(&h, 7, 1024)) { free(rec); free(pred); return ORC_UNREACHABLE; }
  uint64_t seq = 0;
  rec[source] = 0;
  {
    uint64_t hs = HEUR(source);
    uint64_t k[7] = {hs, hs, 0, (source / W) % H, source % W, source / (W * H), seq++};
    heap_push(&h, k);
  }
  uint32_t result = ORC_UNREACHABLE;
  uint64_t nexp = 0;
  while (h.n) {
    uint64_t k[7];
    heap_pop(&h, k);
    uint32_t u = (uint32_t)((k[5] * H + k[3]) * W + k[4]);
    if (rec[u] != k[2]) continue;           /* lazy deletion of stale entries */
    if (u == target) { result = (uint32_t)k[2]; break; }  /* goal on first pop */
    ++nexp;
    /* fixed relaxation order: E, N, W, S, via up, via down */
    uint32_t x = u % W, y = (u / W) % H, l = u / (W * H);
    uint32_t cand[6];
    int nc = 0;
    if (x + 1 < W) cand[nc++] = u + 1;
    if (y + 1 < H) cand[nc++] = u + W;
    if (x > 0) cand[nc++] = u - 1;
    if (y > 0) cand[nc++] = u - W;
    if (g->via && l + 1 < g->L) cand[nc++] = u + W * H;
    if (g->via && l > 0) cand[nc++] = u - W * H;
    for (int i = 0; i < nc; ++i) {
      uint32_t v = cand[i];
      if (!cell_free(g, v)) continue;
      uint64_t nd = k[2] + move_weight(g, u, v);
      if (rec[v] <= nd) continue;             /* strict improvement only */
      rec[v] = nd;
      pred[v] = u;
      uint64_t hv = HEUR(v);
      uint64_t kk[7] = {nd + hv, hv, nd, (v / W) % H, v % W, v / (W * H), seq++};
      if (heap_push(&h, kk)) { nc = 0; h.n = 0; result = ORC_UNREACHABLE; break; }
    }
  }
#undef HEUR
  if (expanded) *expanded = nexp;
  if (result != ORC_UNREACHABLE && path) {
    uint64_t n = 0;
    uint32_t v = target;
    while (n < cap) {
      path[n++] = v;
      if (v == source) break;
      v = pred[v];
    }
    if (path_len) *path_len = n;
  }
  heap_free(&h);
  free(rec);
  free(pred);
  return result;
}

int orc_dense_route_net(const orc_lattice* g, const uint32_t* pins, uint32_t npins,
                        uint32_t* conn_pin, uint32_t* conn_dist, uint32_t* conn_len,
                        uint32_t* cells, uint64_t cell_cap, uint64_t* cells_used) {
  uint64_t N = (uint64_t)g->W * g->H * g->L;
  if (npins < 2 || npins > 32) return -1;
  uint32_t* tree = (uint32_t*)malloc(sizeof(uint32_t) * (N < 1 ? 1 : N));
  uint32_t* dist = (uint32_t*)malloc(sizeof(uint32_t) * N);
  uint8_t* in_tree = (uint8_t*)calloc(N, 1);
  if (!tree || !dist || !in_tree) { free(tree); free(dist); free(in_tree); return -1; }
  uint64_t ntree = 0, used = 0;
  tree[ntree++] = pins[0];
  in_tree[pins[0]] = 1;
  uint32_t connected = 1u;
  int status = 0;
  for (uint32_t step = 0; step + 1 < npins; ++step) {
    if (orc_dense_dijkstra(g, tree, (uint32_t)ntree, dist)) { status = -1; break; }
    uint32_t best = ORC_UNREACHABLE, best_pin = 0xFFFFFFFFu;
    for (uint32_t p = 0; p < npins; ++p) {
      if (connected & (1u << p)) continue;
      if (dist[pins[p]] < best) { best = dist[pins[p]]; best_pin = p; }
    }
    if (best == ORC_UNREACHABLE) { status = 1; break; }
    int64_t len = orc_dense_canonical_path(g, dist, pins[best_pin], cells + used, cell_cap - used);
    if (len < 0) { status = -1; break; }
    conn_pin[step] = best_pin;
    conn_dist[step] = best;
    conn_len[step] = (uint32_t)len;
    for (int64_t i = 0; i < len; ++i) {
      uint32_t c = cells[used + i];
      if (!in_tree[c]) { in_tree[c] = 1; tree[ntree++] = c; }
    }
    used += (uint64_t)len;
    connected |= 1u << best_pin;
    /* pins that coincide with tree cells are connected at distance 0 by the
     * next search; nothing special to do here */
  }
  if (cells_used) *cells_used = used;
  free(tree);
  free(dist);
  free(in_tree);
  return status;
}
