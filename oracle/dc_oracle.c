/*
 * dc_oracle.c -- CPU oracle for the DynColor recolor path (TEST INFRASTRUCTURE
 * ONLY: used by tests/ as the checker and by bench.py's CPU baseline; the
 * product library never links or calls it).
 *
 * Plain C, single thread.  Restates the reference's rules under round-
 * synchronous snapshot semantics; every function cites the reference lines it
 * follows (paths relative to the reference repo root):
 *   HI = DynColorHeuristicIndependent/DynColorHeuristicIndependent/
 *   HB = DynColorHeuristicBased/DynColorHeuristicBased/
 *   MG = Multi-GPU/
 */
#include "dc_oracle.h"

#include <stdlib.h>
#include <string.h>

typedef struct {
  int32_t a, b;
  uint8_t eq0; /* batch-start colors equal (insertions only) */
} orc_edge;

struct orc_graph {
  int32_t n, n_owned;
  /* base CSR built from the creation edge list (both directions, multi-edges
   * kept, -1 = tombstone), plus per-vertex appended insertion entries */
  int64_t* off;
  int32_t* adj;
  int32_t** ext;
  int32_t* ext_len;
  int32_t* ext_cap;
  int32_t* prop; /* ldf priority: running +-1 line counter */

  /* staged batch */
  orc_edge* del;  int64_t ndel;
  orc_edge* ins;  int64_t nins;
  orc_edge* xdel; int64_t nxdel;
  orc_edge* xins; int64_t nxins;
  int32_t staged;

  /* scratch */
  int32_t* vstamp; int32_t vgen;   /* vertex set membership */
  int32_t* cstamp; int32_t cgen;   /* color marks for mex (size n+2) */
  int32_t* list_a; int32_t* list_b;
  int32_t* newc;
  int32_t* nbuf; int64_t nbuf_cap;
  int32_t* pstamp; int32_t pgen;   /* parallel replay: atomic set membership */
};

static int is_frozen(const orc_graph* g, int32_t v) { return v >= g->n_owned; }

/* ---- adjacency iteration ------------------------------------------------ */
/* HI/supportingFunctions.cu:29-33 flattens per-vertex lists (base edges first,
 * then insertions appended by readin_changes, HI/all_structure_undir.cuh:96-105)
 * into one row; a row entry is live when its flag is not -1
 * (HI/gpuFunctions_undir.cuh:61). Here the row is base part + ext part. */
#define FOR_LIVE(g, v, w, BODY)                                              \
  do {                                                                       \
    for (int64_t _j = (g)->off[(v)]; _j < (g)->off[(v) + 1]; ++_j) {         \
      int32_t w = (g)->adj[_j];                                              \
      if (w < 0) continue;                                                   \
      BODY                                                                   \
    }                                                                        \
    for (int32_t _j = 0; _j < (g)->ext_len[(v)]; ++_j) {                     \
      int32_t w = (g)->ext[(v)][_j];                                         \
      if (w < 0) continue;                                                   \
      BODY                                                                   \
    }                                                                        \
  } while (0)

int64_t orc_live_degree(const orc_graph* g, int32_t v) {
  int64_t d = 0;
  if (v < 0 || v >= g->n) return -1;
  FOR_LIVE(g, v, w, { (void)w; ++d; });
  return d;
}

int64_t orc_neighbors(const orc_graph* g, int32_t v, int32_t* out, int64_t cap) {
  int64_t d = 0;
  if (v < 0 || v >= g->n) return -1;
  FOR_LIVE(g, v, w, { if (d < cap) out[d] = w; ++d; });
  return d;
}

void orc_get_prop(const orc_graph* g, int32_t* prop) {
  memcpy(prop, g->prop, sizeof(int32_t) * (size_t)g->n);
}

/* ---- mex / distinct count ----------------------------------------------- */
/* FirstZeroBit (HI/gpuFunctions_undir.cuh:36-52, HB/gpuFunctions_undir.cuh:36-40)
 * is the least non-negative integer absent from the neighbour color set.  The
 * reference encodes it in 32-bit words via exp2(double) which overflows for
 * colors >= 31; the restatement is the exact, unbounded mex. */
int32_t orc_mex(const int32_t* c, int64_t k) {
  /* standalone version (tests): O(k) with a local mark array */
  uint8_t* seen = (uint8_t*)calloc((size_t)k + 1, 1);
  int32_t r = 0;
  if (!seen) return -1;
  for (int64_t i = 0; i < k; ++i)
    if (c[i] >= 0 && c[i] <= k) seen[c[i]] = 1;
  while (r <= k && seen[r]) ++r;
  free(seen);
  return r;
}

static void next_cgen(orc_graph* g) {
  if (++g->cgen == 0x7fffffff) {
    memset(g->cstamp, 0, sizeof(int32_t) * ((size_t)g->n + 2));
    g->cgen = 1;
  }
}
static void next_vgen(orc_graph* g) {
  if (++g->vgen == 0x7fffffff) {
    memset(g->vstamp, 0, sizeof(int32_t) * (size_t)g->n);
    g->vgen = 1;
  }
}

/* ldf priority: u precedes v  (HB/gpuFunctions_undir.cuh:104-113) */
static int prior(const orc_graph* g, int32_t u, int32_t v) {
  if (g->prop[u] != g->prop[v]) return g->prop[u] > g->prop[v];
  return u < v;
}

/* mex over colors of live neighbours of t (all of them: computeSC,
 * HI/gpuFunctions_undir.cuh:57-69) or only of neighbours preceding t
 * (findMinAvailableColor, HB/gpuFunctions_undir.cuh:118-129).
 * Only colors <= (number of counted neighbours) can matter. */
static int32_t mex_of(orc_graph* g, const int32_t* color, int32_t t, int ldf_only_prior) {
  int64_t cnt = 0;
  next_cgen(g);
  FOR_LIVE(g, t, w, {
    if (ldf_only_prior && !prior(g, w, t)) continue;
    int32_t c = color[w];
    if (c >= 0 && c <= g->n) g->cstamp[c] = g->cgen;
    ++cnt;
  });
  int32_t r = 0;
  while (r <= g->n && g->cstamp[r] == g->cgen) ++r;
  return r;
}

static int cmp_i32(const void* x, const void* y) {
  int32_t a = *(const int32_t*)x, b = *(const int32_t*)y;
  return (a > b) - (a < b);
}

/* SC size = number of distinct colors among live neighbours (BitCount summed
 * over mask words, HI/gpuFunctions_undir.cuh:203-207). */
static int64_t distinct_colors(orc_graph* g, const int32_t* color, int32_t t) {
  int64_t d = 0;
  FOR_LIVE(g, t, w, { (void)w; ++d; });
  if (d > g->nbuf_cap) {
    free(g->nbuf);
    g->nbuf_cap = d * 2;
    g->nbuf = (int32_t*)malloc(sizeof(int32_t) * (size_t)g->nbuf_cap);
  }
  int64_t k = 0;
  FOR_LIVE(g, t, w, { g->nbuf[k++] = color[w]; });
  if (k > 1) qsort(g->nbuf, (size_t)k, sizeof(int32_t), cmp_i32);
  int64_t distinct = 0;
  for (int64_t i = 0; i < k; ++i)
    if (i == 0 || g->nbuf[i] != g->nbuf[i - 1]) ++distinct;
  return distinct;
}

/* ---- construction --------------------------------------------------------- */
/* read_graphEdges (HI/all_structure_undir.cuh:156-181; HB adds the +1 degree
 * counter per endpoint, HB/all_structure_undir.cuh:177-178).  Partition share:
 * an edge with one ghost endpoint is stored in the owned row only
 * (MG/all_structure_undir.cuh:250-308, unified local IDs). */
orc_graph* orc_graph_create(int32_t n, int32_t n_owned, int64_t m,
                            const int32_t* u, const int32_t* v) {
  if (n < 0 || n_owned < 0 || n_owned > n || m < 0) return NULL;
  orc_graph* g = (orc_graph*)calloc(1, sizeof(orc_graph));
  if (!g) return NULL;
  g->n = n;
  g->n_owned = n_owned;
  g->off = (int64_t*)calloc((size_t)n + 1, sizeof(int64_t));
  g->ext = (int32_t**)calloc((size_t)n + 1, sizeof(int32_t*));
  g->ext_len = (int32_t*)calloc((size_t)n + 1, sizeof(int32_t));
  g->ext_cap = (int32_t*)calloc((size_t)n + 1, sizeof(int32_t));
  g->prop = (int32_t*)calloc((size_t)n + 1, sizeof(int32_t));
  g->vstamp = (int32_t*)calloc((size_t)n + 1, sizeof(int32_t));
  g->cstamp = (int32_t*)calloc((size_t)n + 2, sizeof(int32_t));
  g->list_a = (int32_t*)malloc(sizeof(int32_t) * ((size_t)n + 1));
  g->list_b = (int32_t*)malloc(sizeof(int32_t) * ((size_t)n + 1));
  g->newc = (int32_t*)malloc(sizeof(int32_t) * ((size_t)n + 1));
  if (!g->off || !g->ext || !g->ext_len || !g->ext_cap || !g->prop || !g->vstamp ||
      !g->cstamp || !g->list_a || !g->list_b || !g->newc) {
    orc_graph_destroy(g);
    return NULL;
  }
  for (int64_t i = 0; i < m; ++i) {
    int32_t a = u[i], b = v[i];
    if (a < 0 || b < 0 || a >= n || b >= n || a == b ||
        (is_frozen(g, a) && is_frozen(g, b))) {
      orc_graph_destroy(g);
      return NULL;
    }
    if (!is_frozen(g, a)) g->off[a + 1]++;
    if (!is_frozen(g, b)) g->off[b + 1]++;
    g->prop[a]++;
    g->prop[b]++;
  }
  for (int32_t i = 0; i < n; ++i) g->off[i + 1] += g->off[i];
  g->adj = (int32_t*)malloc(sizeof(int32_t) * (size_t)(g->off[n] + 1));
  int64_t* pos = (int64_t*)malloc(sizeof(int64_t) * ((size_t)n + 1));
  if (!g->adj || !pos) {
    free(pos);
    orc_graph_destroy(g);
    return NULL;
  }
  memcpy(pos, g->off, sizeof(int64_t) * (size_t)n);
  for (int64_t i = 0; i < m; ++i) {
    int32_t a = u[i], b = v[i];
    if (!is_frozen(g, a)) g->adj[pos[a]++] = b;
    if (!is_frozen(g, b)) g->adj[pos[b]++] = a;
  }
  free(pos);
  return g;
}

void orc_graph_destroy(orc_graph* g) {
  if (!g) return;
  if (g->ext)
    for (int32_t i = 0; i < g->n; ++i) free(g->ext[i]);
  free(g->ext); free(g->ext_len); free(g->ext_cap);
  free(g->off); free(g->adj); free(g->prop);
  free(g->del); free(g->ins); free(g->xdel); free(g->xins);
  free(g->vstamp); free(g->cstamp); free(g->list_a); free(g->list_b);
  free(g->newc); free(g->nbuf); free(g->pstamp);
  free(g);
}

static int ext_push(orc_graph* g, int32_t a, int32_t b) {
  if (g->ext_len[a] == g->ext_cap[a]) {
    int32_t nc = g->ext_cap[a] ? g->ext_cap[a] * 2 : 4;
    int32_t* p = (int32_t*)realloc(g->ext[a], sizeof(int32_t) * (size_t)nc);
    if (!p) return ORC_ENOMEM;
    g->ext[a] = p;
    g->ext_cap[a] = nc;
  }
  g->ext[a][g->ext_len[a]++] = b;
  return ORC_OK;
}

/* deleteEdgeFromAdj: tombstone EVERY entry equal to the pair in both rows
 * (HI/gpuFunctions_undir.cuh:118-142), after all insertions were appended, so a
 * deletion also removes re-inserted copies. */
static void tombstone_row(orc_graph* g, int32_t a, int32_t b) {
  if (is_frozen(g, a)) return;
  for (int64_t j = g->off[a]; j < g->off[a + 1]; ++j)
    if (g->adj[j] == b) g->adj[j] = -1;
  for (int32_t j = 0; j < g->ext_len[a]; ++j)
    if (g->ext[a][j] == b) g->ext[a][j] = -1;
}

/* readin_changes (HI/all_structure_undir.cuh:78-123, HB/all_structure_undir.cuh:
 * 78-129, MG/all_structure_undir.cuh:78-192) + the tombstone launch
 * (HI/supportingFunctions.cu:93).  Insertions are appended to both rows, the
 * ldf counter moves by +-1 per line, deletions are collected and tombstoned.
 * eq0 records the batch-start color equality used by the single-graph
 * insertion filter (HI/all_structure_undir.cuh:106-108).  In partition mode a
 * cross edge is normalised to (owned, ghost) like MG/all_structure_undir.cuh:
 * 130-131,161-162. */
int orc_apply_batch(orc_graph* g, const int32_t* color, int64_t k,
                    const int32_t* u, const int32_t* v, const uint8_t* ins) {
  for (int64_t i = 0; i < k; ++i) {
    int32_t a = u[i], b = v[i];
    if (a < 0 || b < 0 || a >= g->n || b >= g->n || a == b ||
        (is_frozen(g, a) && is_frozen(g, b)))
      return ORC_EINVAL;
  }
  free(g->del); free(g->ins); free(g->xdel); free(g->xins);
  g->del = (orc_edge*)malloc(sizeof(orc_edge) * ((size_t)k + 1));
  g->ins = (orc_edge*)malloc(sizeof(orc_edge) * ((size_t)k + 1));
  g->xdel = (orc_edge*)malloc(sizeof(orc_edge) * ((size_t)k + 1));
  g->xins = (orc_edge*)malloc(sizeof(orc_edge) * ((size_t)k + 1));
  if (!g->del || !g->ins || !g->xdel || !g->xins) return ORC_ENOMEM;
  g->ndel = g->nins = g->nxdel = g->nxins = 0;
  for (int64_t i = 0; i < k; ++i) {
    int32_t a = u[i], b = v[i];
    int cross = is_frozen(g, a) || is_frozen(g, b);
    if (cross && is_frozen(g, a)) { int32_t t = a; a = b; b = t; }
    orc_edge e = {a, b, (uint8_t)(color[a] == color[b])};
    if (ins[i]) {
      if (!is_frozen(g, a) && ext_push(g, a, b)) return ORC_ENOMEM;
      if (!is_frozen(g, b) && ext_push(g, b, a)) return ORC_ENOMEM;
      g->prop[a]++;
      g->prop[b]++;
      if (cross) g->xins[g->nxins++] = e; else g->ins[g->nins++] = e;
    } else {
      g->prop[a]--;
      g->prop[b]--;
      if (cross) g->xdel[g->nxdel++] = e; else g->del[g->ndel++] = e;
    }
  }
  for (int64_t i = 0; i < g->ndel; ++i) {
    tombstone_row(g, g->del[i].a, g->del[i].b);
    tombstone_row(g, g->del[i].b, g->del[i].a);
  }
  for (int64_t i = 0; i < g->nxdel; ++i) tombstone_row(g, g->xdel[i].a, g->xdel[i].b);
  g->staged = 1;
  return ORC_OK;
}

/* ---- trace helpers -------------------------------------------------------- */
typedef struct {
  int64_t* t;
  int32_t cap, n;
} trace_t;

static void trace_add(trace_t* tr, int64_t kind, int64_t a, int64_t b, int64_t c) {
  if (tr->t && tr->n < tr->cap) {
    tr->t[4 * tr->n + 0] = kind;
    tr->t[4 * tr->n + 1] = a;
    tr->t[4 * tr->n + 2] = b;
    tr->t[4 * tr->n + 3] = c;
  }
  tr->n++;
}

/* add v to the current vertex set (list, count) if not yet there */
static void set_add(orc_graph* g, int32_t* list, int64_t* cnt, int32_t v) {
  if (g->vstamp[v] != g->vgen) {
    g->vstamp[v] = g->vgen;
    list[(*cnt)++] = v;
  }
}

/* Deletion round (deleteEdge, HI/gpuFunctions_undir.cuh:147-177; cross
 * variant deleteEdge_cross, MG/gpuFunctions_undir.cuh:175-200): target =
 * endpoint with the larger color (tie -> node2); recolor to the mex of all
 * live neighbours when it differs.  A frozen target is skipped. */
static int64_t round_del(orc_graph* g, const orc_edge* e, int64_t ne, int32_t* color,
                         int32_t* prev, int32_t* marked, int64_t* nmarked,
                         int64_t* ntargets) {
  int64_t nt = 0, changed = 0;
  next_vgen(g);
  for (int64_t i = 0; i < ne; ++i) {
    int32_t t = color[e[i].a] > color[e[i].b] ? e[i].a : e[i].b;
    if (is_frozen(g, t)) continue;
    set_add(g, g->list_a, &nt, t);
  }
  for (int64_t i = 0; i < nt; ++i) g->newc[i] = mex_of(g, color, g->list_a[i], 0);
  for (int64_t i = 0; i < nt; ++i) {
    int32_t t = g->list_a[i];
    if (g->newc[i] != color[t]) {
      prev[t] = color[t];
      color[t] = g->newc[i];
      marked[(*nmarked)++] = t;
      ++changed;
    }
  }
  *ntargets = nt;
  return changed;
}

/* Insertion round (insEdge, HI/gpuFunctions_undir.cuh:182-223; cross variant
 * insEdge_cross, MG/gpuFunctions_undir.cuh:256-283 with the peer's SC size
 * unavailable, i.e. +inf): for an edge whose endpoints share a color (checked
 * on the round snapshot), the endpoint with the smaller SC size (distinct
 * neighbour colors; tie -> node2) is recolored unconditionally. */
static int64_t round_ins(orc_graph* g, const orc_edge* e, int64_t ne, int use_filter,
                         int32_t* color, int32_t* prev, int32_t* marked,
                         int64_t* nmarked, int64_t* ntargets) {
  int64_t nt = 0, changed = 0;
  next_vgen(g);
  for (int64_t i = 0; i < ne; ++i) {
    if (use_filter && !e[i].eq0) continue;
    int32_t a = e[i].a, b = e[i].b;
    if (color[a] != color[b]) continue;
    int64_t ka = is_frozen(g, a) ? INT64_MAX : distinct_colors(g, color, a);
    int64_t kb = is_frozen(g, b) ? INT64_MAX : distinct_colors(g, color, b);
    int32_t t = ka < kb ? a : b;
    if (is_frozen(g, t)) continue;
    set_add(g, g->list_a, &nt, t);
  }
  for (int64_t i = 0; i < nt; ++i) g->newc[i] = mex_of(g, color, g->list_a[i], 0);
  for (int64_t i = 0; i < nt; ++i) {
    int32_t t = g->list_a[i];
    if (g->newc[i] != color[t]) ++changed;
    prev[t] = color[t];
    color[t] = g->newc[i];
    marked[(*nmarked)++] = t;
  }
  *ntargets = nt;
  return changed;
}

static int count_filtered(const orc_edge* e, int64_t ne, int use_filter) {
  if (!use_filter) return ne > 0;
  for (int64_t i = 0; i < ne; ++i)
    if (e[i].eq0) return 1;
  return 0;
}

/* ---- recolor ------------------------------------------------------------- */
int orc_recolor(orc_graph* g, int rule, int32_t* color, int32_t* prev,
                int32_t max_loop, int64_t* trace, int32_t trace_cap,
                int32_t* n_phases) {
  trace_t tr = {trace, trace_cap, 0};
  int32_t* L = g->list_b;
  int64_t nL = 0;
  int change = 0;
  int partition = g->n_owned < g->n;
  if (!g->staged) return ORC_EINVAL;
  if (rule != ORC_RULE_INDEPENDENT && rule != ORC_RULE_LDF) return ORC_EINVAL;
  if (partition && rule != ORC_RULE_INDEPENDENT) return ORC_EINVAL;
  g->staged = 0;

  /* the marked set must be duplicate-free: a vertex can be marked by more
   * than one pre-loop round */
  int32_t* marked = (int32_t*)malloc(sizeof(int32_t) * ((size_t)g->n * 4 + 4));
  int64_t nmarked = 0;
  if (!marked) return ORC_ENOMEM;

  if (rule == ORC_RULE_INDEPENDENT) {
    /* prev reset per run (HI/supportingFunctions.cu:86-87) */
    for (int32_t i = 0; i < g->n; ++i) prev[i] = -1;
    int64_t nt, ch;
    /* HI/Color_main.cu:149-160; MG/kernel.cu:299-312 */
    if (g->ndel > 0) {
      ch = round_del(g, g->del, g->ndel, color, prev, marked, &nmarked, &nt);
      trace_add(&tr, ORC_PH_DEL, nt, ch, 0);
      if (ch) change = 1;
    }
    /* MG/kernel.cu:314-326 */
    if (g->nxdel > 0) {
      ch = round_del(g, g->xdel, g->nxdel, color, prev, marked, &nmarked, &nt);
      trace_add(&tr, ORC_PH_XDEL, nt, ch, 0);
      if (ch) change = 1;
    }
    /* HI/Color_main.cu:163-174 (insertion list pre-filtered by the batch-start
     * colors, HI/all_structure_undir.cuh:106-108); the partition share does not
     * pre-filter (MG/all_structure_undir.cuh:111-120) */
    if (count_filtered(g->ins, g->nins, !partition)) {
      ch = round_ins(g, g->ins, g->nins, !partition, color, prev, marked, &nmarked, &nt);
      trace_add(&tr, ORC_PH_INS, nt, ch, 0);
      if (nt) change = 1;
    }
    /* MG/kernel.cu:475-487 */
    if (g->nxins > 0) {
      ch = round_ins(g, g->xins, g->nxins, 0, color, prev, marked, &nmarked, &nt);
      trace_add(&tr, ORC_PH_XINS, nt, ch, 0);
      if (nt) change = 1;
    }
  } else {
    /* processChangeEdge over deletions + color-equal insertions
     * (HB/gpuFunctions_undir.cuh:222-249, list built at
     * HB/all_structure_undir.cuh:106-109,117): both endpoints are re-evaluated
     * with the predecessor mex; a changed endpoint is written and marked. */
    int64_t nt = 0, ch = 0;
    int any = g->ndel > 0 || count_filtered(g->ins, g->nins, 1);
    next_vgen(g);
    for (int64_t i = 0; i < g->ndel; ++i) {
      set_add(g, g->list_a, &nt, g->del[i].a);
      set_add(g, g->list_a, &nt, g->del[i].b);
    }
    for (int64_t i = 0; i < g->nins; ++i) {
      if (!g->ins[i].eq0) continue;
      set_add(g, g->list_a, &nt, g->ins[i].a);
      set_add(g, g->list_a, &nt, g->ins[i].b);
    }
    for (int64_t i = 0; i < nt; ++i) g->newc[i] = mex_of(g, color, g->list_a[i], 1);
    for (int64_t i = 0; i < nt; ++i) {
      int32_t t = g->list_a[i];
      if (g->newc[i] != color[t]) {
        color[t] = g->newc[i];
        marked[nmarked++] = t;
        ++ch;
      }
    }
    if (any) trace_add(&tr, ORC_PH_CHANGE, nt, ch, 0);
    if (ch) change = 1;
  }

  /* initial frontier = compaction of the mark flags (HI/Color_main.cu:179):
   * a set */
  next_vgen(g);
  for (int64_t i = 0; i < nmarked; ++i) set_add(g, L, &nL, marked[i]);
  free(marked);

  int32_t iter = 0;
  int32_t* C = g->list_a;
  /* HI/Color_main.cu:181-202, HB/Color_main.cu:203-224: the next frontier is
   * the full candidate list of the previous round */
  while (change) {
    int64_t nC = 0, changed = 0;
    if (iter >= max_loop) {
      if (n_phases) *n_phases = tr.n;
      return ORC_ECAP;
    }
    change = 0;
    next_vgen(g);
    for (int64_t i = 0; i < nL; ++i) {
      int32_t v = L[i];
      if (rule == ORC_RULE_INDEPENDENT) {
        /* findEligibleNeighbors, HI/gpuFunctions_undir.cuh:229-260 */
        FOR_LIVE(g, v, w, {
          if (is_frozen(g, w)) continue;
          if ((color[v] == color[w] && v < w) || prev[v] < color[w])
            set_add(g, C, &nC, w);
        });
      } else {
        /* findEligibleNeighbors, HB/gpuFunctions_undir.cuh:259-291 */
        FOR_LIVE(g, v, w, {
          int pv = prior(g, v, w);
          if (color[v] == color[w]) set_add(g, C, &nC, pv ? w : v);
          if (pv) set_add(g, C, &nC, w);
        });
      }
    }
    /* recolorNeighbor on the snapshot (HI/gpuFunctions_undir.cuh:264-284:
     * unconditional; HB/gpuFunctions_undir.cuh:295-314: only if changed) */
    for (int64_t i = 0; i < nC; ++i)
      g->newc[i] = mex_of(g, color, C[i], rule == ORC_RULE_LDF);
    for (int64_t i = 0; i < nC; ++i) {
      int32_t w = C[i];
      if (g->newc[i] != color[w]) ++changed;
      if (rule == ORC_RULE_INDEPENDENT) {
        prev[w] = color[w];
        color[w] = g->newc[i];
      } else if (g->newc[i] != color[w]) {
        color[w] = g->newc[i];
      }
    }
    trace_add(&tr, ORC_PH_LOOP, nL, nC, changed);
    if (rule == ORC_RULE_INDEPENDENT) change = nC > 0;
    else change = changed > 0;
    /* L <- C */
    int32_t* t = L; L = C; C = t;
    nL = nC;
    ++iter;
  }
  if (n_phases) *n_phases = tr.n;
  return ORC_OK;
}

/* ---- static greedy --------------------------------------------------------- */
typedef struct { int32_t prop, id; } pk_t;
static int cmp_pk(const void* x, const void* y) {
  const pk_t* a = (const pk_t*)x;
  const pk_t* b = (const pk_t*)y;
  if (a->prop != b->prop) return a->prop > b->prop ? -1 : 1;
  return (a->id > b->id) - (a->id < b->id);
}

/* First-fit greedy: each owned vertex, in order, takes the least color unused
 * by its already-colored neighbours.  The reference takes its initial colors
 * from an external tool (Gunrock, MG/README.txt:21-23); this is the
 * deterministic textbook replacement.  ID order for the independent rule, the
 * ldf priority order (HB/gpuFunctions_undir.cuh:104-113) for ldf. */
void orc_static_greedy(orc_graph* g, int rule, int32_t* color) {
  int32_t n = g->n_owned;
  int32_t* order = (int32_t*)malloc(sizeof(int32_t) * ((size_t)n + 1));
  uint8_t* done = (uint8_t*)calloc((size_t)g->n + 1, 1);
  for (int32_t i = n; i < g->n; ++i) done[i] = 1; /* ghosts keep their colors */
  if (rule == ORC_RULE_LDF) {
    pk_t* pk = (pk_t*)malloc(sizeof(pk_t) * ((size_t)n + 1));
    for (int32_t i = 0; i < n; ++i) { pk[i].prop = g->prop[i]; pk[i].id = i; }
    qsort(pk, (size_t)n, sizeof(pk_t), cmp_pk);
    for (int32_t i = 0; i < n; ++i) order[i] = pk[i].id;
    free(pk);
  } else {
    for (int32_t i = 0; i < n; ++i) order[i] = i;
  }
  for (int32_t i = 0; i < n; ++i) {
    int32_t v = order[i];
    next_cgen(g);
    FOR_LIVE(g, v, w, {
      if (done[w] && color[w] >= 0 && color[w] <= g->n) g->cstamp[color[w]] = g->cgen;
    });
    int32_t r = 0;
    while (g->cstamp[r] == g->cgen) ++r;
    color[v] = r;
    done[v] = 1;
  }
  free(order);
  free(done);
}

/* validate (HI/gpuFunctions_undir.cuh:289-305) as a count instead of printf:
 * monochromatic live entries (v,w) with v < w (or w a ghost); max color over
 * owned vertices (HI/Color_main.cu:218-226). */
int64_t orc_validate(const orc_graph* g, const int32_t* color, int32_t* max_color) {
  int64_t bad = 0;
  int32_t mc = -1;
  for (int32_t v = 0; v < g->n_owned; ++v) {
    if (color[v] > mc) mc = color[v];
    FOR_LIVE(g, v, w, {
      if ((w > v || is_frozen(g, w)) && color[w] == color[v]) ++bad;
    });
  }
  if (max_color) *max_color = mc;
  return bad;
}

/* ---- parallel replay (OpenMP) ----------------------------------------------
 * The same independent-rule rounds as orc_recolor, each round's work split
 * over threads: a round reads only the snapshot and writes only its own
 * result lists, so the results, the trace counts and the colours equal the
 * single-thread replay's (sets are order-free).  Timed as the all-cores CPU
 * baseline; tests compare it with orc_recolor. */
#ifdef _OPENMP
#include <omp.h>
#else
static int omp_get_thread_num(void) { return 0; }
#endif

typedef struct {
  int32_t* stamp; int32_t gen; int64_t cap;   /* mex marks, indexed by colour <= row length */
  int32_t* buf; int64_t bcap;                 /* distinct-colour sort buffer */
  int32_t* out; int64_t nout, ocap;           /* this thread's appended vertices */
} par_scratch;

static int64_t row_len(const orc_graph* g, int32_t v) {
  return (g->off[v + 1] - g->off[v]) + g->ext_len[v];
}

static int out_push(par_scratch* s, int32_t v) {
  if (s->nout == s->ocap) {
    int64_t nc = s->ocap ? 2 * s->ocap : 1024;
    int32_t* p = (int32_t*)realloc(s->out, sizeof(int32_t) * (size_t)nc);
    if (!p) return ORC_ENOMEM;
    s->out = p;
    s->ocap = nc;
  }
  s->out[s->nout++] = v;
  return ORC_OK;
}

/* true for exactly one caller per generation */
static int par_claim(orc_graph* g, int32_t v) {
  return __atomic_exchange_n(&g->pstamp[v], g->pgen, __ATOMIC_RELAXED) != g->pgen;
}

static void par_next_gen(orc_graph* g) {
  if (++g->pgen == 0x7fffffff) {
    memset(g->pstamp, 0, sizeof(int32_t) * (size_t)g->n);
    g->pgen = 1;
  }
}

static int32_t par_mex(const orc_graph* g, const int32_t* color, int32_t t, par_scratch* s) {
  if (++s->gen == 0x7fffffff) {
    memset(s->stamp, 0, sizeof(int32_t) * (size_t)s->cap);
    s->gen = 1;
  }
  const int64_t lim = s->cap - 1;   /* the mex never exceeds the row length */
  FOR_LIVE(g, t, w, {
    int32_t c = color[w];
    if (c >= 0 && c <= lim) s->stamp[c] = s->gen;
  });
  int32_t r = 0;
  while (r <= lim && s->stamp[r] == s->gen) ++r;
  return r;
}

static int64_t par_distinct(const orc_graph* g, const int32_t* color, int32_t t, par_scratch* s) {
  int64_t k = 0;
  FOR_LIVE(g, t, w, { s->buf[k++] = color[w]; });
  if (k > 1) qsort(s->buf, (size_t)k, sizeof(int32_t), cmp_i32);
  int64_t distinct = 0;
  for (int64_t i = 0; i < k; ++i)
    if (i == 0 || s->buf[i] != s->buf[i - 1]) ++distinct;
  return distinct;
}

/* gather every thread's out list into list (count into *cnt), clearing them */
static void par_gather(par_scratch* sc, int nt, int32_t* list, int64_t* cnt) {
  int64_t n = 0;
  for (int i = 0; i < nt; ++i) {
    memcpy(list + n, sc[i].out, sizeof(int32_t) * (size_t)sc[i].nout);
    n += sc[i].nout;
    sc[i].nout = 0;
  }
  *cnt = n;
}

int orc_recolor_par(orc_graph* g, int rule, int32_t* color, int32_t* prev, int32_t max_loop,
                    int64_t* trace, int32_t trace_cap, int32_t* n_phases, int32_t threads) {
  trace_t tr = {trace, trace_cap, 0};
  const int partition = g->n_owned < g->n;
  if (!g->staged || rule != ORC_RULE_INDEPENDENT || threads < 1) return ORC_EINVAL;
  g->staged = 0;
  if (!g->pstamp) {
    g->pstamp = (int32_t*)calloc((size_t)g->n + 1, sizeof(int32_t));
    if (!g->pstamp) return ORC_ENOMEM;
    g->pgen = 0;
  }
  int64_t maxlen = 1;
  for (int32_t v = 0; v < g->n; ++v) {
    const int64_t l = row_len(g, v);
    if (l > maxlen) maxlen = l;
  }
  const int nt = threads;
  par_scratch* sc = (par_scratch*)calloc((size_t)nt, sizeof(par_scratch));
  int32_t* marked = (int32_t*)malloc(sizeof(int32_t) * ((size_t)g->n + 1));
  int32_t* newc = g->newc;
  int32_t* L = g->list_b;
  int32_t* C = g->list_a;
  int rc = sc && marked ? ORC_OK : ORC_ENOMEM;
  for (int i = 0; rc == ORC_OK && i < nt; ++i) {
    sc[i].cap = maxlen + 2;
    sc[i].stamp = (int32_t*)calloc((size_t)sc[i].cap, sizeof(int32_t));
    sc[i].bcap = maxlen + 1;
    sc[i].buf = (int32_t*)malloc(sizeof(int32_t) * (size_t)sc[i].bcap);
    if (!sc[i].stamp || !sc[i].buf) rc = ORC_ENOMEM;
  }
  int64_t nmarked = 0, nL = 0;
  int change = 0;
  if (rc) goto done;

  for (int32_t i = 0; i < g->n; ++i) prev[i] = -1;
  /* pre-loop rounds: targets of the lines (a set), their mex on the snapshot,
   * then the commit (deletion rounds: only changed targets are written and
   * marked; insertion rounds: every target) */
  for (int kind = 0; kind < 4; ++kind) {
    const orc_edge* e = kind == 0 ? g->del : kind == 1 ? g->xdel : kind == 2 ? g->ins : g->xins;
    const int64_t ne = kind == 0 ? g->ndel : kind == 1 ? g->nxdel : kind == 2 ? g->nins : g->nxins;
    const int ins = kind >= 2;
    const int use_filter = kind == 2 && !partition;
    if (ins ? (kind == 2 ? !count_filtered(e, ne, use_filter) : ne == 0) : ne == 0) continue;
    par_next_gen(g);
#pragma omp parallel for num_threads(nt) schedule(dynamic, 256)
    for (int64_t i = 0; i < ne; ++i) {
      par_scratch* s = &sc[omp_get_thread_num()];
      int32_t a = e[i].a, b = e[i].b, t;
      if (!ins) {
        t = color[a] > color[b] ? a : b;
      } else {
        if (use_filter && !e[i].eq0) continue;
        if (color[a] != color[b]) continue;
        int64_t ka = is_frozen(g, a) ? INT64_MAX : par_distinct(g, color, a, s);
        int64_t kb = is_frozen(g, b) ? INT64_MAX : par_distinct(g, color, b, s);
        t = ka < kb ? a : b;
      }
      if (is_frozen(g, t) || !par_claim(g, t)) continue;
      if (out_push(s, t)) continue;
    }
    int64_t nt_ = 0;
    par_gather(sc, nt, C, &nt_);
    int64_t changed = 0;
#pragma omp parallel for num_threads(nt) schedule(dynamic, 64)
    for (int64_t i = 0; i < nt_; ++i) newc[i] = par_mex(g, color, C[i], &sc[omp_get_thread_num()]);
    for (int64_t i = 0; i < nt_; ++i) {
      const int32_t t = C[i];
      const int ch = newc[i] != color[t];
      changed += ch;
      if (ins || ch) {
        prev[t] = color[t];
        color[t] = newc[i];
        marked[nmarked++] = t;
      }
    }
    trace_add(&tr, kind == 0 ? ORC_PH_DEL : kind == 1 ? ORC_PH_XDEL : kind == 2 ? ORC_PH_INS : ORC_PH_XINS,
              nt_, changed, 0);
    if (ins ? nt_ > 0 : changed > 0) change = 1;
  }
  /* the marked set (a vertex can be marked by more than one round) */
  par_next_gen(g);
  for (int64_t i = 0; i < nmarked; ++i)
    if (par_claim(g, marked[i])) L[nL++] = marked[i];

  int32_t iter = 0;
  while (change) {
    if (iter >= max_loop) {
      rc = ORC_ECAP;
      goto done;
    }
    par_next_gen(g);
#pragma omp parallel for num_threads(nt) schedule(dynamic, 64)
    for (int64_t i = 0; i < nL; ++i) {
      par_scratch* s = &sc[omp_get_thread_num()];
      const int32_t v = L[i];
      FOR_LIVE(g, v, w, {
        if (is_frozen(g, w)) continue;
        if (((color[v] == color[w] && v < w) || prev[v] < color[w]) && par_claim(g, w))
          (void)out_push(s, w);
      });
    }
    int64_t nC = 0, changed = 0;
    par_gather(sc, nt, C, &nC);
#pragma omp parallel for num_threads(nt) schedule(dynamic, 64)
    for (int64_t i = 0; i < nC; ++i) newc[i] = par_mex(g, color, C[i], &sc[omp_get_thread_num()]);
#pragma omp parallel for num_threads(nt) reduction(+ : changed)
    for (int64_t i = 0; i < nC; ++i) {
      const int32_t w = C[i];
      changed += newc[i] != color[w];
      prev[w] = color[w];
      color[w] = newc[i];
    }
    trace_add(&tr, ORC_PH_LOOP, nL, nC, changed);
    change = nC > 0;
    int32_t* t = L; L = C; C = t;
    nL = nC;
    ++iter;
  }
done:
  if (n_phases) *n_phases = tr.n;
  for (int i = 0; sc && i < nt; ++i) {
    free(sc[i].stamp);
    free(sc[i].buf);
    free(sc[i].out);
  }
  free(sc);
  free(marked);
  return rc;
}
