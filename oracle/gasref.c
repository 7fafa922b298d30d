/*
 * gasref.c -- CPU oracle for the GAS hot path (see gasref.h).
 * TEST INFRASTRUCTURE ONLY: never linked into or called by libgas.so.
 *
 * The engine below restates GASEngineRef (refgas.h:30-225): the program is
 * a table of callbacks on opaque records so one engine serves all four
 * programs, as the template does.
 */
#include "gasref.h"

#include <limits.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ---- CSR / CSC: util.h:79-121 ---------------------------------------- */
/* Stable counting sort of the edges by key; one valid ordering of the
 * reference's unstable std::sort on (key, index) pairs (util.h:48-72). */
static void compress(int nV, int nE, const int32_t *keys, const int32_t *vals,
                     int32_t *offsets, int32_t *out_vals, int32_t *eid) {
  int32_t *cur = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nV + 1));
  memset(offsets, 0, sizeof(int32_t) * (size_t)(nV + 1));
  for (int k = 0; k < nE; ++k) offsets[keys[k] + 1]++;
  for (int v = 0; v < nV; ++v) offsets[v + 1] += offsets[v];
  memcpy(cur, offsets, sizeof(int32_t) * (size_t)nV);
  for (int k = 0; k < nE; ++k) {
    int32_t p = cur[keys[k]]++;
    if (out_vals) out_vals[p] = vals[k];
    if (eid) eid[p] = k;
  }
  free(cur);
}

void gasref_edge_list_to_csr(int nV, int nE, const int32_t *srcs,
                             const int32_t *dsts, int32_t *offsets,
                             int32_t *out_dsts, int32_t *eid) {
  compress(nV, nE, srcs, dsts, offsets, out_dsts, eid); /* util.h:79-110 */
}

void gasref_edge_list_to_csc(int nV, int nE, const int32_t *srcs,
                             const int32_t *dsts, int32_t *offsets,
                             int32_t *out_srcs, int32_t *eid) {
  compress(nV, nE, dsts, srcs, offsets, out_srcs, eid); /* util.h:114-121 */
}

/* ---- the engine: refgas.h:30-225 --------------------------------------- */
typedef struct {
  size_t vsize, esize, gsize;
  void (*gather_zero)(void *out);
  void (*gather_map)(const void *dst, const void *src, const void *edge,
                     void *out);
  void (*gather_reduce)(const void *a, const void *b, void *out);
  int (*apply)(void *vertex, const void *gathered, int iteration);
} program;

typedef struct {
  const program *p;
  int nV, nE;
  char *vdata;       /* caller's vertex records, mutated in place (refgas.h:90) */
  const char *edata; /* caller's edge records, edge-list order (refgas.h:91)   */
  int32_t *srcs, *src_off, *eidx_csc; /* CSC, refgas.h:44-46 */
  int32_t *dsts, *dst_off, *eidx_csr; /* CSR, refgas.h:49-51 */
  char *gather_results;               /* refgas.h:54 */
  int32_t *active;                    /* refgas.h:55 */
  int n_active;
  int32_t *apply_ret;                 /* refgas.h:56 */
  unsigned char *active_flags;        /* refgas.h:57 */
  int iteration;                      /* the BFS driver's counter, bfs.cu:98 */
  gasref_stats st;
} engine;

/* refgas.h:81-113 */
static void eng_set_graph(engine *e, const program *p, int nV, void *vdata,
                          int nE, const void *edata, const int32_t *srcs,
                          const int32_t *dsts) {
  memset(e, 0, sizeof(*e));
  e->p = p;
  e->nV = nV;
  e->nE = nE;
  e->vdata = (char *)vdata;
  e->edata = (const char *)edata;
  e->dst_off = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nV + 1));
  e->dsts = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nE ? nE : 1));
  e->eidx_csr = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nE ? nE : 1));
  gasref_edge_list_to_csr(nV, nE, srcs, dsts, e->dst_off, e->dsts,
                          e->eidx_csr);
  e->src_off = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nV + 1));
  e->srcs = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nE ? nE : 1));
  e->eidx_csc = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nE ? nE : 1));
  gasref_edge_list_to_csc(nV, nE, srcs, dsts, e->src_off, e->srcs,
                          e->eidx_csc);
  e->active = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nV ? nV : 1));
  e->apply_ret = (int32_t *)calloc((size_t)(nV ? nV : 1), sizeof(int32_t));
  e->active_flags = (unsigned char *)calloc((size_t)(nV ? nV : 1), 1);
  e->gather_results = (char *)malloc(p->gsize * (size_t)(nV ? nV : 1));
}

static void eng_free(engine *e) {
  free(e->dst_off); free(e->dsts); free(e->eidx_csr);
  free(e->src_off); free(e->srcs); free(e->eidx_csc);
  free(e->active); free(e->apply_ret); free(e->active_flags);
  free(e->gather_results);
}

/* refgas.h:128-133 */
static void eng_set_active(engine *e, int start, int end) {
  e->n_active = 0;
  for (int i = start; i < end; ++i) e->active[e->n_active++] = i;
}

/* refgas.h:143-167: a gather for every active vertex over its CSC in-edges
 * seeded with gatherZero, then (separate loop, bulk synchronous) apply.
 * As in the reference, the gather loop runs even when haveGather is false;
 * its result is then simply ignored by programs without a gather. */
static void eng_gather_apply(engine *e, int have_gather) {
  const program *p = e->p;
  char tmp[64], sum[64];
  (void)have_gather;
  for (int i = 0; i < e->n_active; ++i) {
    int dv = e->active[i];
    p->gather_zero(sum);
    int es = e->src_off[dv], ee = e->src_off[dv + 1];
    for (int ie = es; ie < ee; ++ie) {
      int src = e->srcs[ie];
      p->gather_map(e->vdata + (size_t)dv * p->vsize,
                    e->vdata + (size_t)src * p->vsize,
                    e->edata ? e->edata + (size_t)e->eidx_csc[ie] * p->esize
                             : NULL,
                    tmp);
      p->gather_reduce(sum, tmp, sum);
    }
    memcpy(e->gather_results + (size_t)i * p->gsize, sum, p->gsize);
    e->st.gathered += ee - es;
  }
  for (int i = 0; i < e->n_active; ++i) {
    int dv = e->active[i];
    e->apply_ret[i] = p->apply(e->vdata + (size_t)dv * p->vsize,
                               e->gather_results + (size_t)i * p->gsize,
                               e->iteration);
  }
  e->st.active_total += e->n_active;
}

/* refgas.h:171-196: flags of all out-neighbours of vertices whose apply
 * returned true.  Program::scatter is a no-op in all four programs. */
static void eng_scatter_activate(engine *e) {
  memset(e->active_flags, 0, (size_t)e->nV);
  for (int i = 0; i < e->n_active; ++i) {
    if (!e->apply_ret[i]) continue;
    int sv = e->active[i];
    int es = e->dst_off[sv], ee = e->dst_off[sv + 1];
    for (int ie = es; ie < ee; ++ie) e->active_flags[e->dsts[ie]] = 1;
    e->st.returned += 1;
    e->st.scattered += ee - es;
  }
}

/* refgas.h:201-210: next frontier in ascending vertex order. */
static int eng_next_iter(engine *e) {
  e->n_active = 0;
  for (int i = 0; i < e->nV; ++i)
    if (e->active_flags[i]) e->active[e->n_active++] = i;
  e->st.supersteps += 1;
  return e->n_active;
}

/* refgas.h:216-224 */
static void eng_run(engine *e) {
  while (e->n_active) {
    eng_gather_apply(e, 1);
    eng_scatter_activate(e);
    eng_next_iter(e);
  }
}

/* ---- programs ------------------------------------------------------------ */
/* BFS, bfs.cu:31-78 (host path: level from the driver's counter) */
static void bfs_zero(void *o) { *(int *)o = INT_MAX - 1; }
static void bfs_map(const void *d, const void *s, const void *e, void *o) {
  (void)d; (void)s; (void)e;
  *(int *)o = 0;
}
static void bfs_reduce(const void *a, const void *b, void *o) {
  (void)a; (void)b;
  *(int *)o = 0;
}
static int bfs_apply(void *v, const void *g, int it) {
  (void)g;
  int *depth = (int *)v;
  if (*depth == -1) {
    *depth = it;
    return 1;
  }
  return 0;
}
static const program BFS_P = {4, 1, 4, bfs_zero, bfs_map, bfs_reduce,
                              bfs_apply};

/* SSSP, sssp.cu:25-64 */
static void sssp_zero(void *o) { *(int *)o = INT_MAX - 100000; }
static void sssp_map(const void *d, const void *s, const void *e, void *o) {
  (void)d;
  *(int *)o = *(const int *)s + *(const int *)e;
}
static void sssp_reduce(const void *a, const void *b, void *o) {
  int x = *(const int *)a, y = *(const int *)b;
  *(int *)o = x < y ? x : y;
}
static int sssp_apply(void *v, const void *g, int it) {
  (void)it;
  int *cur = (int *)v, dist = *(const int *)g;
  int changed = dist < *cur;
  if (dist < *cur) *cur = dist;
  return changed;
}
static const program SSSP_P = {4, 4, 4, sssp_zero, sssp_map, sssp_reduce,
                               sssp_apply};

/* CC, connected_component.cu:25-63 */
static void cc_zero(void *o) { *(int *)o = INT_MAX; }
static void cc_map(const void *d, const void *s, const void *e, void *o) {
  (void)d; (void)e;
  *(int *)o = *(const int *)s;
}
static const program CC_P = {4, 1, 4, cc_zero, cc_map, sssp_reduce,
                             sssp_apply};

/* PageRank, pagerank.cu:27-73 */
typedef struct {
  float rank;
  int num_out_edges;
} pr_vertex;
static void pr_zero(void *o) { *(float *)o = 0.0f; }
static void pr_map(const void *d, const void *s, const void *e, void *o) {
  (void)d; (void)e;
  const pr_vertex *src = (const pr_vertex *)s;
  *(float *)o = src->rank / (float)src->num_out_edges;
}
static void pr_reduce(const void *a, const void *b, void *o) {
  *(float *)o = *(const float *)a + *(const float *)b;
}
static int pr_apply(void *v, const void *g, int it) {
  (void)it;
  pr_vertex *x = (pr_vertex *)v;
  float new_rank = 0.15f + (1.0f - 0.15f) * *(const float *)g;
  int ret = fabsf(new_rank - x->rank) >= 0.01f;
  x->rank = new_rank;
  return ret;
}
static const program PR_P = {sizeof(pr_vertex), 1, 4, pr_zero, pr_map,
                             pr_reduce, pr_apply};

static void add_stats(gasref_stats *dst, const engine *e) {
  if (dst) *dst = e->st;
}

/* bfs.cu:91-112 with the initialisation of bfs.cu:144-146 */
int gasref_bfs(int nV, int nE, const int32_t *srcs, const int32_t *dsts,
               int source, int32_t *depth, gasref_stats *st) {
  if (source < 0 || source >= nV) return 1;
  engine e;
  for (int v = 0; v < nV; ++v) depth[v] = -1;
  eng_set_graph(&e, &BFS_P, nV, depth, nE, NULL, srcs, dsts);
  eng_set_active(&e, source, source + 1);
  e.iteration = 0;
  while (e.n_active) {
    eng_gather_apply(&e, 0);
    eng_scatter_activate(&e);
    eng_next_iter(&e);
    e.iteration++;
  }
  add_stats(st, &e);
  eng_free(&e);
  return 0;
}

/* sssp.cu:67-82 with the initialisation of sssp.cu:358-362 */
int gasref_sssp(int nV, int nE, const int32_t *srcs, const int32_t *dsts,
                const int32_t *weights, int source, int32_t *dist,
                gasref_stats *st) {
  if (source < 0 || source >= nV) return 1;
  engine e;
  for (int v = 0; v < nV; ++v) dist[v] = INT_MAX - 100000;
  dist[source] = 0;
  eng_set_graph(&e, &SSSP_P, nV, dist, nE, weights, srcs, dsts);
  eng_set_active(&e, 0, nV);
  eng_run(&e);
  add_stats(st, &e);
  eng_free(&e);
  return 0;
}

/* connected_component.cu:66-81 with the initialisation of :112-114 */
int gasref_cc(int nV, int nE, const int32_t *srcs, const int32_t *dsts,
              int32_t *label, gasref_stats *st) {
  engine e;
  for (int v = 0; v < nV; ++v) label[v] = v;
  eng_set_graph(&e, &CC_P, nV, label, nE, NULL, srcs, dsts);
  eng_set_active(&e, 0, nV);
  eng_run(&e);
  add_stats(st, &e);
  eng_free(&e);
  return 0;
}

/* pagerank.cu:85-101 with numOutEdges from the CSR (pagerank.cu:126-132) */
int gasref_pagerank(int nV, int nE, const int32_t *srcs, const int32_t *dsts,
                    int iterations, float *rank, gasref_stats *st) {
  engine e;
  pr_vertex *vd = (pr_vertex *)malloc(sizeof(pr_vertex) * (size_t)(nV + 1));
  int32_t *off = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nV + 1));
  gasref_edge_list_to_csr(nV, nE, srcs, dsts, off, NULL, NULL);
  for (int v = 0; v < nV; ++v) {
    vd[v].rank = 0.15f;
    vd[v].num_out_edges = off[v + 1] - off[v];
  }
  free(off);
  eng_set_graph(&e, &PR_P, nV, vd, nE, NULL, srcs, dsts);
  if (iterations > 0) {
    for (int it = 0; it < iterations; ++it) {
      eng_set_active(&e, 0, nV);
      eng_gather_apply(&e, 1);
      e.st.supersteps += 1;
    }
  } else {
    eng_set_active(&e, 0, nV);
    eng_run(&e);
  }
  for (int v = 0; v < nV; ++v) rank[v] = vd[v].rank;
  add_stats(st, &e);
  eng_free(&e);
  free(vd);
  return 0;
}

/* ---- independent checkers ----------------------------------------------- */
int gasref_check_bfs(int nV, int nE, const int32_t *srcs,
                     const int32_t *dsts, int source, int32_t *depth) {
  if (source < 0 || source >= nV) return 1;
  int32_t *off = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nV + 1));
  int32_t *adj = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nE ? nE : 1));
  int32_t *queue = (int32_t *)malloc(sizeof(int32_t) * (size_t)nV);
  gasref_edge_list_to_csr(nV, nE, srcs, dsts, off, adj, NULL);
  for (int v = 0; v < nV; ++v) depth[v] = -1;
  int head = 0, tail = 0;
  depth[source] = 0;
  queue[tail++] = source;
  while (head < tail) {
    int u = queue[head++];
    for (int k = off[u]; k < off[u + 1]; ++k) {
      int w = adj[k];
      if (depth[w] < 0) {
        depth[w] = depth[u] + 1;
        queue[tail++] = w;
      }
    }
  }
  free(off); free(adj); free(queue);
  return 0;
}

/* Binary-heap Dijkstra; unreached vertices get 2147383647 like the
 * reference's SSSP sentinel. */
typedef struct {
  int64_t d;
  int v;
} heap_item;

static void heap_push(heap_item *h, int *n, heap_item x) {
  int i = (*n)++;
  while (i > 0) {
    int p = (i - 1) / 2;
    if (h[p].d <= x.d) break;
    h[i] = h[p];
    i = p;
  }
  h[i] = x;
}

static heap_item heap_pop(heap_item *h, int *n) {
  heap_item top = h[0], last = h[--(*n)];
  int i = 0;
  for (;;) {
    int c = 2 * i + 1;
    if (c >= *n) break;
    if (c + 1 < *n && h[c + 1].d < h[c].d) ++c;
    if (last.d <= h[c].d) break;
    h[i] = h[c];
    i = c;
  }
  if (*n > 0) h[i] = last;
  return top;
}

int gasref_check_dijkstra(int nV, int nE, const int32_t *srcs,
                          const int32_t *dsts, const int32_t *weights,
                          int source, int32_t *dist) {
  if (source < 0 || source >= nV) return 1;
  int32_t *off = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nV + 1));
  int32_t *adj = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nE ? nE : 1));
  int32_t *eid = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nE ? nE : 1));
  int64_t *d = (int64_t *)malloc(sizeof(int64_t) * (size_t)nV);
  heap_item *h = (heap_item *)malloc(sizeof(heap_item) * (size_t)(nE + 1));
  int hn = 0;
  gasref_edge_list_to_csr(nV, nE, srcs, dsts, off, adj, eid);
  for (int v = 0; v < nV; ++v) d[v] = INT64_MAX;
  d[source] = 0;
  heap_item s0 = {0, source};
  heap_push(h, &hn, s0);
  while (hn) {
    heap_item x = heap_pop(h, &hn);
    if (x.d != d[x.v]) continue;
    for (int k = off[x.v]; k < off[x.v + 1]; ++k) {
      int w = adj[k];
      int64_t nd = x.d + weights[eid[k]];
      if (nd < d[w]) {
        d[w] = nd;
        heap_item y = {nd, w};
        heap_push(h, &hn, y);
      }
    }
  }
  for (int v = 0; v < nV; ++v)
    dist[v] = d[v] == INT64_MAX ? INT_MAX - 100000 : (int32_t)d[v];
  free(off); free(adj); free(eid); free(d); free(h);
  return 0;
}

/* Ascending sources: the first source that reaches v is the minimum.  A
 * search stops at labelled vertices, since everything they reach was
 * already labelled by a smaller source. */
int gasref_check_min_reaching(int nV, int nE, const int32_t *srcs,
                              const int32_t *dsts, int32_t *label) {
  int32_t *off = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nV + 1));
  int32_t *adj = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nE ? nE : 1));
  int32_t *stack = (int32_t *)malloc(sizeof(int32_t) * (size_t)(nV ? nV : 1));
  gasref_edge_list_to_csr(nV, nE, srcs, dsts, off, adj, NULL);
  for (int v = 0; v < nV; ++v) label[v] = -1;
  for (int u = 0; u < nV; ++u) {
    if (label[u] >= 0) continue;
    int sp = 0;
    label[u] = u;
    stack[sp++] = u;
    while (sp) {
      int x = stack[--sp];
      for (int k = off[x]; k < off[x + 1]; ++k) {
        int w = adj[k];
        if (label[w] < 0) {
          label[w] = u;
          stack[sp++] = w;
        }
      }
    }
  }
  free(off); free(adj); free(stack);
  return 0;
}

static int power_iteration(int nV, int nE, const int32_t *srcs,
                           const int32_t *dsts, int source, int iterations,
                           double *rank) {
  int32_t *outdeg = (int32_t *)calloc((size_t)(nV ? nV : 1), sizeof(int32_t));
  double *next = (double *)malloc(sizeof(double) * (size_t)(nV ? nV : 1));
  for (int k = 0; k < nE; ++k) outdeg[srcs[k]]++;
  for (int v = 0; v < nV; ++v)
    rank[v] = (source < 0) ? 0.15 : (v == source ? 0.15 : 0.0);
  for (int it = 0; it < iterations; ++it) {
    for (int v = 0; v < nV; ++v) next[v] = 0.0;
    for (int k = 0; k < nE; ++k)
      next[dsts[k]] += rank[srcs[k]] / (double)outdeg[srcs[k]];
    for (int v = 0; v < nV; ++v) {
      double tele = (source < 0) ? 0.15 : (v == source ? 0.15 : 0.0);
      rank[v] = tele + 0.85 * next[v];
    }
  }
  free(outdeg);
  free(next);
  return 0;
}

int gasref_check_pagerank_fp64(int nV, int nE, const int32_t *srcs,
                               const int32_t *dsts, int iterations,
                               double *rank) {
  return power_iteration(nV, nE, srcs, dsts, -1, iterations, rank);
}

int gasref_check_ppr_fp64(int nV, int nE, const int32_t *srcs,
                          const int32_t *dsts, int source, int iterations,
                          double *rank) {
  if (source < 0 || source >= nV) return 1;
  return power_iteration(nV, nE, srcs, dsts, source, iterations, rank);
}

/* ---- certificates (one pass over the edge list, no search) --------------- */
/* 1 iff depth is the BFS depth from source: depth[source] = 0; every edge
 * out of a reached vertex u ends at a reached vertex of depth <= depth[u] +
 * 1; every other reached vertex has an in-edge from depth - 1.  Any array
 * that passes is the unique BFS depth vector (induction on depth). */
int gasref_cert_bfs(int nV, int nE, const int32_t *srcs, const int32_t *dsts,
                    int source, const int32_t *depth) {
  if (source < 0 || source >= nV || depth[source] != 0) return 0;
  unsigned char *parent = (unsigned char *)calloc((size_t)(nV ? nV : 1), 1);
  int ok = 1;
  for (int k = 0; k < nE && ok; ++k) {
    int32_t du = depth[srcs[k]], dv = depth[dsts[k]];
    if (du < 0) continue;
    if (dv < 0 || dv > du + 1) ok = 0;
    else if (dv == du + 1) parent[dsts[k]] = 1;
  }
  for (int v = 0; v < nV && ok; ++v) {
    if (depth[v] < -1) ok = 0;
    else if (v != source && depth[v] >= 0 && !parent[v]) ok = 0;
  }
  free(parent);
  return ok;
}

/* 1 iff dist is the shortest-path distance from source under positive
 * weights with 2147383647 for unreached vertices: dist[source] = 0; no edge
 * out of a reached vertex can lower its head; every other reached vertex
 * has a tight in-edge (dist[v] = dist[u] + w).  With positive weights the
 * tight edges form a tree rooted at source, so a passing array is the
 * unique distance vector. */
int gasref_cert_sssp(int nV, int nE, const int32_t *srcs, const int32_t *dsts,
                     const int32_t *weights, int source, const int32_t *dist) {
  const int32_t inf = INT_MAX - 100000;
  if (source < 0 || source >= nV || dist[source] != 0) return 0;
  unsigned char *parent = (unsigned char *)calloc((size_t)(nV ? nV : 1), 1);
  int ok = 1;
  for (int k = 0; k < nE && ok; ++k) {
    int32_t du = dist[srcs[k]];
    if (du >= inf) continue;
    if (weights[k] <= 0) { ok = 0; break; }
    int64_t cand = (int64_t)du + weights[k];
    int64_t dv = dist[dsts[k]];
    if (dv > cand) ok = 0;
    else if (dv == cand) parent[dsts[k]] = 1;
  }
  for (int v = 0; v < nV && ok; ++v) {
    if (dist[v] < 0 || dist[v] > inf) ok = 0;
    else if (v != source && dist[v] < inf && !parent[v]) ok = 0;
  }
  free(parent);
  return ok;
}
