/*
 * gasref.h -- CPU oracle for the GAS hot path.  TEST INFRASTRUCTURE ONLY:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and only as the checker.  The product (libgas.so) never calls it.
 *
 * gasref_engine_* is a plain-C restatement of the reference's CPU engine
 * GASEngineRef (refgas.h:30-225) and of the CSR/CSC construction it uses
 * (util.h:48-121).  gasref_bfs / _sssp / _cc / _pagerank restate the app
 * drivers on top of it.  The remaining functions are independent checkers
 * that do not share code with the engine (textbook BFS, Dijkstra, minimum
 * reaching id, FP64 power iteration), used to pin the restatement.
 */
#ifndef GASREF_H_
#define GASREF_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Per-run counters of the oracle engine (sums over supersteps). */
typedef struct {
  int64_t supersteps;
  int64_t active_total;   /* sum |A_i|                                   */
  int64_t gathered;       /* sum of in-degrees of A_i (gather ran)       */
  int64_t returned;       /* sum |R_i|, vertices whose apply returned 1  */
  int64_t scattered;      /* sum of out-degrees of R_i                   */
} gasref_stats;

/* ---- CSR/CSC (util.h:79-121) ----------------------------------------- */
/* offsets: V+1; out_vals/eid: E.  Stable in edge-list order. */
void gasref_edge_list_to_csr(int nV, int nE, const int32_t *srcs,
                             const int32_t *dsts, int32_t *offsets,
                             int32_t *out_dsts, int32_t *eid);
void gasref_edge_list_to_csc(int nV, int nE, const int32_t *srcs,
                             const int32_t *dsts, int32_t *offsets,
                             int32_t *out_srcs, int32_t *eid);

/* ---- app drivers over the restated engine ----------------------------- */
int gasref_bfs(int nV, int nE, const int32_t *srcs, const int32_t *dsts,
               int source, int32_t *depth, gasref_stats *st);
int gasref_sssp(int nV, int nE, const int32_t *srcs, const int32_t *dsts,
                const int32_t *weights, int source, int32_t *dist,
                gasref_stats *st);
int gasref_cc(int nV, int nE, const int32_t *srcs, const int32_t *dsts,
              int32_t *label, gasref_stats *st);
/* iterations > 0: that many all-active sweeps; 0: engine.run() with the
 * program's tolerance (pagerank.cu:89-97). */
int gasref_pagerank(int nV, int nE, const int32_t *srcs, const int32_t *dsts,
                    int iterations, float *rank, gasref_stats *st);

/* ---- independent checkers --------------------------------------------- */
int gasref_check_bfs(int nV, int nE, const int32_t *srcs,
                     const int32_t *dsts, int source, int32_t *depth);
int gasref_check_dijkstra(int nV, int nE, const int32_t *srcs,
                          const int32_t *dsts, const int32_t *weights,
                          int source, int32_t *dist);
/* label[v] = smallest id u with a directed path u -> v (u = v allowed). */
int gasref_check_min_reaching(int nV, int nE, const int32_t *srcs,
                              const int32_t *dsts, int32_t *label);
/* r_0 = 0.15; r_{k+1}(v) = 0.15 + 0.85 * sum_{u->v} r_k(u)/outdeg(u), in
 * FP64 with pairwise-free plain summation. */
int gasref_check_pagerank_fp64(int nV, int nE, const int32_t *srcs,
                               const int32_t *dsts, int iterations,
                               double *rank);
/* Personalised PageRank with teleport to `source`: r_0 = 0.15 e_s;
 * r_{k+1}(v) = 0.15 [v == s] + 0.85 * sum_{u->v} r_k(u)/outdeg(u), FP64. */
int gasref_check_ppr_fp64(int nV, int nE, const int32_t *srcs,
                          const int32_t *dsts, int source, int iterations,
                          double *rank);

/* Certificates: 1 iff the array is the exact BFS depth / shortest-path
 * distance (unreached 2147383647) from source, checked in one pass over the
 * edge list (properties that fix the answer uniquely; no search). */
int gasref_cert_bfs(int nV, int nE, const int32_t *srcs, const int32_t *dsts,
                    int source, const int32_t *depth);
int gasref_cert_sssp(int nV, int nE, const int32_t *srcs, const int32_t *dsts,
                     const int32_t *weights, int source, const int32_t *dist);

#ifdef __cplusplus
}
#endif

#endif
