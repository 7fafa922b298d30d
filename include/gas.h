/*
 * gas.h -- C ABI of the MI355X Gather-Apply-Scatter engine (libgas.so).
 *
 * Every entry point returns an int status (GAS_OK == 0) and leaves a
 * human-readable message for the calling thread in gas_last_error().
 * Pointers are plain host or device pointers; sizes are element counts
 * unless the name says bytes.  Vertex and edge ids are int32 (the
 * reference's default Int, gpugas.h:75 / refgas.h:31).
 *
 * Groups:
 *   gasrt_*      thin HIP runtime layer.  Replaces the reference's missing
 *                cuda_util.h helpers gpuAlloc/gpuFree/copyToGPU/copyToHost/
 *                CHECK/SYNC_CHECK (used at gpugas.h:177-195, 227-318).
 *   gas_edges_*  host edge lists: seeded generators and the MatrixMarket
 *                loader (graphio.cpp:190-262).
 *   gas_graph_*  a graph resident in HBM as CSC (gather) + CSR (scatter),
 *                the layout GASEngineGPU::setGraph builds (gpugas.h:212-322).
 *   gas_engine_* the reference engine API, step by step, for the four
 *                programs (gpugas.h:212-772, refgas.h:81-224).
 *   gas_bfs/gas_sssp/gas_cc/gas_pagerank  whole runs, the app drivers'
 *                run<Engine>() (bfs.cu:91-112, sssp.cu:67-82,
 *                connected_component.cu:66-81, pagerank.cu:85-101).
 *   gas_batch_*  64-source batched BFS / SSSP / personalised PageRank.
 *   gas_server_* continuous-batching query server over resident graphs.
 */
#ifndef GAS_H_
#define GAS_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define GAS_OK              0
#define GAS_ERR_ARG         1   /* invalid argument                          */
#define GAS_ERR_HIP         2   /* HIP runtime error                         */
#define GAS_ERR_NOMEM       3   /* allocation refused (device or host)       */
#define GAS_ERR_IO          4   /* file could not be read / parsed           */
#define GAS_ERR_STATE       5   /* call not valid in the object's state      */
#define GAS_ERR_FULL        6   /* server admission refused (no slot/memory) */

/* Message describing the last failure on the calling thread ("" if none). */
const char *gas_last_error(void);

/* ---- runtime layer (replaces cuda_util.h, gpugas.h:65) ------------------ */
typedef void *gasrt_stream_t;
typedef void *gasrt_event_t;
typedef void *gasrt_graph_exec_t;

int gasrt_device_count(int *count);
int gasrt_set_device(int device);
/* Device properties used by the engine's launch sizing. */
int gasrt_device_info(int device, int *cu_count, int *lds_bytes_per_block,
                      size_t *hbm_bytes, char *name, size_t name_len);
int gasrt_malloc(void **ptr, size_t bytes);                 /* gpuAlloc   */
int gasrt_free(void *ptr);                                  /* gpuFree    */
int gasrt_host_alloc(void **ptr, size_t bytes);             /* pinned     */
int gasrt_host_free(void *ptr);
/* Pinned host memory that kernels can write (coherent, mapped), with the
 * device address to hand to kernels; freed with gasrt_host_free.  No
 * counterpart in the reference beyond its mapped host scalar
 * (gpugas.h:120-122). */
int gasrt_host_alloc_mapped(void **host_ptr, void **dev_ptr, size_t bytes);
int gasrt_memcpy_h2d(void *dst, const void *src, size_t bytes,
                     gasrt_stream_t stream);                /* copyToGPU  */
int gasrt_memcpy_d2h(void *dst, const void *src, size_t bytes,
                     gasrt_stream_t stream);                /* copyToHost */
int gasrt_memcpy_d2d(void *dst, const void *src, size_t bytes,
                     gasrt_stream_t stream);
/* Asynchronous device-to-host copy (no synchronisation; dst should be
 * pinned memory from gasrt_host_alloc). */
int gasrt_memcpy_d2h_async(void *dst, const void *src, size_t bytes,
                           gasrt_stream_t stream);
int gasrt_memset(void *dst, int value, size_t bytes, gasrt_stream_t stream);
int gasrt_stream_create(gasrt_stream_t *stream);
int gasrt_stream_destroy(gasrt_stream_t stream);
int gasrt_stream_synchronize(gasrt_stream_t stream);        /* SYNC_CHECK */
/* *idle = 1 when all work queued on the stream has finished. */
int gasrt_stream_query(gasrt_stream_t stream, int *idle);
int gasrt_event_create(gasrt_event_t *event);
int gasrt_event_destroy(gasrt_event_t event);
int gasrt_event_record(gasrt_event_t event, gasrt_stream_t stream);
int gasrt_event_synchronize(gasrt_event_t event);
int gasrt_event_query(gasrt_event_t event, int *done);
int gasrt_event_elapsed_ms(gasrt_event_t start, gasrt_event_t stop, float *ms);
/* Launch a __global__ function by its host stub pointer. */
int gasrt_launch(const void *func, unsigned grid_x, unsigned block_x,
                 void **args, size_t shared_bytes, gasrt_stream_t stream);
int gasrt_capture_begin(gasrt_stream_t stream);
int gasrt_capture_end(gasrt_stream_t stream, gasrt_graph_exec_t *exec);
int gasrt_graph_launch(gasrt_graph_exec_t exec, gasrt_stream_t stream);
int gasrt_graph_destroy(gasrt_graph_exec_t exec);
int gasrt_mem_info(size_t *free_bytes, size_t *total_bytes);

/* ---- host edge lists ---------------------------------------------------- */
typedef struct gas_edges *gas_edges_t;

/* Borrowed view; arrays stay valid until gas_edges_free. */
typedef struct {
  int32_t n_vertices;
  int32_t n_edges;
  const int32_t *srcs;
  const int32_t *dsts;
  const int32_t *weights; /* NULL when the list carries no edge values */
} gas_edges_view;

int gas_edges_from_arrays(int32_t n_vertices, int32_t n_edges,
                          const int32_t *srcs, const int32_t *dsts,
                          const int32_t *weights /* may be NULL */,
                          gas_edges_t *out);
int gas_edges_view_get(gas_edges_t e, gas_edges_view *view);
int gas_edges_free(gas_edges_t e);
/* rows x cols 4-neighbour grid, id = r*cols + c, both directions, unit
 * weights. */
int gas_gen_grid(int32_t rows, int32_t cols, gas_edges_t *out);
/* Road-like grid: drop floor(drop_frac*L) of the L undirected grid links,
 * add floor(long_frac*L) random long links (distinct, no self loops, no
 * duplicates), symmetric weights U[wmin,wmax]. */
int gas_gen_road(int32_t rows, int32_t cols, double drop_frac,
                 double long_frac, int32_t wmin, int32_t wmax, uint64_t seed,
                 gas_edges_t *out);
/* R-MAT: edge_factor * 2^scale samples with quadrant probabilities a,b,c
 * (d = 1-a-b-c); self loops and duplicates removed; optionally
 * symmetrised; weights U[wmin,wmax] shared by both directions of a pair
 * (wmax == 0: no weights). */
int gas_gen_rmat(int32_t scale, int32_t edge_factor, double a, double b,
                 double c, uint64_t seed, int symmetrize, int32_t wmin,
                 int32_t wmax, gas_edges_t *out);
/* Graph files (replace loadGraph and its loaders / writers,
 * graphio.cpp:178-473).  Every file may be gzip-compressed ("x.mtx.gz").
 * want_values: 1 as the reference's edgeValues argument, 0 for none, -1
 * for the values the file holds, if any (the converters' use).
 *
 * MatrixMarket coordinate loader with the reference's rules
 * (graphio.cpp:69-262): n_vertices = max id + 1, self loops dropped,
 * symmetric files expanded to both directions, values read as float and
 * stored as int.  want_values mirrors passing edgeValues (sssp.cu:112):
 * with it a self loop does not count toward max id, as in the reference. */
int gas_load_mtx(const char *path, int want_values, gas_edges_t *out);
/* GraphLab SNAP edge list (loadGraph_GraphLabSnap, graphio.cpp:178-188):
 * '#' comments, 0-based "src dst" lines, no values, self loops dropped
 * (still counted toward max id). */
int gas_load_edge(const char *path, gas_edges_t *out);
/* Lonestar binary CSR version 1 (loadGraph_binaryCSR, graphio.cpp:266-362):
 * u64 version, u64 edge-value bytes (0 or 4), u64 vertices, u64 edges,
 * u64 end offset per vertex, u32 destination per edge (+ 4 bytes of padding
 * when the count is odd), u32 value per edge; expanded to an edge list in
 * CSR order.  want_values on a file without values fails as in the
 * reference; unlike it (graphio.cpp:305-309), a file without values loads
 * when none are wanted. */
int gas_load_gr(const char *path, int want_values, gas_edges_t *out);
/* Writes the list as Lonestar binary CSR (writeGraph_binaryCSR,
 * graphio.cpp:365-414; with its values if it has them), each source's
 * edges in list order (mtx2gr.cpp:41-51). */
int gas_write_gr(const char *path, gas_edges_t edges);
/* Writes the list as MatrixMarket (writeGraph_mtx, graphio.cpp:416-439):
 * the reference's banner "%MatrixMarket matrix coordinate Integer general"
 * (one '%'; "pattern" for a list without values), "V V E", then 1-based
 * "src dst [value]" lines in list order. */
int gas_write_mtx(const char *path, gas_edges_t edges);
/* loadGraph (graphio.cpp:441-473): by extension, ".edge", ".mtx" or ".gr",
 * after stripping a final ".gz". */
int gas_load_graph(const char *path, int want_values, gas_edges_t *out);
/* The reference's -m option (bfs.cu:148-166): vertex of maximum out-degree,
 * lowest id on ties. */
int gas_edges_max_outdegree_vertex(gas_edges_t e, int32_t *vertex,
                                   int32_t *degree);

/* ---- device-resident graph ---------------------------------------------- */
typedef struct gas_graph *gas_graph_t;

typedef struct {
  int32_t n_vertices;
  int32_t n_edges;
  const int32_t *csc_offsets; /* device, V+1: in-edges of v at [o[v],o[v+1]) */
  const int32_t *csc_srcs;    /* device, E                                  */
  const int32_t *csr_offsets; /* device, V+1: out-edges                      */
  const int32_t *csr_dsts;    /* device, E                                  */
  const int32_t *csc_weights; /* device, E, or NULL                         */
  const int32_t *csr_weights; /* device, E, or NULL (same weights, CSR order) */
  size_t device_bytes;
} gas_graph_view;

/* Builds CSC and CSR (the layout of util.h:79-121: offsets, neighbour ids
 * grouped by key vertex, equal keys in edge-list order) and keeps them in
 * HBM, with host copies.  Lists of at least the device-build threshold of
 * edges are sorted on the GPU (stable radix sort of (key, edge index)),
 * smaller ones on the host (stable counting sort); both give the same
 * layout. */
int gas_graph_create(gas_edges_t edges, gas_graph_t *out);
/* Process-wide: edge lists of at least min_edges edges (default 2^20) are
 * laid out on the GPU by gas_graph_create, smaller ones on the host; 0
 * builds every graph on the GPU.  No counterpart in the reference, which
 * sorts on the host (util.h:68). */
int gas_graph_set_device_build(int64_t min_edges);
/* The whole-run entry points of BFS, SSSP, PageRank, CC and personalised
 * PageRank (gas_bfs[_device], gas_sssp[_device], gas_pagerank[_device],
 * gas_cc[_device], gas_ppr[_device], and gas_graph_autotune /
 * gas_graph_set_tiers for them) and the batch engines on a skewed graph (at least 2^20 edges, largest in-degree at
 * least 32 times the mean) run on a copy of it with the vertices
 * renumbered by descending in-degree (mode 1, default; built at
 * first use, results returned in the caller's numbering, CC labels are the
 * caller's ids, sources are mapped), or on the graph as given (mode 0).
 * The order puts hub destinations together and, on skewed graphs, the
 * most-read sources in shared cache lines.  Integer results are identical
 * in both modes; PageRank sums differ only by FP32 rounding order.
 * Changing the mode frees the ordered copy: free the batch engines and
 * servers created on g before.  No counterpart in the reference. */
int gas_graph_set_vertex_order(gas_graph_t g, int32_t mode);
/* The graph the batch engines run on: g's degree-ordered copy (built if
 * needed) with each of g's vertices' id in it (device and host arrays of
 * n_vertices, valid while g lives and its order mode is unchanged), or g
 * itself and NULLs when g is not reordered (mode 0, or not skewed).  Batch
 * sources and results stay in g's numbering. */
int gas_graph_ordered_view(gas_graph_t g, gas_graph_t *ordered,
                           const int32_t **new_of_dev, const int32_t **new_of_host);
int gas_graph_view_get(gas_graph_t g, gas_graph_view *view);
/* Out-neighbour table of a graph whose out-degrees are all at most 8
 * (device arrays of 8 * n_vertices int32: row v holds v's out-neighbours in
 * CSR order, then -1; and their weights, then 0, or NULL without weights),
 * built at the first call and kept with the graph; NULLs for graphs with
 * larger out-degrees.  The whole-run engines read a vertex's out-edges from
 * it without first loading its CSR offsets.  No counterpart in the
 * reference. */
int gas_graph_ell(gas_graph_t g, const int32_t **ell_dst, const int32_t **ell_w);
/* Host copies for checking: arrays of V+1 / E int32 (any may be NULL). */
int gas_graph_host_csc(gas_graph_t g, int32_t *offsets, int32_t *srcs,
                       int32_t *edge_index);
int gas_graph_host_csr(gas_graph_t g, int32_t *offsets, int32_t *dsts,
                       int32_t *edge_index);
/* Reorder per-edge records of elem_bytes each from edge-list order into
 * CSC order (host to host). */
int gas_graph_permute_to_csc(gas_graph_t g, const void *in, void *out,
                             size_t elem_bytes);
int gas_graph_free(gas_graph_t g);

/* ---- run statistics ----------------------------------------------------- */
/* Kernel slots of gas_run_stats.kernel_ms / kernel_launches. */
enum {
  GAS_K_GATHER = 0,    /* k_gather_vp: thread and wave64 gather tiers      */
  GAS_K_APPLY = 1,     /* k_apply: apply + fused thread/wave activation    */
  GAS_K_SCATTER = 2,   /* k_scatter_vp: stepwise scatterActivate           */
  GAS_K_HSCAN1 = 3,    /* k_hscan1: hub degree scan, pass 1                */
  GAS_K_HSCAN2 = 4,    /* k_hscan2: hub degree scan, pass 2                */
  GAS_K_HGATHER = 5,   /* k_hgather: merge-path gather of hub in-edges     */
  GAS_K_HFIX = 6,      /* k_hfix: fold of hub tile partials                */
  GAS_K_HSCATTER = 7,  /* k_hscatter: merge-path activation over hub edges */
  GAS_K_COMPACT1 = 8,  /* k_compact1: frontier count per block             */
  GAS_K_COMPACT2 = 9,  /* k_compact2: ordered frontier write               */
  GAS_K_FUSED = 10,    /* k_fused: gather + apply + scatter in one launch  */
  GAS_K_SMALL = 11,    /* k_small: supersteps of a small frontier, one WG  */
  GAS_K_PUSH = 12,     /* k_push: min-program superstep pushed by changes  */
  GAS_K_SLOTS = 13
};

typedef struct {
  int32_t supersteps;        /* BSP supersteps executed                    */
  int32_t reserved;
  int64_t gathered_edges;    /* sum over supersteps of in-edges gathered   */
  int64_t scattered_edges;   /* sum over supersteps of out-edges scattered */
  int64_t active_total;      /* sum over supersteps of |active|            */
  double  ms;                /* device time of the run (HIP events)        */
  /* The fields below are filled only while gas_set_kernel_timing is on. */
  double  gather_ms;         /* gather phase: k_gather_vp + hub gather     */
  double  scatter_ms;        /* k_apply / k_scatter_vp + hub scatter       */
  int32_t gather_launches;   /* k_gather_vp launches                       */
  int32_t scatter_launches;  /* k_apply + k_scatter_vp launches            */
  double  kernel_ms[GAS_K_SLOTS];        /* per kernel, summed over launches */
  int32_t kernel_launches[GAS_K_SLOTS];
  int64_t returned_total;    /* sum over supersteps of vertices whose apply */
                             /* returned true (|R_i|)                      */
  int64_t hub_vertices_total; /* of active_total: gathered by the hub      */
  int64_t hub_edges_total;    /* kernels, and their in-edges               */
} gas_run_stats;

/* Set to nonzero to bracket every engine kernel launch with HIP events on
 * the engine's stream (adds event packets between launches; used for the
 * per-kernel roofline figures). */
int gas_set_kernel_timing(int enable);

/* Launches an empty kernel (k_trace_mark) of `tag` workgroups of 64 lanes
 * on the default stream and waits for it, so that a kernel trace taken by a
 * profiler can be cut at this point: the mark's grid size is 64 * tag.
 * tag in 1..65535.  No counterpart in the reference. */
int gas_trace_mark(uint32_t tag);

/* Degree tiers of the engine the whole-run entry points keep for (graph,
 * program): tiers[0..1] for the gather, tiers[2..3] for the scatter.  A
 * vertex of degree d is walked by one thread if d <= t1, by one wave64 if
 * t1 < d <= t2, and split into merge-path tiles above t2 (t1 <= 16).
 * No counterpart in the reference, whose gather is one merge path
 * (gpugas.h:421-513). */
int gas_graph_set_tiers(gas_graph_t g, int program, const int32_t *tiers);
int gas_graph_get_tiers(gas_graph_t g, int program, int32_t *tiers);
/* Frontier representation of whole runs: a superstep's scatter builds the
 * next frontier as a list of ids (claimed with an atomic OR, appended, work
 * proportional to the frontier) while the last reported frontier has
 * |A| * divisor < V, and as activation bytes scanned in O(V) otherwise.
 * divisor 0 = always activation bytes; default 64 for graphs of at least
 * 2^24 vertices, 0 below.  Replaces the reference's dense / sparse scatter
 * choice (gpugas.h:669-751). */
int gas_graph_set_frontier_switch(gas_graph_t g, int program, int32_t divisor);
/* Small-frontier mode of whole runs on graphs whose in- and out-degrees are
 * all at most 64: while a list frontier has at most max_vertices vertices,
 * consecutive supersteps run inside one workgroup (barriers instead of
 * kernel launches).  0 turns it off; default 4096 for graphs of at most
 * 8192 vertices, 0 above (measured slower there: each vertex's loads and
 * claims form one long dependent chain inside the workgroup).  No
 * counterpart in the reference, which launches every superstep
 * (gpugas.h:764-772). */
int gas_graph_set_small_frontier(gas_graph_t g, int program,
                                 int32_t max_vertices);
/* Push supersteps of whole SSSP / CC runs (enable 1, the default, or 0):
 * the vertices whose value changed push it along their out-edges with
 * atomic minima into the other state buffer, instead of every
 * out-neighbour of theirs pulling over all its in-edges.  The values after
 * every superstep, hence the results and the superstep counts, are the
 * same; used when every out-degree is at most 64, the frontier switch is 0
 * and the small-frontier mode is off for the graph.  No counterpart in the
 * reference, whose engine only pulls (gpugas.h:421-513). */
int gas_graph_set_push(gas_graph_t g, int program, int32_t enable);
/* Autotuner: times candidate tiers on this graph for the program (gather
 * and scatter phases separately, vertex data untouched) and keeps the
 * fastest; tiers_out (may be NULL) receives them. */
int gas_graph_autotune(gas_graph_t g, int program, int32_t *tiers_out);

/* ---- whole runs (app drivers) ------------------------------------------- */
/* BFS (bfs.cu:31-112): depth[v] = superstep at which v is first reached,
 * -1 if unreached.  stats may be NULL. */
int gas_bfs(gas_graph_t g, int32_t source, int32_t *depth_out,
            gas_run_stats *stats);
/* SSSP (sssp.cu:25-82): dist = min over paths of summed int32 weights;
 * unreached vertices keep gatherZero = 2147383647. */
int gas_sssp(gas_graph_t g, int32_t source, int32_t *dist_out,
             gas_run_stats *stats);
/* The same two runs with the result left in HBM: the initial vertex state
 * is written on the device (no upload) and the result copied device to
 * device into depth_dev / dist_dev (n_vertices int32 each).  stats->ms
 * covers initialisation, the run and the copy. */
int gas_bfs_device(gas_graph_t g, int32_t source, int32_t *depth_dev,
                   gas_run_stats *stats);
int gas_sssp_device(gas_graph_t g, int32_t source, int32_t *dist_dev,
                    gas_run_stats *stats);
/* Connected components by min-label propagation over in-edges
 * (connected_component.cu:25-81); label[v] initialised to v. */
int gas_cc(gas_graph_t g, int32_t *label_out, gas_run_stats *stats);
/* PageRank (pagerank.cu:27-101): rank = 0.15 + 0.85 * sum rank/outdeg,
 * FP32 with compensated summation.  iterations > 0: that many sweeps with
 * every vertex active; iterations == 0: the reference's tolerance-driven
 * run (tol 0.01).  residual_out (may be NULL) receives, per sweep, the
 * FP64 sum of |new - old| reduced on the GPU. */
int gas_pagerank(gas_graph_t g, int32_t iterations, float *rank_out,
                 double *residual_out, gas_run_stats *stats);
/* The same two runs with the result left in HBM (label_dev: n_vertices
 * int32; rank_dev: n_vertices floats, residual_dev: `iterations` doubles or
 * NULL, in which case the per-sweep residual is not reduced); the initial
 * state is written on the device.  PageRank here runs fixed sweeps only
 * (iterations >= 1). */
int gas_cc_device(gas_graph_t g, int32_t *label_dev, gas_run_stats *stats);
int gas_pagerank_device(gas_graph_t g, int32_t iterations, float *rank_dev,
                        double *residual_dev, gas_run_stats *stats);

/* gas_pagerank_device without allocating: no counterpart in the reference;
 * for callers that cannot allocate device or pinned memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid allocation).
 * Every buffer the sweeps need comes from 64 MiB of device memory the
 * library holds on each GPU it runs on until the process ends, enough for
 * graphs up to about 1M vertices (a larger graph fails with GAS_ERR_ARG
 * before any work); kernels run on `stream` (NULL: the null stream) and the
 * call returns when the ranks are in rank_dev.  Runs on g's degree-ordered
 * copy only if one already exists (gas_pagerank_device builds it), with the
 * degree tiers of g's whole-run PageRank engine when one exists.  All calls
 * on one GPU share that memory: calls that overlap on one GPU race. */
int gas_pagerank_device_noalloc(gas_graph_t g, int32_t iterations, float *rank_dev,
                                double *residual_dev, gasrt_stream_t stream);

/* Personalised PageRank from one source, `iterations` all-active sweeps in
 * FP32 with compensated summation: r_0 = 0.15 e_s, r_{k+1}(v) = 0.15 [v==s]
 * + 0.85 sum_{u->v} r_k(u)/outdeg(u) (the PageRank program of
 * pagerank.cu:27-73 with the teleport paid at the source only). */
int gas_ppr(gas_graph_t g, int32_t source, int32_t iterations,
            float *rank_out, gas_run_stats *stats);
/* The same with the ranks left in HBM (rank_dev: n_vertices floats); the
 * initial state is written on the device. */
int gas_ppr_device(gas_graph_t g, int32_t source, int32_t iterations,
                   float *rank_dev, gas_run_stats *stats);

/* ---- batched multi-source queries ------------------------------------------
 * Up to 64 BFS or SSSP queries on one resident graph, one per lane, sharing
 * every edge read (no counterpart in the reference, which runs one source
 * per engine, bfs.cu:91-112, sssp.cu:67-82).  Lanes can be added between
 * rounds while others are running (continuous batching); a lane is finished
 * when its bit leaves gas_batch_alive.  Results per lane are identical to
 * the single-source runs: BFS depth (-1 unreached, at most 65534 levels),
 * SSSP distance (2147383647 unreached). */
#define GAS_BATCH_BFS   0
#define GAS_BATCH_SSSP  1
#define GAS_BATCH_PPR   2   /* personalised PageRank, FP32 lanes */
typedef struct gas_batch *gas_batch_t;

/* kind GAS_BATCH_PPR runs 10 iterations per query; see gas_batch_create_ppr. */
int gas_batch_create(gas_graph_t g, int kind, gas_batch_t *out);
/* Personalised PageRank lanes with `iterations` sweeps each (the recurrence
 * of gas_ppr); a lane's result must be read before the next step after it
 * finishes. */
int gas_batch_create_ppr(gas_graph_t g, int32_t iterations, gas_batch_t *out);
/* Starts n queries: lanes[i] (0..63, free) runs from sources[i]. */
int gas_batch_add(gas_batch_t b, const int32_t *lanes, const int32_t *sources,
                  int32_t n);
/* Runs `rounds` rounds (BFS levels / SSSP relaxation rounds) of all lanes. */
int gas_batch_step(gas_batch_t b, int32_t rounds);
/* Lanes in use whose query still has work. */
int gas_batch_alive(gas_batch_t b, uint64_t *alive);
/* Lanes in use (running or finished but not yet released). */
int gas_batch_used(gas_batch_t b, uint64_t *used);
int gas_batch_release(gas_batch_t b, uint64_t lanes);
/* Rounds until every lane in use is finished; stats may be NULL. */
int gas_batch_run(gas_batch_t b, gas_run_stats *stats);
/* Result of one lane: V int32 to host memory / to device memory (complete
 * when the call returns). */
int gas_batch_result(gas_batch_t b, int32_t lane, int32_t *host_out);
int gas_batch_result_device(gas_batch_t b, int32_t lane, int32_t *dev_out);
/* Results of every lane set in `lanes` in one pass over the lane state:
 * lane s into dev_outs[s] (device, n_vertices int32 each; entries of lanes
 * not in the mask are ignored).  BFS / SSSP batches. */
int gas_batch_results_device(gas_batch_t b, uint64_t lanes,
                             int32_t *const *dev_outs);
/* PPR lanes: V float to host memory / to device memory. */
int gas_batch_result_f32(gas_batch_t b, int32_t lane, float *host_out);
int gas_batch_result_f32_device(gas_batch_t b, int32_t lane, float *dev_out);
/* PPR lanes in `lanes` into dev_outs[lane] (V float each, device), in one
 * pass over the rank rows (only the 64-byte sectors holding those lanes). */
int gas_batch_results_f32_device(gas_batch_t b, uint64_t lanes,
                                 float *const *dev_outs);
/* Push-versus-pull switch of batched BFS: pull when the frontier's
 * out-edges times `divisor` reach |E| (default 0 = never pull: with 64
 * lanes a pull rarely stops early, measured slower on every R-MAT round). */
int gas_batch_set_pull_divisor(gas_batch_t b, int32_t divisor);
/* Batched SSSP: in round r a lane relaxes the out-edges of a changed vertex
 * only while its distance there is below (r - r0 + 1) * delta (r0: the
 * lane's first round); changed lanes above the bound wait.  Fewer
 * re-relaxations than Bellman-Ford rounds, same (unique) distances.
 * 0 = no bound. */
int gas_batch_set_sssp_delta(gas_batch_t b, int32_t delta);
/* Batch autotuner: times 64-lane batches from six fixed samples of
 * sources (uniform over the vertices with out-edges) on this graph for each
 * candidate setting of `kind` and keeps in the graph the fastest in total
 * among the default (the first candidate) and the settings no slower than
 * 1.1x the default on any sample -- GAS_BATCH_BFS: pull divisor in {0, 2, 8, 32};
 * GAS_BATCH_SSSP: distance step in {0, w/4, w/2, w} for the largest edge
 * weight w -- then, with that setting, the hub threshold
 * (gas_graph_batch_hub_items), then the push item size
 * (gas_graph_batch_item_edges).  Batch engines created on the graph
 * afterwards (gas_batch_create, the server's engines) start with all three;
 * param_out (may be NULL) receives the first. */
int gas_graph_autotune_batch(gas_graph_t g, int kind, int32_t *param_out);
/* The setting batch engines on this graph start with (0 until tuned). */
int gas_graph_batch_param(gas_graph_t g, int kind, int32_t *param);
/* Hub threshold of the batch engines on this graph: a list vertex with more
 * push items (16 out-edges each for BFS, 64 for SSSP) than this has its
 * items written by a whole workgroup.  The tuned value (the autotuner's
 * second stage, {32, 16, 64, 128} under the same rule, with the first
 * stage's setting), 32 until tuned.  No counterpart in the reference. */
int gas_graph_batch_hub_items(gas_graph_t g, int kind, int32_t *items);
/* Sets one batch engine's hub threshold (>= 16 items). */
int gas_batch_set_hub_items(gas_batch_t b, int32_t items);
/* Out-edges per push item of the batch engines on this graph: the tile a
 * list vertex's out-edges are cut into (BFS 16 or 8, SSSP 64 or 32).  The
 * tuned value (the autotuner's third stage, under the same rule), the
 * larger size until tuned.  No counterpart in the reference. */
int gas_graph_batch_item_edges(gas_graph_t g, int kind, int32_t *edges);
/* Sets one batch engine's item size (a power of two, BFS 8..16, SSSP
 * 32..64); takes effect from the next list. */
int gas_batch_set_item_edges(gas_batch_t b, int32_t edges);
/* Rounds per launch of a BFS / SSSP batch engine's captured round graph:
 * k in 1..8 records k consecutive rounds' kernel launches once (hipGraph
 * stream capture) and issues rounds as replays of that graph; 0 launches
 * every round's kernels directly.  The round, its parity and the lanes in
 * use live in device memory, so replays need no new arguments; results are
 * identical either way.  The graph's value is what batch engines created on
 * it afterwards (gas_batch_create, the server's engines) start with; 0 unless
 * set (on one MI355X replays were no faster than direct launches on the road
 * and R-MAT batches).  Replaces the reference's host-driven superstep loop
 * (gpugas.h:764-772) for batches; the reference has no batched engine. */
int gas_graph_batch_graph_rounds(gas_graph_t g, int kind, int32_t *rounds);
int gas_graph_set_batch_graph_rounds(gas_graph_t g, int kind, int32_t rounds);
int gas_batch_set_graph_rounds(gas_batch_t b, int32_t rounds);
int gas_batch_free(gas_batch_t b);

/* ---- continuous-batching query server -------------------------------------
 * Keeps graphs resident and serves BFS / SSSP / personalised-PageRank
 * queries (kinds GAS_BATCH_*) through one batch engine per (kind, graph),
 * each on its own stream.  Queries are admitted into free lanes between
 * rounds while their state fits the memory budget, so work of queries that
 * arrive while others run overlaps on the device.  No counterpart in the
 * reference. */
typedef struct gas_server *gas_server_t;

typedef struct {
  double  elapsed_s;            /* first arrival to last completion window */
  int64_t rounds;               /* host loop iterations that stepped work  */
  int64_t admissions_deferred;  /* times a waiting query did not fit      */
  int32_t max_concurrent;       /* most queries running at once           */
  int32_t solo;                 /* BFS/SSSP queries answered by the plain  */
                                /* engine alone (own engine idle)         */
  int64_t evictions;            /* idle batch engines freed to make room  */
  int64_t device_bytes;         /* device memory the server holds at end  */
} gas_serve_stats;

/* Device bytes of one query's lane state and result (an estimate of the
 * per-query share; admission charges the actual allocations, see below). */
size_t gas_server_query_bytes(int kind, int64_t n_vertices);
/* memory_budget_bytes bounds the device memory the server itself holds:
 * each batch engine's state for its 64 lanes (charged when the engine is
 * built), each lane's result buffer (charged when first used) and the solo
 * threads' result buffers.  A query is admitted only while what it needs
 * fits, evicting batch engines with no running query (least recently used
 * first) to make room; the graphs and their whole-run engines are outside
 * the budget.  0: 80% of the device memory free at creation. */
int gas_server_create(const gas_graph_t *graphs, int32_t n_graphs,
                      size_t memory_budget_bytes, int32_t ppr_iterations,
                      gas_server_t *out);
/* The budget and the device bytes the server holds now. */
int gas_server_budget(gas_server_t s, size_t *budget_bytes, size_t *held_bytes);
/* A new budget (not while serving online); frees the batch engines, which
 * are rebuilt as queries need them. */
int gas_server_set_budget(gas_server_t s, size_t budget_bytes);
/* State bytes of the batch engine of (kind, graph), measured when it was
 * first built (0 before the first replay or start). */
int gas_server_engine_bytes(gas_server_t s, int32_t kind, int32_t graph_id,
                            size_t *bytes);
/* Serves n queries arriving at arrival_s[i] seconds (sorted, relative to
 * the call) in real time.  latency_ms[i] receives arrival-to-result time;
 * host_results[i], if non-NULL, receives the V int32 (BFS, SSSP) or float
 * (PPR) result.  Fails with GAS_ERR_FULL, serving nothing, if some query can
 * never fit the budget.  During the call two more host threads run the
 * graphs' whole-run engines (gas_bfs_device, ...) for batches of one, the
 * second with its own set of them, so the caller must not run whole-run
 * entry points on the same graphs concurrently. */
int gas_server_replay(gas_server_t s, int32_t n, const int32_t *kinds,
                      const int32_t *graph_ids, const int32_t *sources,
                      const double *arrival_s, void *const *host_results,
                      double *latency_ms, gas_serve_stats *stats);
/* Batch admission under load: while at least `backlog` queries wait for a
 * BFS or SSSP engine, they join its running batch only within `rounds`
 * rounds of the batch's start and otherwise wait for it to drain, so a
 * batch's lanes move level by level together and share edge reads.
 * backlog 0 = always admit into free lanes.  Default 16 and 1. */
int gas_server_set_gang(gas_server_t s, int32_t backlog, int32_t rounds);
/* Batches of one: a query whose plain-engine time (measured per kind and
 * graph at the first replay) is at most max_ms, or at most idle_max_ms
 * while the server is lightly loaded, is answered on the plain engine by a
 * solo host thread beside the batch rounds (two under light load, one
 * otherwise).  0, 0 = always batch.
 * Default 2 ms and 50 ms. */
int gas_server_set_solo(gas_server_t s, double max_ms, double idle_max_ms);
/* Online serving: gas_server_start runs the serving loop on a thread of the
 * server's own (after measuring the engines on first use, as the first
 * replay does); gas_server_submit hands it one query, arriving now, and
 * returns a ticket; the result (V int32 for BFS / SSSP, V float for PPR)
 * goes to host_result, if non-NULL, before the ticket completes.
 * gas_server_poll reports a ticket's state (0 waiting or running, 1 done,
 * -1 failed, with the error returned) and latency (arrival to result in
 * HBM); gas_server_wait blocks until it is done or timeout_s passes
 * (GAS_ERR_STATE then).  gas_server_stop stops taking queries, finishes
 * those submitted, joins the thread and returns the session's statistics.
 * Submission fails with GAS_ERR_FULL when the query can never fit the
 * budget.  Replays are refused while the server runs online; the settings
 * (gas_server_set_*) are read by the loop and are set before the start.
 * Thread-safe: any host thread may submit, poll and wait. */
int gas_server_start(gas_server_t s);
int gas_server_submit(gas_server_t s, int32_t kind, int32_t graph_id,
                      int32_t source, void *host_result, int64_t *ticket);
int gas_server_poll(gas_server_t s, int64_t ticket, int32_t *state,
                    double *latency_ms);
int gas_server_wait(gas_server_t s, int64_t ticket, double timeout_s,
                    double *latency_ms);
int gas_server_stop(gas_server_t s, gas_serve_stats *stats);
int gas_server_free(gas_server_t s);

/* ---- the engine API, step by step --------------------------------------- */
#define GAS_PROGRAM_BFS       0  /* VertexData int depth; no EdgeData      */
#define GAS_PROGRAM_SSSP      1  /* VertexData int;  EdgeData int weight   */
#define GAS_PROGRAM_PAGERANK  2  /* VertexData {float rank; int outdeg}    */
#define GAS_PROGRAM_CC        3  /* VertexData int label                   */

typedef struct gas_engine *gas_engine_t;

/* GASEngine*::setGraph (gpugas.h:212, refgas.h:81).  vertex_data is the
 * caller's array of n_vertices program VertexData records, retained and
 * written back by gas_engine_get_results; edge_data (n_edges records in
 * edge-list order, or NULL) is read here only. */
int gas_engine_create(int program, int32_t n_vertices, void *vertex_data,
                      int32_t n_edges, const void *edge_data,
                      const int32_t *srcs, const int32_t *dsts,
                      gas_engine_t *out);
/* Same, on a graph already resident in HBM; edge_data may be NULL to use
 * the graph's own int32 weights. */
int gas_engine_create_on_graph(int program, gas_graph_t g, void *vertex_data,
                               const void *edge_data, gas_engine_t *out);
int gas_engine_set_active(gas_engine_t e, int32_t start, int32_t end);
int gas_engine_count_active(gas_engine_t e, int32_t *count);
int gas_engine_gather_apply(gas_engine_t e, int have_gather);
int gas_engine_scatter_activate(gas_engine_t e, int have_scatter);
int gas_engine_next_iter(gas_engine_t e, int32_t *count);
int gas_engine_run(gas_engine_t e);
int gas_engine_get_results(gas_engine_t e);
/* Superstep counter the BFS program reads as its level (the reference's
 * g_iterationCount, bfs.cu:27-28). */
int gas_engine_set_iteration(gas_engine_t e, int32_t iteration);
int gas_engine_stats(gas_engine_t e, gas_run_stats *stats);
int gas_engine_destroy(gas_engine_t e);

#ifdef __cplusplus
}
#endif

#endif /* GAS_H_ */
