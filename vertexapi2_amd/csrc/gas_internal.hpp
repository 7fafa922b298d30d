// Internal declarations shared by the libgas.so translation units.
#pragma once

#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "gas.h"

namespace gas {

int set_error(int code, const char *fmt, ...)
    __attribute__((format(printf, 2, 3)));
void clear_error();
int hip_status(int hip_err, const char *what);

// Returns from the enclosing function with the status of `expr` if nonzero.
#define GAS_TRY(expr)              \
  do {                             \
    int gas_rc_ = (expr);          \
    if (gas_rc_) return gas_rc_;   \
  } while (0)

}  // namespace gas

// Host edge list (owned).
struct gas_edges {
  int32_t nV = 0;
  std::vector<int32_t> src, dst, w;  // w empty when the list has no values
};

// Graph resident in HBM plus the host-side index maps needed to reorder
// per-edge user data.
struct gas_graph {
  int32_t nV = 0, nE = 0;
  std::vector<int32_t> h_csc_off, h_csc_src, h_csc_eid;
  std::vector<int32_t> h_csr_off, h_csr_dst, h_csr_eid;
  int32_t *csc_off = nullptr, *csc_src = nullptr;
  int32_t *csr_off = nullptr, *csr_dst = nullptr;
  int32_t *csc_w = nullptr, *csr_w = nullptr;
  int32_t max_w = 0;  // largest edge weight (0 without weights)
  size_t device_bytes = 0;
  // batch settings new batch engines on this graph start with, by kind:
  // BFS pull divisor, SSSP distance step (gas_graph_autotune_batch)
  int32_t batch_param[2] = {0, 0};
  // hub threshold of those batch engines, in push items (0: the default)
  int32_t batch_hub[2] = {0, 0};
  // rounds per replay of their captured round graph (0: direct launches)
  int32_t batch_graph[2] = {0, 0};
  // out-edges per push item of those batch engines (0: the default)
  int32_t batch_items[2] = {0, 0};
  // Engines kept bound to this graph by the whole-run entry points
  // (gas_bfs, ...), one per program and engine set, released with the
  // graph.  A host thread uses the set engine_set() names, so threads of
  // different sets can run whole-run queries on one graph at once.
  static constexpr int kEngineSets = 2;
  void *engine_cache[kEngineSets][8] = {};
  void (*engine_release[kEngineSets][8])(void *) = {};
  // Copy of the graph with the vertices renumbered by descending in-degree
  // (gas_graph_set_vertex_order), built at first use by PageRank and CC,
  // with each vertex's id in it (new_of) and the original id of each of
  // its vertices (old_of), on the host and the device
  int32_t vertex_order = 1;
  std::mutex order_mu;          // guards the lazy build of the ordered copy
  bool order_unskewed = false;  // checked once: too regular to reorder
  gas_graph *ordered = nullptr;
  std::vector<int32_t> new_of, old_of;
  int32_t *d_new_of = nullptr, *d_old_of = nullptr;
  // out-neighbour table of graphs whose out-degrees are all <= kEllWidth
  // (gas_graph_ell): built at first use, guarded by order_mu
  int32_t ell_state = 0;  // 0 not looked at, 1 built, -1 degrees too large
  int32_t *ell_dst = nullptr, *ell_w = nullptr;
};

namespace gas {

// Engine set of the calling host thread's whole-run calls (0 unless the
// thread chose another; the server's second solo thread uses 1).
int &engine_set();

// The graph PageRank's and CC's whole-run entry points run on: g itself,
// or its degree-ordered copy (built here at first use) when
// g->vertex_order is set, g has at least kOrderEdges edges and its largest
// in-degree is at least kOrderSkew times the mean.
constexpr int64_t kOrderEdges = (int64_t)1 << 20;
constexpr int64_t kOrderSkew = 32;
gas_graph *ordered_graph(gas_graph *g);

// Device-side build of one layout (graph_build.hip): offsets (V+1), values
// and weights (if d_w) in key order, the edge-list index of each slot;
// d_sorted_keys (E) and d_temp (device_build_temp_bytes(E)) are scratch.
// Enqueued on `stream`.
int device_build_layout(int32_t nV, int64_t nE, const int32_t *d_keys,
                        const int32_t *d_vals, const int32_t *d_w,
                        int32_t *d_off, int32_t *d_out_vals, int32_t *d_eid,
                        int32_t *d_out_w, uint32_t *d_sorted_keys, void *d_temp,
                        size_t temp_bytes, gasrt_stream_t stream);
size_t device_build_temp_bytes(int64_t nE);

// Width of the out-neighbour table (gas_graph_ell).
constexpr int kEllWidth = 8;

// Fills the out-neighbour table from the CSR on the device (graph_build.hip):
// row v holds v's out-neighbours (and weights, if d_w) in CSR order, -1
// (and 0) after the last; every out-degree must be <= kEllWidth.
int device_build_ell(int32_t nV, const int32_t *d_off, const int32_t *d_dst,
                     const int32_t *d_w, int32_t *ell_dst, int32_t *ell_w,
                     gasrt_stream_t stream);

// Stable counting sort of edges by key: offsets (V+1), values permuted into
// key order, and the original edge index of each slot.  Equal keys keep
// edge-list order (one valid order of the reference's unstable std::sort,
// util.h:48-72).
void build_compressed(int32_t nV, int32_t nE, const int32_t *keys,
                      const int32_t *vals, std::vector<int32_t> &off,
                      std::vector<int32_t> &out_vals,
                      std::vector<int32_t> &eid);

}  // namespace gas
