// C ABI over GASEngineGPU for the four vertex programs: the step-by-step
// engine API (gas_engine_*) and the app drivers' whole runs (gas_bfs,
// gas_sssp, gas_cc, gas_pagerank).
//
// Whole runs follow the drivers' initialisation and loops:
//   BFS       bfs.cu:91-112, 144-146   depth = -1, setActive(src, src+1),
//                                      gatherApply(false) +
//                                      scatterActivate(false) per level
//   SSSP      sssp.cu:67-82, 141-144   dist = gatherZero, dist[src] = 0,
//                                      setActive(0, V), run()
//   CC        connected_component.cu:66-81, 112-114   label = id, run()
//   PageRank  pagerank.cu:85-101, 126-132   rank = 0.15, numOutEdges from
//                                      CSR, setActive(0, V), run(); or a
//                                      fixed number of all-active sweeps
#include <hip/hip_runtime.h>

#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "gas.h"
#include "gas/engine.hpp"
#include "gas/programs.hpp"
#include "gas_internal.hpp"

namespace gas {

// Sum of the apply kernel's per-chunk FP64 residuals into hist[slot]
// (single workgroup, fixed order).
__global__ __launch_bounds__(NT) void k_residual_sum(const double *rpart,
                                                     double *hist, int slot) {
  __shared__ double lds[NT / WAVE];
  double s = 0;
  for (int i = threadIdx.x; i < SCAN_BLOCKS; i += NT) s += rpart[i];
  s = block_sum(s, lds);
  if (threadIdx.x == 0) hist[slot] = s;
}

// Empty kernel whose grid size carries a tag into profiler traces
// (gas_trace_mark).
__global__ void k_trace_mark() {}

// Type-erased engine for the C ABI.
struct AnyEngine {
  virtual ~AnyEngine() {}
  virtual void setActive(int s, int e) = 0;
  virtual int countActive() = 0;
  virtual void gatherApply(bool g) = 0;
  virtual void scatterActivate(bool s) = 0;
  virtual int nextIter() = 0;
  virtual void run() = 0;
  virtual void getResults() = 0;
  virtual void setIteration(int it) = 0;
  virtual void stats(gas_run_stats *st) = 0;
};

template <class P>
struct EngineImpl : AnyEngine {
  GASEngineGPU<P> e;
  void setActive(int s, int en) override { e.setActive(s, en); }
  int countActive() override { return e.countActive(); }
  void gatherApply(bool g) override { e.gatherApply(g); }
  void scatterActivate(bool s) override { e.scatterActivate(s); }
  int nextIter() override { return e.nextIter(); }
  void run() override { e.run(); }
  void getResults() override { e.getResults(); }
  void setIteration(int it) override { e.setIteration(it); }
  void stats(gas_run_stats *st) override {
    e.stats(st);
    st->ms = 0;
  }
};

template <class F>
static int guarded(F f) {
  try {
    clear_error();
    f();
    return GAS_OK;
  } catch (const Error &x) {
    return set_error(x.code ? x.code : GAS_ERR_HIP, "%s", x.what());
  } catch (const std::bad_alloc &) {
    return set_error(GAS_ERR_NOMEM, "host allocation failed");
  } catch (const std::exception &x) {
    return set_error(GAS_ERR_STATE, "%s", x.what());
  }
}

template <class P>
static GASEngineGPU<P> &cached_engine(gas_graph_t g, int slot,
                                      typename P::VertexData *vd) {
  void *&cache = g->engine_cache[engine_set()][slot];
  if (!cache) {
    auto *e = new GASEngineGPU<P>();
    try {
      e->setGraph(g, vd, nullptr);
    } catch (...) {
      delete e;
      throw;
    }
    cache = e;
    g->engine_release[engine_set()][slot] = [](void *p) {
      delete static_cast<GASEngineGPU<P> *>(p);
    };
  } else {
    if (vd) static_cast<GASEngineGPU<P> *>(cache)->resetVertexData(vd);
  }
  return *static_cast<GASEngineGPU<P> *>(cache);
}

// Times a whole run on the engine's stream with HIP events.
// Device time of a whole run on the engine's stream, with the engine's two
// timing events; stop() waits for the run.
struct RunTimer {
  gasrt_event_t a = nullptr, b = nullptr;
  gasrt_stream_t s;
  template <class E>
  explicit RunTimer(E &e) : s(e.stream()) {
    gasrt_event_t *ev = e.runEvents();
    a = ev[0];
    b = ev[1];
    check(gasrt_event_record(a, s), "event");
  }
  double stop() {
    check(gasrt_event_record(b, s), "event");
    check(gasrt_event_synchronize(b), "event");
    float ms = 0;
    check(gasrt_event_elapsed_ms(a, b, &ms), "elapsed");
    return ms;
  }
};

// Personalised PageRank initial state on the device: rank 0.15 and the
// source flag at `source`, 0 elsewhere, out-degree from the CSR offsets.
__global__ void k_ppr_init(Dev<PersonalizedPageRank> d, int32_t source) {
  for (int64_t v = blockIdx.x * (int64_t)NT + threadIdx.x; v < d.nV;
       v += (int64_t)gridDim.x * NT) {
    PersonalizedPageRank::VertexData x;
    x.rank = v == source ? 0.15f : 0.0f;
    x.numOutEdges = d.csr_off[v + 1] - d.csr_off[v];
    x.isSource = v == source;
    d.vdata[v] = x;
  }
}

// PageRank's initial state (pagerank.cu:89-90, 126-132) written on the device.
__global__ void k_pr_init(Dev<PageRank> d) {
  for (int64_t v = blockIdx.x * (int64_t)NT + threadIdx.x; v < d.nV;
       v += (int64_t)gridDim.x * NT) {
    PageRank::VertexData x;
    x.rank = PageRank::pageConst;
    x.numOutEdges = d.csr_off[v + 1] - d.csr_off[v];
    d.vdata[v] = x;
  }
}

// Ranks in the caller's numbering: out[v] = rank of v's id in the engine's
// graph (new_of = NULL: the same numbering).
__global__ void k_pr_rank_out(Dev<PageRank> d, const int32_t *new_of, float *out) {
  for (int64_t v = blockIdx.x * (int64_t)NT + threadIdx.x; v < d.nV;
       v += (int64_t)gridDim.x * NT)
    out[v] = d.vdata[new_of ? new_of[v] : v].rank;
}

// CC's initial labels, label[v] = v (connected_component.cu:112-114), in
// the caller's numbering: old_of[v] is v's original id (NULL: v itself).
__global__ void k_cc_init(Dev<CC> d, const int32_t *old_of) {
  for (int64_t v = blockIdx.x * (int64_t)NT + threadIdx.x; v < d.nV;
       v += (int64_t)gridDim.x * NT)
    d.vdata[v] = old_of ? old_of[v] : (int)v;
}

// A per-vertex int32 result in the caller's numbering: out[v] = in[new_of[v]].
__global__ void k_gather_result(const int32_t *in, const int32_t *new_of, int32_t *out,
                                int32_t n) {
  for (int64_t v = blockIdx.x * (int64_t)NT + threadIdx.x; v < n;
       v += (int64_t)gridDim.x * NT)
    out[v] = in[new_of[v]];
}

// Copies a per-vertex int32 result of the engine on `og` (g or its
// degree-ordered copy) to the caller's device array, in g's numbering.
static void result_to_caller(gas_graph_t g, gas_graph_t og, const int32_t *in,
                             int32_t *out, gasrt_stream_t s) {
  if (og == g) {
    check(gasrt_memcpy_d2d(out, in, sizeof(int32_t) * (size_t)g->nV, s), "result");
    return;
  }
  const unsigned grid = (unsigned)std::min<int64_t>(
      2048, std::max<int64_t>(1, ((int64_t)g->nV + NT - 1) / NT));
  launch(k_gather_result, grid, s, in, (const int32_t *)g->d_new_of, out, g->nV);
}

// Labels in the caller's numbering: out[v] = label of v's id in the
// engine's graph (new_of = NULL: the same numbering).
__global__ void k_cc_label_out(Dev<CC> d, const int32_t *new_of, int32_t *out) {
  for (int64_t v = blockIdx.x * (int64_t)NT + threadIdx.x; v < d.nV;
       v += (int64_t)gridDim.x * NT)
    out[v] = d.vdata[new_of ? new_of[v] : v];
}

// Ranks in the caller's numbering (new_of = NULL: the same numbering).
__global__ void k_ppr_rank_out(Dev<PersonalizedPageRank> d, const int32_t *new_of,
                               float *out) {
  for (int64_t v = blockIdx.x * (int64_t)NT + threadIdx.x; v < d.nV;
       v += (int64_t)gridDim.x * NT)
    out[v] = d.vdata[new_of ? new_of[v] : v].rank;
}

// Device memory of the allocation-free entry point
// (gas_pagerank_device_noalloc): every buffer its engine needs is carved
// from this array, which each GPU the library runs on holds from module
// load until the process ends.  In host code `scratch` only names the
// symbol; its device address comes from scratch_address().
constexpr size_t kNoAllocScratchBytes = (size_t)64 << 20;
__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

// Device address of `scratch` on the current device, looked up with
// hipGetSymbolAddress once per device (each GPU has its own copy).
static int scratch_address(void **out) {
  static std::mutex mu;
  static void *addr[64] = {};
  int dev = 0;
  GAS_TRY(hip_status(hipGetDevice(&dev), "hipGetDevice"));
  if (dev < 0 || dev >= 64) return set_error(GAS_ERR_ARG, "device %d out of range", dev);
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[dev]) {
    void *p = nullptr;
    GAS_TRY(hip_status(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)),
                       "hipGetSymbolAddress(scratch)"));
    addr[dev] = p;
  }
  *out = addr[dev];
  return GAS_OK;
}

// Upper bound of the scratch the PageRank engine carves for a graph of nV
// vertices and nE edges (every buffer of GASEngineGPU::bindGraph, each
// rounded up to 256 bytes, and the hub lists at their largest).
static size_t noalloc_pagerank_bytes(int64_t nV, int64_t nE) {
  const int64_t tiles = std::max((nE + nV) / MP_NV, nE / SPAN_TE) + 2;
  const int64_t per_vertex = 2 * 8 /* states */ + 4 /* srcval */ + 2 * 4 /* lists */ +
                             4 * 1 /* activation, changed */ + 1 /* apply flags */ +
                             4 /* gather results */ + 4 * 4 /* hub lists, offsets */ +
                             2 * 4 /* all-active hubs, offsets */;
  const int64_t per_tile = 3 * 4 /* partitions */ + 4 /* span tiles */ +
                           2 * (int64_t)sizeof(Acc<PageRank>);
  return (size_t)(per_vertex * (nV + 4) + per_tile * tiles +
                  SCAN_BLOCKS * (4 + 8 * STAT_SLOTS + 8 + 8) + 64 * 256);
}

}  // namespace gas

using namespace gas;

extern "C" {

int gas_set_kernel_timing(int enable) {
  kernel_timing_flag() = enable ? 1 : 0;
  return GAS_OK;
}

int gas_trace_mark(uint32_t tag) {
  if (tag == 0 || tag > 65535) return set_error(GAS_ERR_ARG, "trace tag %u outside 1..65535", tag);
  return guarded([&] {
    launch_wg(k_trace_mark, tag, WAVE, nullptr);
    check(gasrt_stream_synchronize(nullptr), "sync");
    return GAS_OK;
  });
}

}  // extern "C"

// Runs f on the cached engine of (graph, program), creating it with
// placeholder vertex data when no whole run has created it yet.
template <class F>
static int with_engine(gas_graph_t g, int program, F f) {
  if (!g) return set_error(GAS_ERR_ARG, "null graph");
  return guarded([&] {
    switch (program) {
      case GAS_PROGRAM_BFS: {
        gas_graph_t og = ordered_graph(g);
        std::vector<BFS::VertexData> vd(g->nV, BFS::VertexData{-1});
        f(cached_engine<BFS>(og, program, vd.data()));
        static_cast<GASEngineGPU<BFS> *>(og->engine_cache[engine_set()][program])
            ->detachHost();
        break;
      }
      case GAS_PROGRAM_SSSP: {
        if (!g->csc_w && g->nE > 0)
          throw Error(GAS_ERR_ARG, "graph has no edge weights");
        gas_graph_t og = ordered_graph(g);
        std::vector<int> vd(g->nV, SSSP::gatherZero);
        f(cached_engine<SSSP>(og, program, vd.data()));
        static_cast<GASEngineGPU<SSSP> *>(og->engine_cache[engine_set()][program])
            ->detachHost();
        break;
      }
      case GAS_PROGRAM_PAGERANK: {
        gas_graph_t pg = ordered_graph(g);
        std::vector<PageRank::VertexData> vd(pg->nV);
        for (int32_t v = 0; v < pg->nV; ++v)
          vd[v] = {PageRank::pageConst,
                   std::max(1, pg->h_csr_off[v + 1] - pg->h_csr_off[v])};
        f(cached_engine<PageRank>(pg, program, vd.data()));
        static_cast<GASEngineGPU<PageRank> *>(pg->engine_cache[engine_set()][program])
            ->detachHost();
        break;
      }
      case GAS_PROGRAM_CC: {
        gas_graph_t cg = ordered_graph(g);
        std::vector<int> vd(cg->nV);
        for (int32_t v = 0; v < cg->nV; ++v) vd[v] = cg == g ? v : g->old_of[v];
        f(cached_engine<CC>(cg, program, vd.data()));
        static_cast<GASEngineGPU<CC> *>(cg->engine_cache[engine_set()][program])
            ->detachHost();
        break;
      }
      default:
        throw Error(GAS_ERR_ARG, "unknown program");
    }
  });
}

extern "C" {

int gas_graph_set_tiers(gas_graph_t g, int program, const int32_t *tiers) {
  if (!tiers) return set_error(GAS_ERR_ARG, "gas_graph_set_tiers: null");
  return with_engine(g, program, [&](auto &e) {
    e.setThresholds(tiers[0], tiers[1], tiers[2], tiers[3]);
  });
}

int gas_graph_set_frontier_switch(gas_graph_t g, int program,
                                  int32_t divisor) {
  if (divisor < 0)
    return set_error(GAS_ERR_ARG, "gas_graph_set_frontier_switch: need >= 0");
  return with_engine(g, program,
                     [&](auto &e) { e.setSwitchDivisor(divisor); });
}

int gas_graph_set_small_frontier(gas_graph_t g, int program,
                                 int32_t max_vertices) {
  if (max_vertices < 0)
    return set_error(GAS_ERR_ARG, "gas_graph_set_small_frontier: need >= 0");
  return with_engine(g, program,
                     [&](auto &e) { e.setSmallFrontier(max_vertices); });
}

int gas_graph_set_push(gas_graph_t g, int program, int32_t enable) {
  if (enable < 0 || enable > 1)
    return set_error(GAS_ERR_ARG, "gas_graph_set_push: enable must be 0 or 1");
  return with_engine(g, program, [&](auto &e) { e.setPush(enable != 0); });
}

int gas_graph_get_tiers(gas_graph_t g, int program, int32_t *tiers) {
  if (!tiers) return set_error(GAS_ERR_ARG, "gas_graph_get_tiers: null");
  return with_engine(g, program, [&](auto &e) {
    int t[4];
    e.thresholds(t);
    for (int i = 0; i < 4; ++i) tiers[i] = t[i];
  });
}

int gas_graph_autotune(gas_graph_t g, int program, int32_t *tiers_out) {
  return with_engine(g, program, [&](auto &e) {
    e.autotune(program != GAS_PROGRAM_BFS, true);
    if (tiers_out) {
      int t[4];
      e.thresholds(t);
      for (int i = 0; i < 4; ++i) tiers_out[i] = t[i];
    }
  });
}

int gas_bfs(gas_graph_t g, int32_t source, int32_t *depth_out,
            gas_run_stats *stats) {
  if (!g || !depth_out || source < 0 || source >= g->nV)
    return set_error(GAS_ERR_ARG, "gas_bfs: bad arguments");
  return guarded([&] {
    gas_graph_t og = ordered_graph(g);  // g or its degree-ordered copy
    const int32_t src = og == g ? source : g->new_of[source];
    std::vector<BFS::VertexData> vd(g->nV);
    for (auto &x : vd) x.depth = -1;
    auto &e = cached_engine<BFS>(og, GAS_PROGRAM_BFS, vd.data());
    e.resetStats();
    RunTimer tm(e);
    e.setActive(src, src + 1);
    e.setIteration(0);
    e.runUntilEmpty(false, false);
    if (stats) e.enqueueStatTotals();
    double ms = tm.stop();
    e.getResults();
    for (int32_t v = 0; v < g->nV; ++v)
      depth_out[v] = vd[og == g ? v : g->new_of[v]].depth;
    if (stats) {
      e.stats(stats);
      stats->ms = ms;
    }
  });
}

int gas_bfs_device(gas_graph_t g, int32_t source, int32_t *depth_dev,
                   gas_run_stats *stats) {
  if (!g || !depth_dev || source < 0 || source >= g->nV)
    return set_error(GAS_ERR_ARG, "gas_bfs_device: bad arguments");
  return guarded([&] {
    gas_graph_t og = ordered_graph(g);  // g or its degree-ordered copy
    const int32_t src = og == g ? source : g->new_of[source];
    auto &e = cached_engine<BFS>(og, GAS_PROGRAM_BFS, nullptr);
    e.resetStats();
    RunTimer tm(e);
    e.fillVertexData(BFS::VertexData{-1}, -1, BFS::VertexData{-1});
    e.setActive(src, src + 1);
    e.setIteration(0);
    e.runUntilEmpty(false, false);
    result_to_caller(g, og, reinterpret_cast<const int32_t *>(e.deviceVertexData()),
                     depth_dev, e.stream());
    if (stats) e.enqueueStatTotals();
    double ms = tm.stop();
    if (stats) {
      e.stats(stats);
      stats->ms = ms;
    }
  });
}

int gas_sssp(gas_graph_t g, int32_t source, int32_t *dist_out,
             gas_run_stats *stats) {
  if (!g || !dist_out || source < 0 || source >= g->nV)
    return set_error(GAS_ERR_ARG, "gas_sssp: bad arguments");
  if (!g->csc_w && g->nE > 0)
    return set_error(GAS_ERR_ARG, "gas_sssp: graph has no edge weights");
  return guarded([&] {
    gas_graph_t og = ordered_graph(g);  // g or its degree-ordered copy
    std::vector<int> vd(g->nV, SSSP::gatherZero);
    vd[og == g ? source : g->new_of[source]] = 0;
    auto &e = cached_engine<SSSP>(og, GAS_PROGRAM_SSSP, vd.data());
    e.resetStats();
    RunTimer tm(e);
    e.setActive(0, g->nV);
    e.setIteration(0);
    e.runUntilEmpty(true, true);
    if (stats) e.enqueueStatTotals();
    double ms = tm.stop();
    e.getResults();
    for (int32_t v = 0; v < g->nV; ++v) dist_out[v] = vd[og == g ? v : g->new_of[v]];
    if (stats) {
      e.stats(stats);
      stats->ms = ms;
    }
  });
}

int gas_sssp_device(gas_graph_t g, int32_t source, int32_t *dist_dev,
                    gas_run_stats *stats) {
  if (!g || !dist_dev || source < 0 || source >= g->nV)
    return set_error(GAS_ERR_ARG, "gas_sssp_device: bad arguments");
  if (!g->csc_w && g->nE > 0)
    return set_error(GAS_ERR_ARG, "gas_sssp_device: graph has no edge weights");
  return guarded([&] {
    gas_graph_t og = ordered_graph(g);  // g or its degree-ordered copy
    auto &e = cached_engine<SSSP>(og, GAS_PROGRAM_SSSP, nullptr);
    e.resetStats();
    RunTimer tm(e);
    e.fillVertexData(SSSP::gatherZero, og == g ? source : g->new_of[source], 0);
    e.setActive(0, g->nV);
    e.setIteration(0);
    e.runUntilEmpty(true, true);
    result_to_caller(g, og, e.deviceVertexData(), dist_dev, e.stream());
    if (stats) e.enqueueStatTotals();
    double ms = tm.stop();
    if (stats) {
      e.stats(stats);
      stats->ms = ms;
    }
  });
}

int gas_cc(gas_graph_t g, int32_t *label_out, gas_run_stats *stats) {
  if (!g || !label_out) return set_error(GAS_ERR_ARG, "gas_cc: bad arguments");
  return guarded([&] {
    // labels start as the caller's ids, also on the degree-ordered copy
    gas_graph_t cg = ordered_graph(g);
    std::vector<int> vd(cg->nV);
    for (int32_t v = 0; v < cg->nV; ++v) vd[v] = cg == g ? v : g->old_of[v];
    auto &e = cached_engine<CC>(cg, GAS_PROGRAM_CC, vd.data());
    e.resetStats();
    RunTimer tm(e);
    e.setActive(0, g->nV);
    e.setIteration(0);
    e.runUntilEmpty(true, true);
    if (stats) e.enqueueStatTotals();
    double ms = tm.stop();
    e.getResults();
    for (int32_t v = 0; v < g->nV; ++v) label_out[v] = vd[cg == g ? v : g->new_of[v]];
    if (stats) {
      e.stats(stats);
      stats->ms = ms;
    }
  });
}

int gas_pagerank(gas_graph_t g, int32_t iterations, float *rank_out,
                 double *residual_out, gas_run_stats *stats) {
  if (!g || !rank_out || iterations < 0)
    return set_error(GAS_ERR_ARG, "gas_pagerank: bad arguments");
  return guarded([&] {
    gas_graph_t pg = ordered_graph(g);  // g or its degree-ordered copy
    std::vector<PageRank::VertexData> vd(pg->nV);
    for (int32_t v = 0; v < pg->nV; ++v) {
      vd[v].rank = PageRank::pageConst;
      vd[v].numOutEdges = pg->h_csr_off[v + 1] - pg->h_csr_off[v];
    }
    auto &e = cached_engine<PageRank>(pg, GAS_PROGRAM_PAGERANK, vd.data());
    e.resetStats();
    DevBuf<double> hist;
    hist.alloc(iterations > 0 ? (size_t)iterations : 1);
    RunTimer tm(e);
    if (iterations > 0) {
      // every vertex stays active: one setActive serves all sweeps (a
      // gatherApply without a scatter leaves the input as it was)
      e.setActive(0, g->nV);
      for (int it = 0; it < iterations; ++it) {
        e.gatherApply(true);
        launch(k_residual_sum, 1, e.stream(), e.dev().rpart, hist.p, it);
      }
    } else {
      e.setActive(0, g->nV);
      e.setIteration(0);
      e.runUntilEmpty(true, true);
    }
    if (stats) e.enqueueStatTotals();
    double ms = tm.stop();
    e.getResults();
    for (int32_t v = 0; v < g->nV; ++v)
      rank_out[v] = vd[pg == g ? v : g->new_of[v]].rank;
    if (residual_out && iterations > 0)
      check(gasrt_memcpy_d2h(residual_out, hist.p,
                             sizeof(double) * (size_t)iterations, e.stream()),
            "residuals");
    if (stats) {
      e.stats(stats);
      stats->ms = ms;
      if (iterations > 0) {
        stats->supersteps = iterations;
        stats->gathered_edges = (int64_t)iterations * g->nE;
        stats->active_total = (int64_t)iterations * g->nV;
      }
    }
  });
}

int gas_cc_device(gas_graph_t g, int32_t *label_dev, gas_run_stats *stats) {
  if (!g || !label_dev) return set_error(GAS_ERR_ARG, "gas_cc_device: bad arguments");
  return guarded([&] {
    gas_graph_t cg = ordered_graph(g);  // g or its degree-ordered copy
    const bool same = cg == g;
    auto &e = cached_engine<CC>(cg, GAS_PROGRAM_CC, nullptr);
    e.resetStats();
    RunTimer tm(e);
    const unsigned grid = (unsigned)std::min<int64_t>(
        2048, std::max<int64_t>(1, ((int64_t)g->nV + NT - 1) / NT));
    launch(k_cc_init, grid, e.stream(), e.dev(),
           same ? (const int32_t *)nullptr : (const int32_t *)g->d_old_of);
    e.stateChanged();
    e.setActive(0, g->nV);
    e.setIteration(0);
    e.runUntilEmpty(true, true);
    launch(k_cc_label_out, grid, e.stream(), e.dev(),
           same ? (const int32_t *)nullptr : (const int32_t *)g->d_new_of, label_dev);
    if (stats) e.enqueueStatTotals();
    double ms = tm.stop();
    if (stats) {
      e.stats(stats);
      stats->ms = ms;
    }
  });
}

int gas_pagerank_device(gas_graph_t g, int32_t iterations, float *rank_dev,
                        double *residual_dev, gas_run_stats *stats) {
  if (!g || !rank_dev || iterations < 1)
    return set_error(GAS_ERR_ARG, "gas_pagerank_device: bad arguments");
  return guarded([&] {
    gas_graph_t pg = ordered_graph(g);  // g or its degree-ordered copy
    auto &e = cached_engine<PageRank>(pg, GAS_PROGRAM_PAGERANK, nullptr);
    e.resetStats();
    double *res = residual_dev;  // residual sums only when asked for
    RunTimer tm(e);
    const unsigned grid = (unsigned)std::min<int64_t>(
        2048, std::max<int64_t>(1, ((int64_t)g->nV + NT - 1) / NT));
    launch(k_pr_init, grid, e.stream(), e.dev());
    e.stateChanged();
    e.setActive(0, g->nV);  // all sweeps: see gas_pagerank
    for (int it = 0; it < iterations; ++it) {
      e.gatherApply(true);
      if (res) launch(k_residual_sum, 1, e.stream(), e.dev().rpart, res, it);
    }
    launch(k_pr_rank_out, grid, e.stream(), e.dev(),
           pg == g ? (const int32_t *)nullptr : (const int32_t *)g->d_new_of,
           rank_dev);
    if (stats) e.enqueueStatTotals();
    double ms = tm.stop();
    if (stats) {
      e.stats(stats);
      stats->ms = ms;
      stats->supersteps = iterations;
      stats->gathered_edges = (int64_t)iterations * g->nE;
      stats->active_total = (int64_t)iterations * g->nV;
    }
  });
}

int gas_pagerank_device_noalloc(gas_graph_t g, int32_t iterations, float *rank_dev,
                                double *residual_dev, gasrt_stream_t stream) {
  if (!g || !rank_dev || iterations < 1)
    return set_error(GAS_ERR_ARG, "gas_pagerank_device_noalloc: bad arguments");
  return guarded([&] {
    void *base = nullptr;
    check(scratch_address(&base), "scratch");
    // the degree-ordered copy is used only when it already exists (building
    // it allocates)
    gas_graph_t pg = g;
    {
      std::lock_guard<std::mutex> lock(g->order_mu);
      if (g->vertex_order && g->ordered) pg = g->ordered;
    }
    const size_t need = noalloc_pagerank_bytes(pg->nV, pg->nE);
    if (need > kNoAllocScratchBytes)
      throw Error(GAS_ERR_ARG,
                  "gas_pagerank_device_noalloc: a graph of " + std::to_string(pg->nV) +
                      " vertices and " + std::to_string(pg->nE) + " edges needs " +
                      std::to_string(need) + " bytes of scratch; the maximum is " +
                      std::to_string(kNoAllocScratchBytes));
    ExternalMemory ext;
    ext.stream = stream;
    ext.base = static_cast<char *>(base);
    ext.bytes = kNoAllocScratchBytes;
    GASEngineGPU<PageRank> e(ext);
    // the degree tiers of the whole-run engine on this graph, when there is one
    if (void *cached = pg->engine_cache[engine_set()][GAS_PROGRAM_PAGERANK]) {
      int tiers[4];
      static_cast<GASEngineGPU<PageRank> *>(cached)->thresholds(tiers);
      e.setThresholds(tiers[0], tiers[1], tiers[2], tiers[3]);
    }
    e.setGraph(pg, nullptr, nullptr);
    const unsigned grid = (unsigned)std::min<int64_t>(
        2048, std::max<int64_t>(1, ((int64_t)g->nV + NT - 1) / NT));
    launch(k_pr_init, grid, stream, e.dev());
    e.stateChanged();
    e.setActive(0, g->nV);  // all sweeps: see gas_pagerank
    for (int it = 0; it < iterations; ++it) {
      e.gatherApply(true);
      if (residual_dev) launch(k_residual_sum, 1, stream, e.dev().rpart, residual_dev, it);
    }
    launch(k_pr_rank_out, grid, stream, e.dev(),
           pg == g ? (const int32_t *)nullptr : (const int32_t *)g->d_new_of, rank_dev);
    check(gasrt_stream_synchronize(stream), "sync");
  });
}

int gas_ppr(gas_graph_t g, int32_t source, int32_t iterations,
            float *rank_out, gas_run_stats *stats) {
  if (!g || !rank_out || source < 0 || source >= g->nV || iterations < 0)
    return set_error(GAS_ERR_ARG, "gas_ppr: bad arguments");
  return guarded([&] {
    typedef PersonalizedPageRank PPR;
    gas_graph_t og = ordered_graph(g);  // g or its degree-ordered copy
    const int32_t src = og == g ? source : g->new_of[source];
    std::vector<PPR::VertexData> vd(g->nV);
    for (int32_t v = 0; v < g->nV; ++v) {
      vd[v].rank = v == src ? 0.15f : 0.0f;
      vd[v].numOutEdges = og->h_csr_off[v + 1] - og->h_csr_off[v];
      vd[v].isSource = v == src;
    }
    auto &e = cached_engine<PPR>(og, 4, vd.data());
    e.resetStats();
    RunTimer tm(e);
    e.setActive(0, g->nV);  // all sweeps: see gas_pagerank
    for (int it = 0; it < iterations; ++it) e.gatherApply(true);
    if (stats) e.enqueueStatTotals();
    double ms = tm.stop();
    e.getResults();
    for (int32_t v = 0; v < g->nV; ++v) rank_out[v] = vd[og == g ? v : g->new_of[v]].rank;
    if (stats) {
      e.stats(stats);
      stats->ms = ms;
      stats->supersteps = iterations;
      stats->gathered_edges = (int64_t)iterations * g->nE;
    }
  });
}

int gas_ppr_device(gas_graph_t g, int32_t source, int32_t iterations,
                   float *rank_dev, gas_run_stats *stats) {
  if (!g || !rank_dev || source < 0 || source >= g->nV || iterations < 0)
    return set_error(GAS_ERR_ARG, "gas_ppr_device: bad arguments");
  return guarded([&] {
    typedef PersonalizedPageRank PPR;
    gas_graph_t og = ordered_graph(g);  // g or its degree-ordered copy
    auto &e = cached_engine<PPR>(og, 4, nullptr);
    e.resetStats();
    RunTimer tm(e);
    const unsigned grid = (unsigned)std::min<int64_t>(
        2048, std::max<int64_t>(1, ((int64_t)g->nV + NT - 1) / NT));
    launch(k_ppr_init, grid, e.stream(), e.dev(), og == g ? source : g->new_of[source]);
    e.stateChanged();
    e.setActive(0, g->nV);  // all sweeps: see gas_pagerank
    for (int it = 0; it < iterations; ++it) e.gatherApply(true);
    launch(k_ppr_rank_out, grid, e.stream(), e.dev(),
           og == g ? (const int32_t *)nullptr : (const int32_t *)g->d_new_of, rank_dev);
    if (stats) e.enqueueStatTotals();
    double ms = tm.stop();
    if (stats) {
      e.stats(stats);
      stats->ms = ms;
      stats->supersteps = iterations;
      stats->gathered_edges = (int64_t)iterations * g->nE;
    }
  });
}

// ---- step-by-step engine API ------------------------------------------------

struct gas_engine {
  std::unique_ptr<AnyEngine> impl;
};

static AnyEngine *make_engine(int program) {
  switch (program) {
    case GAS_PROGRAM_BFS: return new EngineImpl<BFS>();
    case GAS_PROGRAM_SSSP: return new EngineImpl<SSSP>();
    case GAS_PROGRAM_PAGERANK: return new EngineImpl<PageRank>();
    case GAS_PROGRAM_CC: return new EngineImpl<CC>();
    default: return nullptr;
  }
}

int gas_engine_create(int program, int32_t n_vertices, void *vertex_data,
                      int32_t n_edges, const void *edge_data,
                      const int32_t *srcs, const int32_t *dsts,
                      gas_engine_t *out) {
  if (!out || !vertex_data || n_vertices < 0 || n_edges < 0)
    return set_error(GAS_ERR_ARG, "gas_engine_create: bad arguments");
  if (program < 0 || program > 3)
    return set_error(GAS_ERR_ARG, "gas_engine_create: unknown program %d",
                     program);
  return guarded([&] {
    std::unique_ptr<gas_engine> h(new gas_engine);
    h->impl.reset(make_engine(program));
    switch (program) {
      case GAS_PROGRAM_BFS:
        static_cast<EngineImpl<BFS> *>(h->impl.get())
            ->e.setGraph(n_vertices, (BFS::VertexData *)vertex_data, n_edges,
                         nullptr, srcs, dsts);
        break;
      case GAS_PROGRAM_SSSP:
        if (!edge_data && n_edges > 0)
          throw Error(GAS_ERR_ARG, "SSSP needs edge data (weights)");
        static_cast<EngineImpl<SSSP> *>(h->impl.get())
            ->e.setGraph(n_vertices, (int *)vertex_data, n_edges,
                         (int *)edge_data, srcs, dsts);
        break;
      case GAS_PROGRAM_PAGERANK:
        static_cast<EngineImpl<PageRank> *>(h->impl.get())
            ->e.setGraph(n_vertices, (PageRank::VertexData *)vertex_data,
                         n_edges, nullptr, srcs, dsts);
        break;
      case GAS_PROGRAM_CC:
        static_cast<EngineImpl<CC> *>(h->impl.get())
            ->e.setGraph(n_vertices, (int *)vertex_data, n_edges, nullptr,
                         srcs, dsts);
        break;
    }
    *out = h.release();
  });
}

int gas_engine_create_on_graph(int program, gas_graph_t g, void *vertex_data,
                               const void *edge_data, gas_engine_t *out) {
  if (!out || !g || !vertex_data)
    return set_error(GAS_ERR_ARG, "gas_engine_create_on_graph: bad arguments");
  if (program < 0 || program > 3)
    return set_error(GAS_ERR_ARG, "gas_engine_create_on_graph: program %d",
                     program);
  return guarded([&] {
    std::unique_ptr<gas_engine> h(new gas_engine);
    h->impl.reset(make_engine(program));
    switch (program) {
      case GAS_PROGRAM_BFS:
        static_cast<EngineImpl<BFS> *>(h->impl.get())
            ->e.setGraph(g, (BFS::VertexData *)vertex_data, nullptr);
        break;
      case GAS_PROGRAM_SSSP:
        static_cast<EngineImpl<SSSP> *>(h->impl.get())
            ->e.setGraph(g, (int *)vertex_data, (const int *)edge_data);
        break;
      case GAS_PROGRAM_PAGERANK:
        static_cast<EngineImpl<PageRank> *>(h->impl.get())
            ->e.setGraph(g, (PageRank::VertexData *)vertex_data, nullptr);
        break;
      case GAS_PROGRAM_CC:
        static_cast<EngineImpl<CC> *>(h->impl.get())
            ->e.setGraph(g, (int *)vertex_data, nullptr);
        break;
    }
    *out = h.release();
  });
}

#define ENGINE_OR_FAIL(e)                                            \
  if (!(e) || !(e)->impl)                                            \
  return set_error(GAS_ERR_ARG, "%s: null engine", __func__)

int gas_engine_set_active(gas_engine_t e, int32_t start, int32_t end) {
  ENGINE_OR_FAIL(e);
  return guarded([&] { e->impl->setActive(start, end); });
}

int gas_engine_count_active(gas_engine_t e, int32_t *count) {
  ENGINE_OR_FAIL(e);
  if (!count) return set_error(GAS_ERR_ARG, "gas_engine_count_active: null");
  return guarded([&] { *count = e->impl->countActive(); });
}

int gas_engine_gather_apply(gas_engine_t e, int have_gather) {
  ENGINE_OR_FAIL(e);
  return guarded([&] { e->impl->gatherApply(have_gather != 0); });
}

int gas_engine_scatter_activate(gas_engine_t e, int have_scatter) {
  ENGINE_OR_FAIL(e);
  return guarded([&] { e->impl->scatterActivate(have_scatter != 0); });
}

int gas_engine_next_iter(gas_engine_t e, int32_t *count) {
  ENGINE_OR_FAIL(e);
  return guarded([&] {
    int n = e->impl->nextIter();
    if (count) *count = n;
  });
}

int gas_engine_run(gas_engine_t e) {
  ENGINE_OR_FAIL(e);
  return guarded([&] { e->impl->run(); });
}

int gas_engine_get_results(gas_engine_t e) {
  ENGINE_OR_FAIL(e);
  return guarded([&] { e->impl->getResults(); });
}

int gas_engine_set_iteration(gas_engine_t e, int32_t iteration) {
  ENGINE_OR_FAIL(e);
  return guarded([&] { e->impl->setIteration(iteration); });
}

int gas_engine_stats(gas_engine_t e, gas_run_stats *stats) {
  ENGINE_OR_FAIL(e);
  if (!stats) return set_error(GAS_ERR_ARG, "gas_engine_stats: null");
  return guarded([&] { e->impl->stats(stats); });
}

int gas_engine_destroy(gas_engine_t e) {
  delete e;
  return GAS_OK;
}

}  // extern "C"
