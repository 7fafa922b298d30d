// GASEngineGPU: the reference engine API (gpugas.h:74-773, refgas.h:30-225)
// over gfx950 kernels.
//
// Include this from a hipcc translation unit; the kernels in kernels.hpp are
// instantiated for the Program there, and every host-side device operation
// goes through the gasrt_* C ABI of libgas.so.
//
//   template <class Program, class Int = int32_t> class GASEngineGPU
//     setGraph(nV, vertexData, nE, edgeData, srcs, dsts)   gpugas.h:212
//     getResults()                                         gpugas.h:329
//     setActive(start, end)                                gpugas.h:342
//     countActive()                                        gpugas.h:353
//     gatherApply(haveGather = true)                       gpugas.h:421
//     scatterActivate(haveScatter = true)                  gpugas.h:669
//     nextIter()                                           gpugas.h:754
//     run()                                                gpugas.h:764
//
// Differences from the reference GPU engine, none of which change results:
//  * sizes stay on the device between kernels; the host reads the active
//    count only in countActive()/nextIter() and, in run(), a few supersteps
//    behind the device;
//  * the next active list is unordered in queue mode (every program's
//    result is independent of the order of the active list);
//  * like the reference GPU engine, Program::scatter is not invoked
//    (gpugas.h:666-668).
// Errors throw gas::Error carrying gas_last_error().
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <deque>
#include <stdexcept>
#include <string>
#include <tuple>
#include <type_traits>
#include <utility>
#include <vector>

#include "../gas.h"
#include "kernels.hpp"

namespace gas {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

inline void check(int rc, const char *what) {
  if (rc) throw Error(rc, std::string(what) + ": " + gas_last_error());
}

template <class Tuple, std::size_t... I>
inline void arg_ptrs(Tuple &t, void **argv, std::index_sequence<I...>) {
  ((argv[I] = (void *)&std::get<I>(t)), ...);
}

// Launch a __global__ function through gasrt_launch with arguments
// converted to the kernel's exact parameter types.
template <class... KP, class... A>
inline void launch(void (*k)(KP...), unsigned grid, gasrt_stream_t s,
                   A &&...a) {
  std::tuple<std::decay_t<KP>...> t(std::forward<A>(a)...);
  void *argv[sizeof...(KP) + 1];
  arg_ptrs(t, argv, std::index_sequence_for<KP...>{});
  check(gasrt_launch((const void *)k, grid, NT, argv, 0, s), "launch");
}

// Process-wide switch for per-launch kernel timing (gas_set_kernel_timing).
inline int &kernel_timing_flag() {
  static int v = 0;
  return v;
}

template <class T>
struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  void alloc(size_t count) {
    release();
    check(gasrt_malloc((void **)&p, (count ? count : 1) * sizeof(T)),
          "gasrt_malloc");
    n = count;
  }
  void release() {
    if (p) gasrt_free(p);
    p = nullptr;
    n = 0;
  }
  ~DevBuf() { release(); }
};

template <class Program, class Int = int32_t>
class GASEngineGPU {
  static_assert(std::is_same<Int, int32_t>::value,
                "this engine stores ids as int32");

 public:
  typedef typename Program::VertexData VertexData;
  typedef typename Program::EdgeData EdgeData;
  typedef typename Program::GatherResult GatherResult;
  static constexpr bool kEmptyEdge = std::is_empty<EdgeData>::value;

  GASEngineGPU() {
    check(gasrt_stream_create(&m_stream), "gasrt_stream_create");
    check(gasrt_host_alloc((void **)&m_hctl, sizeof(Ctl) * kRing),
          "gasrt_host_alloc");
    for (int i = 0; i < kRing; ++i)
      check(gasrt_event_create(&m_ring_ev[i]), "gasrt_event_create");
  }

  ~GASEngineGPU() {
    if (m_stream) gasrt_stream_synchronize(m_stream);
    for (auto &pr : m_timing) {
      gasrt_event_destroy(pr.first);
      gasrt_event_destroy(pr.second);
    }
    for (int i = 0; i < kRing; ++i) gasrt_event_destroy(m_ring_ev[i]);
    gasrt_host_free(m_hctl);
    if (m_own_graph) gas_graph_free(m_own_graph);
    gasrt_stream_destroy(m_stream);
  }

  GASEngineGPU(const GASEngineGPU &) = delete;
  GASEngineGPU &operator=(const GASEngineGPU &) = delete;

  // gpugas.h:212-322: build CSC + CSR from the edge list, copy vertex data
  // and edge data (reordered into CSC order) to HBM.
  void setGraph(Int nVertices, VertexData *vertexData, Int nEdges,
                EdgeData *edgeData, const Int *edgeListSrcs,
                const Int *edgeListDsts) {
    gas_edges_t el = nullptr;
    check(gas_edges_from_arrays(nVertices, nEdges, edgeListSrcs,
                                edgeListDsts, nullptr, &el),
          "gas_edges_from_arrays");
    gas_graph_t g = nullptr;
    int rc = gas_graph_create(el, &g);
    gas_edges_free(el);
    check(rc, "gas_graph_create");
    if (m_own_graph) gas_graph_free(m_own_graph);
    m_own_graph = g;
    bindGraph(g, vertexData, edgeData);
  }

  // Binds a graph already resident in HBM (shared between engines).
  // edgeData, if given, is in edge-list order; without it, an int32
  // EdgeData uses the graph's own weights.
  void setGraph(gas_graph_t g, VertexData *vertexData,
                const EdgeData *edgeData) {
    bindGraph(g, vertexData, edgeData);
  }

  // gpugas.h:329-338
  void getResults() {
    if (m_vd_host)
      check(gasrt_memcpy_d2h(m_vd_host, m_vdata.p,
                             sizeof(VertexData) * (size_t)m_nV, m_stream),
            "getResults");
  }

  // Uploads the caller's vertex data again (reuse of a bound graph).
  void resetVertexData(VertexData *vertexData) {
    m_vd_host = vertexData;
    check(gasrt_memcpy_h2d(m_vdata.p, vertexData,
                           sizeof(VertexData) * (size_t)m_nV, m_stream),
          "resetVertexData");
  }

  void setActive(Int vertexStart, Int vertexEnd) {
    if (vertexStart < 0 || vertexEnd > m_nV || vertexEnd < vertexStart)
      throw Error(GAS_ERR_ARG, "setActive: range outside [0, nVertices)");
    unsigned grid = (unsigned)std::min<int64_t>(
        1024, std::max<int64_t>(1, (std::max<int64_t>(m_dev.nwords,
                                                      vertexEnd - vertexStart) +
                                    NT - 1) / NT));
    launch(k_set_active<Program>, grid, m_stream, m_dev, m_par, vertexStart,
           vertexEnd);
  }

  Int countActive() {
    check(gasrt_memcpy_d2h(m_hctl, m_ctl.p, sizeof(Ctl), m_stream),
          "countActive");
    return m_hctl[0].n_active[m_par];
  }

  void gatherApply(bool haveGather = true) {
    if (haveGather) {
      launch(k_gscan1<Program>, SCAN_BLOCKS, m_stream, m_dev, m_par);
      launch(k_gscan2<Program>, SCAN_BLOCKS, m_stream, m_dev, m_par);
      timedBegin();
      launch(k_gather<Program>, MP_BLOCKS, m_stream, m_dev, m_par);
      timedEnd(0);
      launch(k_apply<Program, true>, SCAN_BLOCKS, m_stream, m_dev, m_par);
    } else {
      launch(k_apply<Program, false>, SCAN_BLOCKS, m_stream, m_dev, m_par);
    }
  }

  void scatterActivate(bool /*haveScatter*/ = true) {
    launch(k_sscan2<Program>, SCAN_BLOCKS, m_stream, m_dev, m_par);
    timedBegin();
    launch(k_scatter<Program>, MP_BLOCKS, m_stream, m_dev, m_par);
    timedEnd(1);
    launch(k_compact1<Program>, SCAN_BLOCKS, m_stream, m_dev, m_par);
    launch(k_compact2<Program>, SCAN_BLOCKS, m_stream, m_dev, m_par);
    m_par ^= 1;
  }

  Int nextIter() { return countActive(); }

  void run() { runUntilEmpty(true, true); }

  // Supersteps until the frontier is empty, checking the device count a
  // few supersteps behind so the host never idles the GPU.
  void runUntilEmpty(bool haveGather, bool haveScatter) {
    if (countActive() == 0) return;
    std::deque<int> inflight;
    int j = 0;
    for (;;) {
      gatherApply(haveGather);
      scatterActivate(haveScatter);
      const int slot = j % kRing;
      m_ring_par[slot] = m_par;
      check(gasrt_memcpy_d2h_async(&m_hctl[slot], m_ctl.p, sizeof(Ctl), m_stream),
            "count readback");
      check(gasrt_event_record(m_ring_ev[slot], m_stream), "event");
      inflight.push_back(j++);
      if ((int)inflight.size() >= kDepth) {
        int o = inflight.front();
        inflight.pop_front();
        check(gasrt_event_synchronize(m_ring_ev[o % kRing]), "event sync");
        if (m_hctl[o % kRing].n_active[m_ring_par[o % kRing]] == 0) break;
      }
    }
    check(gasrt_stream_synchronize(m_stream), "sync");
  }

  void setIteration(int it) {
    launch(k_set_iteration<Program>, 1, m_stream, m_dev, it);
  }

  void setDenseThreshold(int32_t edges) {
    m_dense_min = edges;
    Ctl c = readCtl();
    c.dense_min = edges;
    check(gasrt_memcpy_h2d(m_ctl.p, &c, sizeof(Ctl), m_stream), "ctl");
  }

  Ctl readCtl() {
    check(gasrt_memcpy_d2h(m_hctl, m_ctl.p, sizeof(Ctl), m_stream),
          "readCtl");
    return m_hctl[0];
  }

  void resetStats() {
    Ctl c = readCtl();
    c.stat_gathered = c.stat_scattered = c.stat_active = 0;
    c.stat_supersteps = 0;
    check(gasrt_memcpy_h2d(m_ctl.p, &c, sizeof(Ctl), m_stream), "ctl");
    m_gather_ms = m_scatter_ms = 0;
    m_gather_launches = m_scatter_launches = 0;
  }

  // Sum of per-block FP64 residuals of the last apply (programs with a
  // residual hook).
  double lastResidual() {
    std::vector<double> r(SCAN_BLOCKS);
    check(gasrt_memcpy_d2h(r.data(), m_rpart.p, sizeof(double) * r.size(),
                           m_stream),
          "residual");
    double s = 0;
    for (double x : r) s += x;
    return s;
  }

  void stats(gas_run_stats *st) {
    collectTiming();
    Ctl c = readCtl();
    st->supersteps = (int32_t)c.stat_supersteps;
    st->reserved = 0;
    st->gathered_edges = c.stat_gathered;
    st->scattered_edges = c.stat_scattered;
    st->active_total = c.stat_active;
    st->gather_ms = m_gather_ms;
    st->scatter_ms = m_scatter_ms;
    st->gather_launches = m_gather_launches;
    st->scatter_launches = m_scatter_launches;
  }

  gasrt_stream_t stream() const { return m_stream; }
  Int numVertices() const { return m_nV; }
  Int numEdges() const { return m_nE; }
  VertexData *deviceVertexData() { return m_vdata.p; }
  const Dev<Program> &dev() const { return m_dev; }

 private:
  static constexpr int kRing = 8;
  static constexpr int kDepth = 3;

  void bindGraph(gas_graph_t g, VertexData *vertexData,
                 const EdgeData *edgeData) {
    gas_graph_view gv;
    check(gas_graph_view_get(g, &gv), "gas_graph_view_get");
    m_graph = g;
    m_nV = gv.n_vertices;
    m_nE = gv.n_edges;
    m_vd_host = vertexData;
    m_vdata.alloc((size_t)m_nV);
    if (vertexData)
      check(gasrt_memcpy_h2d(m_vdata.p, vertexData,
                             sizeof(VertexData) * (size_t)m_nV, m_stream),
            "vertex data");
    const EdgeData *ed_dev = nullptr;
    if (!kEmptyEdge) {
      if (edgeData) {
        std::vector<EdgeData> sorted((size_t)m_nE);
        check(gas_graph_permute_to_csc(g, edgeData, sorted.data(),
                                       sizeof(EdgeData)),
              "permute edge data");
        m_edata.alloc((size_t)m_nE);
        check(gasrt_memcpy_h2d(m_edata.p, sorted.data(),
                               sizeof(EdgeData) * (size_t)m_nE, m_stream),
              "edge data");
        ed_dev = m_edata.p;
      } else if (sizeof(EdgeData) == 4 && gv.csc_weights) {
        ed_dev = reinterpret_cast<const EdgeData *>(gv.csc_weights);
      } else if (m_nE > 0) {
        throw Error(GAS_ERR_ARG, "setGraph: program needs edge data");
      }
    }
    const size_t V = (size_t)m_nV;
    const int nwords = (int)((V + 31) / 32);
    m_ctl.alloc(1);
    m_q0.alloc(V);
    m_q1.alloc(V);
    m_bm.alloc((size_t)nwords);
    m_goff.alloc(V + 1);
    m_soff.alloc(V + 1);
    m_pdeg.alloc(V);
    m_bsum.alloc(SCAN_BLOCKS);
    m_rpart.alloc(SCAN_BLOCKS);
    m_gtmp.alloc(V);
    const size_t tiles = ((size_t)m_nE + V) / MP_NV + 2;
    m_pfirst.alloc(tiles);
    m_plast.alloc(tiles);
    m_gpart.alloc(tiles);
    m_apart.alloc(tiles);
    m_spart.alloc(tiles);
    check(gasrt_memset(m_bm.p, 0, sizeof(uint32_t) * nwords, m_stream),
          "memset");
    check(gasrt_memset(m_rpart.p, 0, sizeof(double) * SCAN_BLOCKS, m_stream),
          "memset");
    Ctl c;
    memset(&c, 0, sizeof(c));
    c.dense_min = m_dense_min >= 0 ? m_dense_min : std::max(1, m_nV / 64);
    check(gasrt_memcpy_h2d(m_ctl.p, &c, sizeof(Ctl), m_stream), "ctl");

    Dev<Program> &d = m_dev;
    d.nV = m_nV;
    d.nE = m_nE;
    d.csc_off = gv.csc_offsets;
    d.csc_src = gv.csc_srcs;
    d.csr_off = gv.csr_offsets;
    d.csr_dst = gv.csr_dsts;
    d.edata = ed_dev;
    d.vdata = m_vdata.p;
    d.ctl = m_ctl.p;
    d.q[0] = m_q0.p;
    d.q[1] = m_q1.p;
    d.gpart = m_gpart.p;
    d.apart = m_apart.p;
    d.spart = m_spart.p;
    d.bm = m_bm.p;
    d.nwords = nwords;
    d.goff = m_goff.p;
    d.soff = m_soff.p;
    d.pdeg = m_pdeg.p;
    d.bsum = m_bsum.p;
    d.rpart = m_rpart.p;
    d.gtmp = m_gtmp.p;
    d.part_first = m_pfirst.p;
    d.part_last = m_plast.p;
    m_par = 0;
    launch(k_static_partition<Program>,
           (unsigned)std::min<int64_t>(1024, (V + NT - 1) / NT + 1), m_stream,
           m_dev);
  }

  void timedBegin() {
    if (!kernel_timing_flag()) return;
    if (m_tpos == m_timing.size()) {
      gasrt_event_t a, b;
      check(gasrt_event_create(&a), "event");
      check(gasrt_event_create(&b), "event");
      m_timing.push_back({a, b});
      m_tkind.push_back(0);
    }
    check(gasrt_event_record(m_timing[m_tpos].first, m_stream), "event");
  }
  void timedEnd(int kind) {
    if (!kernel_timing_flag()) return;
    check(gasrt_event_record(m_timing[m_tpos].second, m_stream), "event");
    m_tkind[m_tpos] = kind;
    ++m_tpos;
    if (m_tpos >= 4096) collectTiming();
  }
  void collectTiming() {
    if (m_tpos == 0) return;
    check(gasrt_stream_synchronize(m_stream), "sync");
    for (size_t i = 0; i < m_tpos; ++i) {
      float ms = 0;
      check(gasrt_event_elapsed_ms(m_timing[i].first, m_timing[i].second,
                                   &ms),
            "elapsed");
      if (m_tkind[i] == 0) {
        m_gather_ms += ms;
        ++m_gather_launches;
      } else {
        m_scatter_ms += ms;
        ++m_scatter_launches;
      }
    }
    m_tpos = 0;
  }

  gasrt_stream_t m_stream = nullptr;
  Ctl *m_hctl = nullptr;  // pinned ring of control-block snapshots
  gasrt_event_t m_ring_ev[kRing] = {};
  gas_graph_t m_graph = nullptr;
  gas_graph_t m_own_graph = nullptr;
  Int m_nV = 0, m_nE = 0;
  int32_t m_dense_min = -1;
  VertexData *m_vd_host = nullptr;
  DevBuf<VertexData> m_vdata;
  DevBuf<EdgeData> m_edata;
  DevBuf<Ctl> m_ctl;
  DevBuf<int32_t> m_q0, m_q1, m_goff, m_soff, m_pdeg, m_bsum;
  DevBuf<int32_t> m_gpart, m_apart, m_spart;
  int m_par = 0;  // parity of the next superstep
  int m_ring_par[kRing] = {};
  DevBuf<uint32_t> m_bm;
  DevBuf<double> m_rpart;
  DevBuf<GatherResult> m_gtmp;
  DevBuf<Acc<Program>> m_pfirst, m_plast;
  Dev<Program> m_dev{};
  std::vector<std::pair<gasrt_event_t, gasrt_event_t>> m_timing;
  std::vector<int> m_tkind;
  size_t m_tpos = 0;
  double m_gather_ms = 0, m_scatter_ms = 0;
  int m_gather_launches = 0, m_scatter_launches = 0;
};

}  // namespace gas
