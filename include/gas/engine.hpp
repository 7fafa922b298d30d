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
//  * run() executes apply and the activation of out-neighbours in one
//    kernel (both are per active vertex and nothing reads the activation
//    flags before the superstep ends);
//  * like the reference GPU engine, Program::scatter is not invoked
//    (gpugas.h:666-668) -- unless the program declares
//    `static constexpr bool scatterEdges = true`: then scatterActivate(true)
//    and run() call it for every out-edge of each vertex whose apply
//    returned true, as the reference CPU engine does (refgas.h:170-194), and
//    getResults() also writes the edge data back to the caller's array.
// Errors throw gas::Error carrying gas_last_error().
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
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

// The same with an explicit workgroup size.
template <class... KP, class... A>
inline void launch_wg(void (*k)(KP...), unsigned grid, unsigned block,
                      gasrt_stream_t s, A &&...a) {
  std::tuple<std::decay_t<KP>...> t(std::forward<A>(a)...);
  void *argv[sizeof...(KP) + 1];
  arg_ptrs(t, argv, std::index_sequence_for<KP...>{});
  check(gasrt_launch((const void *)k, grid, block, argv, 0, s), "launch");
}

// Process-wide switch for per-launch kernel timing (gas_set_kernel_timing).
inline int &kernel_timing_flag() {
  static int v = 0;
  return v;
}

// Device memory an engine may use instead of allocating: its buffers are
// carved in order from [base, base + bytes) and its kernels run on the
// caller's stream (the engine then allocates nothing, device or pinned).
struct ExternalMemory {
  gasrt_stream_t stream = nullptr;
  char *base = nullptr;
  size_t bytes = 0;
  size_t used = 0;
  // next 256-byte aligned piece of `bytes`
  void *take(size_t want) {
    const size_t b = ((want ? want : 1) + 255) & ~(size_t)255;
    if (used + b > bytes)
      throw Error(GAS_ERR_ARG, "engine needs more than the " +
                                   std::to_string(bytes) + " bytes of external memory");
    void *p = base + used;
    used += b;
    return p;
  }
};

template <class T>
struct DevBuf {
  T *p = nullptr;
  size_t n = 0;
  bool owned = false;
  // from the device allocator, or from ext when given
  void alloc(size_t count, ExternalMemory *ext = nullptr) {
    release();
    if (ext) {
      p = static_cast<T *>(ext->take((count ? count : 1) * sizeof(T)));
      n = count;
      return;
    }
    check(gasrt_malloc((void **)&p, (count ? count : 1) * sizeof(T)),
          "gasrt_malloc");
    owned = true;
    n = count;
  }
  void upload(const std::vector<T> &h, gasrt_stream_t s, ExternalMemory *ext = nullptr) {
    alloc(h.size(), ext);
    if (!h.empty())
      check(gasrt_memcpy_h2d(p, h.data(), sizeof(T) * h.size(), s), "upload");
  }
  void release() {
    if (p && owned) gasrt_free(p);
    p = nullptr;
    n = 0;
    owned = false;
  }
  size_t bytes() const { return p ? (n ? n : 1) * sizeof(T) : 0; }
  ~DevBuf() { release(); }
};

// Merge-path tile partition of a segment list given its exclusive scan
// (host version of mark_tiles, for lists fixed per graph).
inline std::vector<int32_t> host_tile_partition(const std::vector<int32_t> &off) {
  const int64_t n = (int64_t)off.size() - 1;
  const int64_t N = (n > 0 ? off[n] : 0) + n;
  const int64_t ntiles = (N + MP_NV - 1) / MP_NV;
  std::vector<int32_t> part((size_t)ntiles + 1, 0);
  for (int64_t h = 0; h < n; ++h) {
    const int64_t start = off[h] + h, end = off[h + 1] + h;
    for (int64_t t = (start + MP_NV - 1) / MP_NV; t * MP_NV <= end; ++t)
      part[t] = (int32_t)h;
  }
  part[ntiles] = (int32_t)n;
  return part;
}

template <class Program, class Int = int32_t>
class GASEngineGPU {
  static_assert(std::is_same<Int, int32_t>::value,
                "this engine stores ids as int32");

 public:
  typedef typename Program::VertexData VertexData;
  typedef typename Program::EdgeData EdgeData;
  typedef typename Program::GatherResult GatherResult;
  static constexpr bool kEmptyEdge = std::is_empty<EdgeData>::value;

  // An engine that allocates nothing: device buffers from ext (which must
  // outlive it), kernels on ext.stream, no host status ring (countActive and
  // whole runs to an empty frontier are not available; fixed supersteps
  // from setActive are), no kernel timing.
  explicit GASEngineGPU(const ExternalMemory &ext) : m_ext_mem(ext) {
    m_ext = &m_ext_mem;
    m_stream = ext.stream;
    m_dev.status = static_cast<int64_t *>(m_ext->take(sizeof(int64_t) * STATUS_SLOTS));
  }

  GASEngineGPU() {
    check(gasrt_stream_create(&m_stream), "gasrt_stream_create");
    check(gasrt_host_alloc((void **)&m_hctl, sizeof(Ctl)), "gasrt_host_alloc");
    void *dev_status = nullptr;
    check(gasrt_host_alloc_mapped((void **)&m_status, &dev_status,
                                  sizeof(int64_t) * (STATUS_SLOTS + STAT_SLOTS)),
          "gasrt_host_alloc_mapped");
    m_dev.status = static_cast<int64_t *>(dev_status);
    m_dev_totals = m_dev.status + STATUS_SLOTS;
  }

  ~GASEngineGPU() {
    if (m_ext) return;  // nothing of its own: buffers and stream are the caller's
    if (m_stream) gasrt_stream_synchronize(m_stream);
    for (auto &pr : m_timing) {
      gasrt_event_destroy(pr.first);
      gasrt_event_destroy(pr.second);
    }
    for (gasrt_event_t ev : m_run_ev)
      if (ev) gasrt_event_destroy(ev);
    gasrt_host_free(m_hctl);
    gasrt_host_free((void *)m_status);
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
    m_ed_host = edgeData;
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
      check(gasrt_memcpy_d2h(m_vd_host, m_dev.vdata,
                             sizeof(VertexData) * (size_t)m_nV, m_stream),
            "getResults");
    if constexpr (scatters_edges<Program>::value && !kEmptyEdge) {
      // edge data as the scatters left it, back in edge-list order
      if (m_ed_host && m_nE > 0) {
        std::vector<EdgeData> csc((size_t)m_nE);
        check(gasrt_memcpy_d2h(csc.data(), m_edata.p,
                               sizeof(EdgeData) * (size_t)m_nE, m_stream),
              "getResults edges");
        for (Int k = 0; k < m_nE; ++k) m_ed_host[m_csc_eid[(size_t)k]] = csc[(size_t)k];
      }
    }
  }

  // Forgets the caller's vertex-data array (getResults becomes a no-op
  // until resetVertexData binds a new one).
  void detachHost() { m_vd_host = nullptr; }

  // Uploads the caller's vertex data again (reuse of a bound graph).
  void resetVertexData(VertexData *vertexData) {
    m_vd_host = vertexData;
    check(gasrt_memcpy_h2d(m_dev.vdata, vertexData,
                           sizeof(VertexData) * (size_t)m_nV, m_stream),
          "resetVertexData");
    m_other_stale = true;
    m_srcval_stale = true;
  }

  // Sets every vertex to `value` and vertex `at` to `one` on the device
  // (at < 0: no exception); unbinds any host vertex-data array.
  void fillVertexData(const VertexData &value, Int at, const VertexData &one) {
    if (at >= m_nV) throw Error(GAS_ERR_ARG, "fillVertexData: vertex out of range");
    m_vd_host = nullptr;
    const unsigned grid = (unsigned)std::min<int64_t>(
        2048, std::max<int64_t>(1, ((int64_t)m_nV + NT - 1) / NT));
    // a min-program's run may push from its first superstep: both state
    // buffers start equal
    VertexData *also = is_push_min<Program>::value ? m_dev.vnext : nullptr;
    launch(k_fill_vdata<Program>, grid, m_stream, m_dev, value, (int32_t)at,
           one, also);
    m_other_stale = also == nullptr;
    m_srcval_stale = true;
  }

  // The caller wrote vertex state through deviceVertexData().
  void stateChanged() {
    m_other_stale = true;
    m_srcval_stale = true;
  }

  void setActive(Int vertexStart, Int vertexEnd) {
    if (vertexStart < 0 || vertexEnd > m_nV || vertexEnd < vertexStart)
      throw Error(GAS_ERR_ARG, "setActive: range outside [0, nVertices)");
    // all active: only the counts are written (one workgroup), no list
    const bool all = vertexStart == 0 && vertexEnd == m_nV;
    const int64_t work = std::max<int64_t>(m_flags_dirty ? m_dev.nfwords : 0,
                                           all ? 0 : (int64_t)vertexEnd - vertexStart);
    const unsigned grid =
        (unsigned)std::min<int64_t>(1024, std::max<int64_t>(1, (work + NT - 1) / NT));
    m_dev.seq = (int32_t)m_seq;
    launch(k_set_active<Program>, grid, m_stream, m_dev, m_par, vertexStart,
           vertexEnd, m_flags_dirty ? 1 : 0);
    m_flags_dirty = false;
    m_in = (vertexStart == 0 && vertexEnd == m_nV) ? IN_ALL : IN_LIST;
    m_prev_in = IN_LIST;  // nothing of an earlier superstep left to report
    m_first = true;
    m_host_active = (int64_t)vertexEnd - vertexStart;
  }

  Int countActive() {
    if (m_ext) throw Error(GAS_ERR_STATE, "countActive: engine without host memory");
    if (m_in == IN_BITMAP)
      throw Error(GAS_ERR_STATE, "countActive: frontier held as a bitmap");
    check(gasrt_memcpy_d2h(m_hctl, m_ctl.p, sizeof(Ctl), m_stream),
          "countActive");
    return m_hctl->ring[m_seq & 3].n_active;
  }

  void gatherApply(bool haveGather = true) {
    m_out_list = 0;
    m_other_stale = true;
    if (haveGather) {
      gatherPhase();
      klaunch(GAS_K_APPLY, 1, k_apply<Program, true, false>, SCAN_BLOCKS);
    } else {
      klaunch(GAS_K_APPLY, 1, k_apply<Program, false, false>, SCAN_BLOCKS);
    }
  }

  // Scatter of the step-by-step API: plain activation bytes, then the
  // ordered compaction, so the next frontier is an ascending list.
  void scatterActivate(bool haveScatter = true) {
    m_out_list = 0;
    if constexpr (scatters_edges<Program>::value && !kEmptyEdge)
      if (haveScatter && m_nE > 0)
        klaunch(GAS_K_SCATTER, 1, k_edge_scatter<Program>, SCAN_BLOCKS);
    klaunch(GAS_K_SCATTER, 1, k_scatter_vp<Program>, SCAN_BLOCKS);
    endSuperstep(true);
  }

  Int nextIter() { return countActive(); }

  // refgas.h:215-222; programs that scatter to edges take the step-by-step
  // supersteps (every scatter must see every apply of its superstep)
  void run() {
    if constexpr (scatters_edges<Program>::value && !kEmptyEdge) {
      while (countActive()) {
        gatherApply(true);
        scatterActivate(true);
        nextIter();
      }
    } else {
      runUntilEmpty(true, true);
    }
  }

  // One whole superstep: gatherApply + scatterActivate, with the thread and
  // wave tiers of the activation fused into the apply kernel.  outList:
  // the scatter builds the next frontier as a list (claims + appends),
  // otherwise as activation bytes only (bitmap).
  void superstep(bool haveGather, bool outList) {
    m_out_list = outList ? 1 : 0;
    if (m_push_run && haveGather && m_in != IN_LIST && !outList) {
      // min-program: the changed vertices push along their out-edges into
      // the other state buffer (k_push; same values as the pull)
      if (m_other_stale)
        check(gasrt_memcpy_d2d(m_dev.vnext, m_dev.vdata,
                               sizeof(VertexData) * (size_t)m_nV, m_stream),
              "state copy");
      m_other_stale = false;
      klaunch(GAS_K_PUSH, 0, k_push<Program>, SCAN_BLOCKS);
      std::swap(m_dev.vdata, m_dev.vnext);
      endSuperstep(false, false);
      return;
    }
    if (haveGather && !inHubs() && m_in != IN_LIST &&
        !is_source_only<Program>::value) {
      // gather, apply and scatter in one launch, state double-buffered
      if (m_other_stale)
        check(gasrt_memcpy_d2d(m_dev.vnext, m_dev.vdata,
                               sizeof(VertexData) * (size_t)m_nV, m_stream),
              "state copy");
      m_other_stale = false;
      klaunch(GAS_K_FUSED, 0, k_fused<Program>, SCAN_BLOCKS);
      std::swap(m_dev.vdata, m_dev.vnext);
      endSuperstep(false);
      return;
    }
    m_other_stale = true;
    if (haveGather) gatherPhase();
    if (haveGather)
      klaunch(GAS_K_APPLY, 1, k_apply<Program, true, true>, SCAN_BLOCKS);
    else
      klaunch(GAS_K_APPLY, 1, k_apply<Program, false, true>, SCAN_BLOCKS);
    endSuperstep(false);
  }

  // Supersteps until the frontier is empty.  The host enqueues supersteps
  // without waiting; every superstep's frontier size reaches the mapped
  // status ring (list and all inputs from their own first kernel, bitmap
  // inputs from the next superstep's), and the host consumes the reports in
  // order, staying at most kDepth + 1 supersteps ahead of the last one.  The
  // scatter of each superstep builds a list when the last reported frontier
  // is sparse (|A| * switch < V), a bitmap otherwise.  The run ends at the
  // first reported empty frontier; supersteps enqueued after it are no-ops.
  // On graphs whose degrees are all small, a list frontier of at most
  // m_small_max vertices is handed to k_small, which runs supersteps in one
  // workgroup until the frontier outgrows it or empties.
  void runUntilEmpty(bool haveGather, bool /*haveScatter*/) {
    // exact size of a list / all input: known on the host right after
    // setActive, read back otherwise
    int64_t known = m_host_active >= 0 ? m_host_active : countActive();
    if (known == 0) return;
    // push supersteps for the whole run, or none: the push frontier is the
    // set of vertices that changed, so the run must start from every vertex
    // (setActive(0, V)), not from a list the pull's semantics defined
    m_push_run = haveGather && m_in == IN_ALL && pushUsable();
    struct PushReset {
      bool &f;
      ~PushReset() { f = false; }
    } push_reset{m_push_run};
    const bool small_ok = m_small_max > 0 && m_max_in <= SMALL_DEG_MAX &&
                          m_max_out <= SMALL_DEG_MAX;
    const int64_t s0 = m_seq;
    const int it0 = m_iter;
    int64_t next = s0, last_a = known, empty_at = -1;
    auto reset_status = [&] {
      for (int i = 0; i < STATUS_SLOTS; ++i)
        __atomic_store_n(m_status + i, (int64_t)0, __ATOMIC_RELEASE);
      next = m_seq;
    };
    reset_status();
    // consumes every report that has arrived, in superstep order
    auto poll = [&]() {
      bool got = false;
      while (empty_at < 0) {
        const int64_t v = __atomic_load_n(m_status + (next & (STATUS_SLOTS - 1)),
                                          __ATOMIC_ACQUIRE);
        if ((v >> 32) != next + 1) break;
        last_a = v & 0xffffffffll;
        if (last_a == 0) empty_at = next;
        ++next;
        got = true;
      }
      return got;
    };
    bool drain = false;  // the superstep just enqueued builds a list: sync
    while (empty_at < 0) {
      if (small_ok && m_in == IN_LIST && known >= 0 && known <= m_small_max) {
        smallSteps(haveGather);
        const Ctl c = readCtl();
        if (c.small_empty) {
          empty_at = m_seq;
          break;
        }
        known = c.small_n;
        last_a = known;
        reset_status();
      }
      const bool want_list =
          (m_switch > 0 && last_a * (int64_t)m_switch < (int64_t)m_nV) ||
          (small_ok && last_a <= m_small_max / 2);
      superstep(haveGather, want_list);
      known = -1;
      poll();
      if (empty_at >= 0) break;
      if (small_ok && want_list && last_a <= m_small_max / 2) drain = true;
      if (drain) {
        // the list just built has an exact size once the stream is idle
        check(gasrt_stream_synchronize(m_stream), "sync");
        poll();
        if (empty_at >= 0) break;
        if (next < m_seq && bitmapInputSize(m_par ^ 1) == 0) {
          // the last superstep's input was an (unreported) empty bitmap
          empty_at = m_seq - 1;
          break;
        }
        known = readCtl().ring[m_seq & 3].n_active;
        if (known == 0) {
          empty_at = m_seq;
          break;
        }
        drain = false;
        reset_status();
        continue;
      }
      for (int64_t spin = 0; empty_at < 0 && m_seq - next > kDepth + 1; ++spin) {
        if (poll()) continue;
        if ((spin & 255) == 255) {
          int idle = 0;
          check(gasrt_stream_query(m_stream, &idle), "stream query");
          if (idle && !poll())
            throw Error(GAS_ERR_STATE, "superstep status was not reported");
        }
        __builtin_ia32_pause();
      }
    }
    // the supersteps enqueued past the empty one find empty frontiers and
    // change nothing; later work on the stream runs after them
    const int64_t real = empty_at - s0;  // supersteps with a non-empty frontier
    m_supersteps -= (m_seq - s0) - real;
    m_iter = it0 + (int)real;
    setActive(0, 0);
  }

  // Push supersteps for pushMin programs in whole runs (default on): used
  // when every out-degree is at most WAVE, the frontier stays a bitmap
  // (switch divisor 0) and the small-frontier mode cannot engage; the
  // pull runs otherwise.  Values and superstep counts are the same either
  // way; the run statistics count pushes (scattered), pushing vertices
  // (active) and changed vertices (returned) instead.
  void setPush(bool on) { m_push = on; }
  bool pushUsable() const {
    if constexpr (!is_push_min<Program>::value) {
      return false;
    } else {
      const bool small_ok = m_small_max > 0 && m_max_in <= SMALL_DEG_MAX &&
                            m_max_out <= SMALL_DEG_MAX;
      return m_push && m_max_out <= WAVE && m_switch == 0 && !small_ok &&
             (kEmptyEdge || m_dev.edata_csr != nullptr);
    }
  }

  // Lanes per frontier vertex in k_small_lds: the largest out-degree
  // rounded up to a power of two (at most the wave).
  int smallLdsGroup() const {
    int g = 1;
    while (g < m_max_out && g < WAVE) g <<= 1;
    return g;
  }

  // Gather-free programs on graphs of at most SMALL_LDS_V vertices and
  // SMALL_LDS_E edges run the small-frontier supersteps with the graph in
  // LDS (k_small_lds; default on).
  void setSmallLds(bool on) { m_small_lds = on; }

  // Largest list frontier handed to k_small (0: never).
  void setSmallFrontier(int max_vertices) {
    if (max_vertices < 0) throw Error(GAS_ERR_ARG, "small frontier must be >= 0");
    m_small_max = max_vertices;
  }

  void setIteration(int it) { m_iter = it; }

  // Frontier density switch: the scatter builds a list while
  // |A| * divisor < V, a bitmap otherwise; 0 = always a bitmap.
  void setSwitchDivisor(int divisor) {
    if (divisor < 0) throw Error(GAS_ERR_ARG, "switch divisor must be >= 0");
    m_switch = divisor;
  }
  int switchDivisor() const { return m_switch; }

  // Degree tiers of the gather (g) and the scatter (s):
  // thread <= t1 < wave <= t2 < merge-path hub.
  void setThresholds(int gt1, int gt2, int st1, int st2) {
    if (gt1 < 0 || gt1 > T1_MAX || gt2 < gt1 || gt2 < HUB_MIN_T2 ||
        st1 < 0 || st1 > T1_MAX || st2 < st1 || st2 < HUB_MIN_T2)
      throw Error(GAS_ERR_ARG,
                  "setThresholds: need 0 <= t1 <= 16, t1 <= t2, t2 >= 4");
    m_gt1 = gt1;
    m_gt2 = gt2;
    m_st1 = st1;
    m_st2 = st2;
    if (m_graph) buildHubs();
  }
  void thresholds(int *t) const {
    t[0] = m_gt1;
    t[1] = m_gt2;
    t[2] = m_st1;
    t[3] = m_st2;
  }

  // Autotuner: times the all-active gather phase and an all-returning
  // scatter phase (neither changes vertex data) for candidate tiers on
  // this graph and keeps the fastest.  Leaves the frontier empty.
  void autotune(bool tuneGather, bool tuneScatter, int reps = 3) {
    static const int kT1[] = {4, 8, 16};
    static const int kT2[] = {0, 32, 128, 512, 2048, 1 << 30};
    gasrt_event_t a, b;
    check(gasrt_event_create(&a), "event");
    check(gasrt_event_create(&b), "event");
    // the scatter is timed as supersteps run it, fused into the apply pass
    // over every vertex (from the state saved here, restored before each
    // repetition and at the end; the gather results of one gather pass),
    // then the hub scatter; engines on caller memory time the unfused pass
    DevBuf<VertexData> saved;
    const bool fused = tuneScatter && !m_ext;
    const size_t state_bytes = sizeof(VertexData) * (size_t)m_nV;
    auto restore = [&]() {
      check(gasrt_memcpy_d2d(m_dev.vdata, saved.p, state_bytes, m_stream), "restore");
    };
    auto time_phase = [&](bool gather) {
      float best = 1e30f;
      for (int r = 0; r < reps; ++r) {
        setActive(0, m_nV);
        if (!gather && fused) restore();
        if (!gather && !fused)
          check(gasrt_memset(m_rflag.p, 1, (size_t)m_nV, m_stream), "memset");
        check(gasrt_event_record(a, m_stream), "event");
        if (gather) {
          gatherPhase();
        } else if (fused) {
          m_out_list = 0;
          m_flags_dirty = true;
          if (tuneGather)
            klaunch(GAS_K_APPLY, 1, k_apply<Program, true, true>, SCAN_BLOCKS);
          else
            klaunch(GAS_K_APPLY, 1, k_apply<Program, false, true>, SCAN_BLOCKS);
          if (outHubs()) {
            klaunch(GAS_K_HSCAN1, 1, k_hscan1<Program, 1>, SCAN_BLOCKS);
            klaunch(GAS_K_HSCAN2, 1, k_hscan2<Program, 1>, SCAN_BLOCKS);
            klaunch(GAS_K_HSCATTER, 1, k_hscatter<Program>, MP_BLOCKS);
          }
        } else {
          m_out_list = 0;
          m_flags_dirty = true;
          klaunch(GAS_K_SCATTER, 1, k_scatter_vp<Program>, SCAN_BLOCKS);
          if (outHubs()) {
            klaunch(GAS_K_HSCAN1, 1, k_hscan1<Program, 1>, SCAN_BLOCKS);
            klaunch(GAS_K_HSCAN2, 1, k_hscan2<Program, 1>, SCAN_BLOCKS);
            klaunch(GAS_K_HSCATTER, 1, k_hscatter<Program>, MP_BLOCKS);
          }
        }
        check(gasrt_event_record(b, m_stream), "event");
        check(gasrt_event_synchronize(b), "event");
        float ms = 0;
        check(gasrt_event_elapsed_ms(a, b, &ms), "elapsed");
        best = std::min(best, ms);
      }
      return best;
    };
    auto sweep = [&](bool gather) {
      int bt1 = gather ? m_gt1 : m_st1, bt2 = gather ? m_gt2 : m_st2;
      float best = 1e30f;
      const int maxdeg = gather ? m_max_in : m_max_out;
      for (int t1 : kT1)
        for (int t2r : kT2) {
          const int t2 = std::max(t1, std::min(t2r, std::max(maxdeg, t1)));
          if (gather)
            setThresholds(t1, t2, m_st1, m_st2);
          else
            setThresholds(m_gt1, m_gt2, t1, t2);
          const float ms = time_phase(gather);
          if (ms < best) {
            best = ms;
            bt1 = t1;
            bt2 = t2;
          }
        }
      if (gather)
        setThresholds(bt1, bt2, m_st1, m_st2);
      else
        setThresholds(m_gt1, m_gt2, bt1, bt2);
    };
    if (tuneGather) sweep(true);
    if (fused) {
      saved.alloc((size_t)m_nV);
      check(gasrt_memcpy_d2d(saved.p, m_dev.vdata, state_bytes, m_stream), "save");
      if (tuneGather) {
        setActive(0, m_nV);
        gatherPhase();  // the apply passes' gather results
      }
    }
    if (tuneScatter) sweep(false);
    if (fused) {
      restore();
      m_srcval_stale = true;
      check(gasrt_stream_synchronize(m_stream), "sync");
    }
    gasrt_event_destroy(a);
    gasrt_event_destroy(b);
    setActive(0, 0);
    setIteration(0);
    resetStats();
  }

  Ctl readCtl() {
    if (m_ext) throw Error(GAS_ERR_STATE, "readCtl: engine without host memory");
    check(gasrt_memcpy_d2h(m_hctl, m_ctl.p, sizeof(Ctl), m_stream),
          "readCtl");
    return *m_hctl;
  }

  void resetStats() {
    collectTiming();
    zero(m_stat.p, sizeof(int64_t) * STAT_SLOTS * SCAN_BLOCKS);
    m_supersteps = 0;
    m_gather_ms = m_scatter_ms = 0;
    for (int i = 0; i < GAS_K_SLOTS; ++i) {
      m_kernel_ms[i] = 0;
      m_kernel_launches[i] = 0;
    }
  }

  // Enqueues the sum of the run counters into host-mapped memory, read by
  // the next stats() (enqueue it before the caller's own wait, and stats()
  // needs no copy of its own).
  void enqueueStatTotals() {
    if (m_ext) return;
    launch(k_stat_totals<Program>, 1, m_stream, m_dev, m_dev_totals);
    m_totals_enqueued = true;
  }

  void stats(gas_run_stats *st) {
    collectTiming();
    int64_t tot[STAT_SLOTS] = {};
    if (!m_ext) {
      if (!m_totals_enqueued) enqueueStatTotals();
      check(gasrt_stream_synchronize(m_stream), "sync");
      m_totals_enqueued = false;
      for (int k = 0; k < STAT_SLOTS; ++k)
        tot[k] = __atomic_load_n(m_status + STATUS_SLOTS + k, __ATOMIC_ACQUIRE);
    } else {
      std::vector<int64_t> acc((size_t)STAT_SLOTS * SCAN_BLOCKS);
      check(gasrt_memcpy_d2h(acc.data(), m_stat.p, sizeof(int64_t) * acc.size(),
                             m_stream),
            "stats");
      for (int k = 0; k < STAT_SLOTS; ++k)
        for (int b = 0; b < SCAN_BLOCKS; ++b) tot[k] += acc[(size_t)k * SCAN_BLOCKS + b];
    }
    st->supersteps = (int32_t)m_supersteps;
    st->reserved = 0;
    st->gathered_edges = tot[0];
    st->scattered_edges = tot[1];
    st->active_total = tot[2];
    st->returned_total = tot[3];
    st->hub_vertices_total = tot[4];
    st->hub_edges_total = tot[5];
    st->gather_ms = m_gather_ms;
    st->scatter_ms = m_scatter_ms;
    st->gather_launches = m_kernel_launches[GAS_K_GATHER];
    st->scatter_launches =
        m_kernel_launches[GAS_K_APPLY] + m_kernel_launches[GAS_K_SCATTER];
    for (int i = 0; i < GAS_K_SLOTS; ++i) {
      st->kernel_ms[i] = m_kernel_ms[i];
      st->kernel_launches[i] = m_kernel_launches[i];
    }
  }

  gasrt_stream_t stream() const { return m_stream; }

  // Two events kept for timing whole runs on the engine's stream.
  gasrt_event_t *runEvents() {
    if (m_ext) throw Error(GAS_ERR_STATE, "runEvents: engine without host resources");
    for (gasrt_event_t &ev : m_run_ev)
      if (!ev) check(gasrt_event_create(&ev), "event");
    return m_run_ev;
  }
  Int numVertices() const { return m_nV; }
  Int numEdges() const { return m_nE; }
  int maxInDegree() const { return m_max_in; }
  int maxOutDegree() const { return m_max_out; }
  VertexData *deviceVertexData() { return m_dev.vdata; }
  const Dev<Program> &dev() const { return m_dev; }

 private:
  static constexpr int kDepth = 3;

  bool inHubs() const { return m_max_in > m_gt2; }
  bool outHubs() const { return m_max_out > m_st2; }

  void gatherPhase() {
    refreshSourceValues();
    klaunch(GAS_K_GATHER, 0, k_gather_vp<Program>, SCAN_BLOCKS);
    if (inHubs()) {
      if (m_in != IN_ALL) {  // all active: the hub list is fixed per graph
        klaunch(GAS_K_HSCAN1, 0, k_hscan1<Program, 0>, SCAN_BLOCKS);
        klaunch(GAS_K_HSCAN2, 0, k_hscan2<Program, 0>, SCAN_BLOCKS);
      }
      if constexpr (is_source_only<Program>::value && is_compensated<Program>::value) {
        if (m_in == IN_ALL && m_dev.span_first) {
          klaunch(GAS_K_HGATHER, 0, k_hgather<Program, true>, MP_BLOCKS);
          klaunch(GAS_K_HFIX, 0, k_hfix<Program, true>, SCAN_BLOCKS);
          return;
        }
      }
      klaunch(GAS_K_HGATHER, 0, k_hgather<Program>, MP_BLOCKS);
      klaunch(GAS_K_HFIX, 0, k_hfix<Program>, SCAN_BLOCKS);
    }
  }

  // Active vertices of the bitmap input of the last superstep of parity p
  // (its workgroups' counts; the stream must be idle).
  int64_t bitmapInputSize(int p) {
    std::vector<int32_t> c(SCAN_BLOCKS);
    check(gasrt_memcpy_d2h(c.data(), m_bcnt.p + (size_t)p * SCAN_BLOCKS,
                           sizeof(int32_t) * SCAN_BLOCKS, m_stream),
          "bitmap count");
    int64_t t = 0;
    for (int32_t x : c) t += x;
    return t;
  }

  // Supersteps of k_small from the current list frontier, then the host's
  // bookkeeping of the supersteps it ran (synchronises the stream).
  void smallSteps(bool haveGather) {
    m_host_active = -1;
    refreshSourceValues();
    m_dev.seq = (int32_t)m_seq;
    m_dev.iter = m_iter;
    m_dev.in_mode = IN_LIST;
    m_dev.out_list = 1;
    m_dev.first = 0;
    m_dev.pub = 0;
    const bool timed = kernel_timing_flag() != 0;
    if (timed) timedMark(true);
    if (haveGather) {
      launch_wg(k_small<Program, true>, 1, SMALL_NT, m_stream, m_dev, m_par,
                1 << 20, m_small_max);
    } else {
      bool lds = false;
      if constexpr (sizeof(VertexData) <= 4)
        lds = m_nV <= SMALL_LDS_V && m_nE <= SMALL_LDS_E && m_small_lds;
      if constexpr (sizeof(VertexData) <= 4) {
        if (lds)
          launch_wg(k_small_lds<Program>, 1, SMALL_LDS_NT, m_stream, m_dev, m_par,
                    1 << 20, m_small_max, smallLdsGroup());
      }
      if (!lds)
        launch_wg(k_small<Program, false>, 1, SMALL_NT, m_stream, m_dev, m_par,
                  1 << 20, m_small_max);
    }
    if (timed) timedMark(false);
    const Ctl c = readCtl();  // synchronises
    const int k = c.small_steps;
    m_seq += k;
    m_iter += k;
    m_supersteps += k;
    if (k & 1) m_par ^= 1;
    m_prev_in = IN_LIST;
    m_in = IN_LIST;
    m_first = true;
    m_flags_dirty = true;
    m_other_stale = true;
  }

  // gatherMap(v) of every vertex for source-only programs, after the state
  // changed outside apply (upload, device fill, setGraph).
  void refreshSourceValues() {
    if (!is_source_only<Program>::value || !m_srcval_stale) return;
    const unsigned grid = (unsigned)std::min<int64_t>(
        2048, std::max<int64_t>(1, ((int64_t)m_nV + NT - 1) / NT));
    launch(k_srcval_all<Program>, grid, m_stream, m_dev);
    m_srcval_stale = false;
  }

  // Hub scatter, optional ordered compaction, and the move to the next
  // superstep: parity, counter-ring index, input mode of the next frontier.
  void endSuperstep(bool compact, bool hubScatter = true) {
    if (hubScatter && outHubs()) {
      klaunch(GAS_K_HSCAN1, 1, k_hscan1<Program, 1>, SCAN_BLOCKS);
      klaunch(GAS_K_HSCAN2, 1, k_hscan2<Program, 1>, SCAN_BLOCKS);
      klaunch(GAS_K_HSCATTER, 1, k_hscatter<Program>, MP_BLOCKS);
    }
    if (compact) {
      klaunch(GAS_K_COMPACT1, 2, k_compact1<Program>, SCAN_BLOCKS);
      klaunch(GAS_K_COMPACT2, 2, k_compact2<Program>, SCAN_BLOCKS);
    }
    m_flags_dirty = true;
    m_prev_in = m_in;
    m_in = (compact || m_out_list) ? IN_LIST : IN_BITMAP;
    m_par ^= 1;
    ++m_seq;
    ++m_iter;
    ++m_supersteps;
    m_first = true;
    m_host_active = -1;
  }

  // What the first kernel of the current superstep reports.
  int pubKind() const {
    int k = 0;
    if (m_in == IN_LIST) k |= PUB_OWN_LIST;
    if (m_in == IN_ALL) k |= PUB_OWN_ALL;
    if (m_prev_in == IN_BITMAP) k |= PUB_PREV_BITMAP;
    return k;
  }

  // Gather hubs of an all-active superstep: vertices with in-degree > t2,
  // their in-degree scan and tile partition (fixed for the graph).
  void buildHubs() {
    std::vector<int32_t> hubs, hoff(1, 0);
    for (int32_t v = 0; v < m_nV; ++v) {
      const int32_t deg = m_h_csc_off[v + 1] - m_h_csc_off[v];
      if (deg > m_gt2) {
        hubs.push_back(v);
        hoff.push_back(hoff.back() + deg);
      }
    }
    std::vector<int32_t> part = host_tile_partition(hoff);
    m_ahub.upload(hubs, m_stream, m_ext);
    m_ahoff.upload(hoff, m_stream, m_ext);
    m_ahpart.upload(part, m_stream, m_ext);
    m_dev.gt1 = m_gt1;
    m_dev.gt2 = m_gt2;
    m_dev.st1 = m_st1;
    m_dev.st2 = m_st2;
    m_dev.ahub = m_ahub.p;
    m_dev.ahoff = m_ahoff.p;
    m_dev.ahpart = m_ahpart.p;
    m_dev.n_ahub = (int32_t)hubs.size();
    m_dev.ahub_items = hoff.back();
    buildSpan(hubs);
  }

  // The all-active hubs as one span of vertices (see hgather_span), when
  // they are consecutive ids and the program can use it: the vertex holding
  // the first edge of each SPAN_TE-edge tile of their in-edges.
  void buildSpan(const std::vector<int32_t> &hubs) {
    m_dev.span_first = nullptr;
    if constexpr (is_source_only<Program>::value && is_compensated<Program>::value) {
      if (hubs.empty() || hubs.back() - hubs.front() + 1 != (int32_t)hubs.size()) return;
      const int32_t lo = hubs.front(), hi = hubs.back() + 1;
      const int32_t e0 = m_h_csc_off[lo], e1 = m_h_csc_off[hi];
      if ((int64_t)e1 + SPAN_TE >= INT32_MAX) return;
      const int32_t ntiles = (e1 - e0 + SPAN_TE - 1) / SPAN_TE;
      std::vector<int32_t> first((size_t)ntiles);
      int32_t v = lo;
      for (int32_t t = 0; t < ntiles; ++t) {
        const int32_t e = e0 + t * SPAN_TE;
        while (m_h_csc_off[v + 1] <= e) ++v;
        first[(size_t)t] = v;
      }
      m_span.upload(first, m_stream, m_ext);
      m_dev.span_first = m_span.p;
      m_dev.span_lo = lo;
      m_dev.span_hi = hi;
      m_dev.span_e0 = e0;
      m_dev.span_e1 = e1;
    }
  }

  void bindGraph(gas_graph_t g, VertexData *vertexData,
                 const EdgeData *edgeData) {
    gas_graph_view gv;
    check(gas_graph_view_get(g, &gv), "gas_graph_view_get");
    m_graph = g;
    m_nV = gv.n_vertices;
    m_nE = gv.n_edges;
    m_vd_host = vertexData;
    m_vdata.alloc((size_t)m_nV, m_ext);
    m_vdata2.alloc((size_t)m_nV, m_ext);
    if (vertexData)
      check(gasrt_memcpy_h2d(m_vdata.p, vertexData,
                             sizeof(VertexData) * (size_t)m_nV, m_stream),
            "vertex data");
    const EdgeData *ed_dev = nullptr;
    m_ed_host = nullptr;
    if constexpr (scatters_edges<Program>::value && !kEmptyEdge) {
      // scatters write the edge data: an engine-owned copy in CSC order,
      // and the CSR slot -> CSC slot map of their out-edges
      if (m_nE > 0 && !edgeData)
        throw Error(GAS_ERR_ARG, "setGraph: program scatters to edge data it needs");
      std::vector<int32_t> off((size_t)m_nV + 1), nb((size_t)m_nE), csr_eid((size_t)m_nE);
      m_csc_eid.assign((size_t)m_nE, 0);
      check(gas_graph_host_csc(g, off.data(), nb.data(), m_csc_eid.data()), "host csc");
      check(gas_graph_host_csr(g, off.data(), nb.data(), csr_eid.data()), "host csr");
      std::vector<int32_t> pos((size_t)m_nE), csr_csc((size_t)m_nE);
      for (Int k = 0; k < m_nE; ++k) pos[(size_t)m_csc_eid[(size_t)k]] = k;
      for (Int k = 0; k < m_nE; ++k) csr_csc[(size_t)k] = pos[(size_t)csr_eid[(size_t)k]];
      m_csr_csc.upload(csr_csc, m_stream, m_ext);
    }
    if (!kEmptyEdge) {
      if (edgeData) {
        std::vector<EdgeData> sorted((size_t)m_nE);
        check(gas_graph_permute_to_csc(g, edgeData, sorted.data(),
                                       sizeof(EdgeData)),
              "permute edge data");
        m_edata.upload(sorted, m_stream, m_ext);
        ed_dev = m_edata.p;
      } else if (sizeof(EdgeData) == 4 && gv.csc_weights) {
        ed_dev = reinterpret_cast<const EdgeData *>(gv.csc_weights);
      } else if (m_nE > 0) {
        throw Error(GAS_ERR_ARG, "setGraph: program needs edge data");
      }
    }
    // degree extremes decide which tiers can occur
    m_h_csc_off.assign((size_t)m_nV + 1, 0);
    std::vector<int32_t> csr_off((size_t)m_nV + 1, 0);
    check(gas_graph_host_csc(g, m_h_csc_off.data(), nullptr, nullptr), "csc");
    check(gas_graph_host_csr(g, csr_off.data(), nullptr, nullptr), "csr");
    m_max_in = m_max_out = 0;
    for (int32_t v = 0; v < m_nV; ++v) {
      m_max_in = std::max(m_max_in, m_h_csc_off[v + 1] - m_h_csc_off[v]);
      m_max_out = std::max(m_max_out, csr_off[v + 1] - csr_off[v]);
    }

    const size_t V = (size_t)m_nV;
    const int nfwords = (int)((V + 3) / 4);
    // tile partials: merge-path tiles of edges + vertices, or span tiles of
    // SPAN_TE edges (buildSpan), whichever form has more
    const size_t tiles =
        std::max(((size_t)m_nE + V) / MP_NV, (size_t)m_nE / SPAN_TE) + 2;
    ExternalMemory *x = m_ext;
    m_ctl.alloc(1, x);
    m_q0.alloc(V, x);
    m_q1.alloc(V, x);
    m_fl0.alloc((size_t)nfwords * 4, x);
    m_fl1.alloc((size_t)nfwords * 4, x);
    m_chg0.alloc((size_t)nfwords * 4, x);
    m_chg1.alloc((size_t)nfwords * 4, x);
    m_rflag.alloc(V, x);
    m_gtmp.alloc(V, x);
    m_ghub.alloc(V, x);
    m_ghoff.alloc(V + 1, x);
    m_ghpart.alloc(tiles, x);
    m_shub.alloc(V, x);
    m_shoff.alloc(V + 1, x);
    m_shpart.alloc(tiles, x);
    m_bsum.alloc(SCAN_BLOCKS, x);
    m_stat.alloc(STAT_SLOTS * SCAN_BLOCKS, x);
    m_bcnt.alloc(2 * SCAN_BLOCKS, x);
    m_rpart.alloc(SCAN_BLOCKS, x);
    m_pfirst.alloc(tiles, x);
    m_plast.alloc(tiles, x);
    zero(m_fl0.p, (size_t)nfwords * 4);
    zero(m_fl1.p, (size_t)nfwords * 4);
    zero(m_stat.p, sizeof(int64_t) * STAT_SLOTS * SCAN_BLOCKS);
    zero(m_bcnt.p, sizeof(int32_t) * 2 * SCAN_BLOCKS);
    zero(m_rpart.p, sizeof(double) * SCAN_BLOCKS);
    zero(m_ctl.p, sizeof(Ctl));

    Dev<Program> &d = m_dev;
    d.nV = m_nV;
    d.nE = m_nE;
    d.csc_off = gv.csc_offsets;
    d.csc_src = gv.csc_srcs;
    d.csr_off = gv.csr_offsets;
    d.csr_dst = gv.csr_dsts;
    d.edata = ed_dev;
    // CSR-order copy of the edge values for push supersteps: the graph's
    // own weights (caller edge data is kept in CSC order only)
    d.edata_csr = nullptr;
    if constexpr (!kEmptyEdge && sizeof(EdgeData) == 4)
      if (!edgeData && gv.csr_weights)
        d.edata_csr = reinterpret_cast<const EdgeData *>(gv.csr_weights);
    d.edata_rw = scatters_edges<Program>::value ? m_edata.p : nullptr;
    // out-neighbour table for the fused activation and the push (graphs of
    // out-degree <= ELL_W; its values only where the engine uses the
    // graph's own weights); not on an engine that must not allocate
    d.ell_dst = nullptr;
    d.ell_w = nullptr;
    if (!m_ext && m_max_out <= ELL_W && m_nV > 0) {
      const int32_t *ed = nullptr, *ew = nullptr;
      check(gas_graph_ell(g, &ed, &ew), "gas_graph_ell");
      if constexpr (kEmptyEdge) {
        d.ell_dst = ed;
      } else if constexpr (sizeof(EdgeData) == 4) {
        if (d.edata_csr && ew) {
          d.ell_dst = ed;
          d.ell_w = reinterpret_cast<const EdgeData *>(ew);
        }
      }
    }
    d.csr_csc = m_csr_csc.p;
    d.vdata = m_vdata.p;
    d.vnext = m_vdata2.p;
    m_srcval.alloc(is_source_only<Program>::value ? (size_t)m_nV : 1, x);
    d.srcval = m_srcval.p;
    m_srcval_stale = true;
    m_flags_dirty = true;
    d.chg[0] = m_chg0.p;
    d.chg[1] = m_chg1.p;
    m_other_stale = true;
    d.ctl = m_ctl.p;
    d.q[0] = m_q0.p;
    d.q[1] = m_q1.p;
    d.fl[0] = m_fl0.p;
    d.fl[1] = m_fl1.p;
    d.nfwords = nfwords;
    d.rflag = m_rflag.p;
    d.gtmp = m_gtmp.p;
    d.ghub = m_ghub.p;
    d.ghoff = m_ghoff.p;
    d.ghpart = m_ghpart.p;
    d.shub = m_shub.p;
    d.shoff = m_shoff.p;
    d.shpart = m_shpart.p;
    d.bsum = m_bsum.p;
    d.bcnt = m_bcnt.p;
    d.stat_acc = m_stat.p;
    d.rpart = m_rpart.p;
    d.part_first = m_pfirst.p;
    d.part_last = m_plast.p;
    m_par = 0;
    m_in = m_prev_in = IN_LIST;
    // Below 2^24 vertices the bitmap scan (V bytes per frontier kernel,
    // staged through LDS) is cheaper than claiming destinations with
    // atomics even for one-vertex frontiers; above, sparse frontiers use
    // the list.
    m_switch = m_nV >= (1 << 24) ? 64 : 0;
    // k_small's chain of dependent loads and claims per vertex costs more
    // than a grid superstep once frontiers reach a few hundred vertices, so
    // it is on by default only for graphs whose whole run stays that small
    m_small_max = m_nV <= kSmallGraph ? 4096 : 0;
    buildHubs();
    setActive(0, 0);
  }

  // Launch of an engine kernel taking (Dev, parity); with kernel timing on,
  // bracketed by events and booked under slot `id` and phase `phase`
  // (0 gather, 1 apply / scatter, 2 other).
  template <class K>
  void klaunch(int id, int phase, K k, unsigned grid) {
    const bool timed = kernel_timing_flag() != 0 && !m_ext;
    if (timed) {
      if (m_tpos == m_timing.size()) {
        gasrt_event_t a, b;
        check(gasrt_event_create(&a), "event");
        check(gasrt_event_create(&b), "event");
        m_timing.push_back({a, b});
        m_tkind.push_back(0);
      }
      check(gasrt_event_record(m_timing[m_tpos].first, m_stream), "event");
    }
    m_dev.seq = (int32_t)m_seq;
    m_dev.iter = m_iter;
    m_dev.in_mode = m_in;
    m_dev.out_list = m_out_list;
    m_dev.first = m_first ? 1 : 0;
    m_dev.pub = m_first ? pubKind() : 0;
    m_first = false;
    launch(k, grid, m_stream, m_dev, m_par);
    if (timed) {
      check(gasrt_event_record(m_timing[m_tpos].second, m_stream), "event");
      m_tkind[m_tpos] = id * 4 + phase;
      ++m_tpos;
      if (m_tpos >= 4096) collectTiming();
    }
  }
  // Event brackets around a launch that does not go through klaunch.
  void timedMark(bool begin) {
    if (begin) {
      if (m_tpos == m_timing.size()) {
        gasrt_event_t a, b;
        check(gasrt_event_create(&a), "event");
        check(gasrt_event_create(&b), "event");
        m_timing.push_back({a, b});
        m_tkind.push_back(0);
      }
      check(gasrt_event_record(m_timing[m_tpos].first, m_stream), "event");
    } else {
      check(gasrt_event_record(m_timing[m_tpos].second, m_stream), "event");
      m_tkind[m_tpos] = GAS_K_SMALL * 4 + 2;
      ++m_tpos;
      if (m_tpos >= 4096) collectTiming();
    }
  }
  // Clears device memory: a memset, or on an engine without its own
  // memory a fill kernel (nothing the caller's stream could not replay)
  void zero(void *p, size_t bytes) {
    if (!m_ext) {
      check(gasrt_memset(p, 0, bytes, m_stream), "memset");
      return;
    }
    const int64_t words = (int64_t)(bytes / 4);
    const unsigned grid =
        (unsigned)std::min<int64_t>(1024, std::max<int64_t>(1, (words + NT - 1) / NT));
    launch(k_zero_words<uint32_t>, grid, m_stream, static_cast<uint32_t *>(p), words);
  }

  void collectTiming() {
    if (m_tpos == 0) return;
    check(gasrt_stream_synchronize(m_stream), "sync");
    for (size_t i = 0; i < m_tpos; ++i) {
      float ms = 0;
      check(gasrt_event_elapsed_ms(m_timing[i].first, m_timing[i].second,
                                   &ms),
            "elapsed");
      const int id = m_tkind[i] / 4, phase = m_tkind[i] % 4;
      m_kernel_ms[id] += ms;
      ++m_kernel_launches[id];
      if (phase == 0) m_gather_ms += ms;
      if (phase == 1) m_scatter_ms += ms;
    }
    m_tpos = 0;
  }

  ExternalMemory m_ext_mem;          // the external memory, when constructed with it
  ExternalMemory *m_ext = nullptr;   // &m_ext_mem on an engine without its own memory
  gasrt_stream_t m_stream = nullptr;
  Ctl *m_hctl = nullptr;            // pinned copy of the control block
  int64_t *m_status = nullptr;      // mapped status ring (STATUS_SLOTS), then
                                    // the run totals (STAT_SLOTS)
  int64_t *m_dev_totals = nullptr;  // device address of the run totals
  bool m_totals_enqueued = false;   // k_stat_totals enqueued since stats()
  gasrt_event_t m_run_ev[2] = {nullptr, nullptr};  // whole-run timer
  int64_t m_host_active = -1;       // frontier size set by setActive (-1: unknown)
  int64_t m_seq = 0;                // superstep number
  int m_iter = 0;                   // program-visible superstep counter
  int64_t m_supersteps = 0;         // supersteps with work since resetStats
  int m_in = IN_LIST, m_prev_in = IN_LIST;  // input mode: this / last superstep
  int m_out_list = 0;               // scatter of this superstep builds a list
  bool m_first = true;              // next launch is the superstep's first
  int m_switch = 0;                 // density switch divisor (0: bitmap only)
  int m_small_max = 0;              // k_small frontier limit (0: off)
  bool m_push = true;               // push supersteps when usable (pushMin)
  bool m_push_run = false;          // the current whole run pushes
  static constexpr int kSmallGraph = 8192;  // default k_small graphs: V <= this
  bool m_other_stale = true;        // vnext does not mirror vdata
  bool m_srcval_stale = true;       // srcval does not match the state
  bool m_small_lds = true;          // setSmallLds
  bool m_flags_dirty = true;        // a scatter ran since the last clear
  gas_graph_t m_graph = nullptr;
  gas_graph_t m_own_graph = nullptr;
  Int m_nV = 0, m_nE = 0;
  int m_gt1 = DEFAULT_T1, m_gt2 = DEFAULT_T2;
  int m_st1 = DEFAULT_T1, m_st2 = DEFAULT_T2;
  int32_t m_max_in = 0, m_max_out = 0;
  std::vector<int32_t> m_h_csc_off;
  int m_par = 0;  // parity of the next superstep
  VertexData *m_vd_host = nullptr;
  EdgeData *m_ed_host = nullptr;        // scatterEdges: caller's edge array
  std::vector<int32_t> m_csc_eid;       // scatterEdges: edge-list index per CSC slot
  DevBuf<int32_t> m_csr_csc;            // scatterEdges: CSR slot -> CSC slot
  DevBuf<VertexData> m_vdata, m_vdata2;  // state, double-buffered by k_fused
  DevBuf<EdgeData> m_edata;
  DevBuf<Ctl> m_ctl;
  DevBuf<int32_t> m_q0, m_q1, m_ghub, m_ghoff, m_ghpart, m_shub, m_shoff,
      m_shpart, m_bsum, m_bcnt, m_ahub, m_ahoff, m_ahpart, m_span;
  DevBuf<int64_t> m_stat;
  DevBuf<uint8_t> m_fl0, m_fl1, m_chg0, m_chg1, m_rflag;
  DevBuf<double> m_rpart;
  DevBuf<GatherResult> m_gtmp, m_srcval;
  DevBuf<Acc<Program>> m_pfirst, m_plast;
  Dev<Program> m_dev{};
  std::vector<std::pair<gasrt_event_t, gasrt_event_t>> m_timing;
  std::vector<int> m_tkind;
  size_t m_tpos = 0;
  double m_gather_ms = 0, m_scatter_ms = 0;
  double m_kernel_ms[GAS_K_SLOTS] = {};
  int m_kernel_launches[GAS_K_SLOTS] = {};
};

}  // namespace gas
