// A vertex program with an edge-writing scatter, compiled against the
// engine's template header the way a reference user compiles a program of
// their own (one hipcc translation unit instantiates the kernels).  Test
// infrastructure: built into tests/template/_lib by tests/template/build.py.
//
// EdgeTrace is SSSP in the reference's pull form (sssp.cu:25-64) whose
// scatter (refgas.h:178-192) counts, on every out-edge of a vertex whose
// distance dropped, how often it was scattered and the distance difference
// between its ends after the superstep's applies.
#include <climits>

#include "gas/engine.hpp"

namespace {

constexpr int kUnreached = INT_MAX - 100000;

struct EdgeTrace {
  struct VertexData {
    int dist;
  };
  struct EdgeData {
    int w;
    int hits;
    int last;
  };
  typedef int GatherResult;
  static constexpr int gatherZero = INT_MAX;
  static constexpr bool scatterEdges = true;

  __host__ __device__ static int gatherReduce(const int &a, const int &b) {
    return a < b ? a : b;
  }
  __host__ __device__ static int gatherMap(const VertexData *,
                                           const VertexData *src,
                                           const EdgeData *e) {
    return src->dist >= kUnreached ? INT_MAX : src->dist + e->w;
  }
  __host__ __device__ static bool apply(VertexData *v, GatherResult g) {
    if (g < v->dist) {
      v->dist = g;
      return true;
    }
    return false;
  }
  __host__ __device__ static void scatter(const VertexData *src,
                                          const VertexData *dst, EdgeData *e) {
    e->hits += 1;
    e->last = src->dist - dst->dist;
  }
};

}  // namespace

extern "C" {

// mode 0: run(); 1: gatherApply / scatterActivate(true) / nextIter loop;
// 2: the same loop with scatterActivate(false) (edges untouched).
// Every vertex starts active (sssp.cu:67-82).  Returns 0 or the error code.
int et_run(int nV, int nE, const int *srcs, const int *dsts, const int *w,
           int source, int mode, int *dist, int *hits, int *last,
           int *supersteps) {
  try {
    std::vector<EdgeTrace::VertexData> vd((size_t)nV);
    for (int i = 0; i < nV; ++i) vd[(size_t)i].dist = i == source ? 0 : kUnreached;
    std::vector<EdgeTrace::EdgeData> ed((size_t)nE);
    for (int i = 0; i < nE; ++i) ed[(size_t)i] = {w[i], 0, 0};
    gas::GASEngineGPU<EdgeTrace> engine;
    engine.setGraph(nV, vd.data(), nE, ed.data(), srcs, dsts);
    engine.setActive(0, nV);
    int steps = 0;
    if (mode == 0) {
      engine.run();
    } else {
      while (engine.countActive()) {
        engine.gatherApply(true);
        engine.scatterActivate(mode == 1);
        engine.nextIter();
        ++steps;
      }
    }
    engine.getResults();
    for (int i = 0; i < nV; ++i) dist[i] = vd[(size_t)i].dist;
    for (int i = 0; i < nE; ++i) {
      hits[i] = ed[(size_t)i].hits;
      last[i] = ed[(size_t)i].last;
    }
    *supersteps = steps;
    return 0;
  } catch (const gas::Error &e) {
    return e.code;
  }
}

}  // extern "C"
