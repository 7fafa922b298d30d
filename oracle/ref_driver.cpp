// ref_driver.cpp -- drives the reference's own CPU engine, GASEngineRef
// (refgas.h) and CSR/CSC helpers (util.h), compiled from the reference
// sources where they lie (see oracle/Makefile).  TEST INFRASTRUCTURE ONLY:
// it validates the C restatement in gasref.c and is the CPU baseline timed
// by bench.py.  Nothing here is part of libgas.so.
//
// The four program structs restate bfs.cu:31-78, sssp.cu:25-64,
// pagerank.cu:27-73 and connected_component.cu:25-63 on the host path
// (those .cu files also include gpugas.h, which needs moderngpu, so they are
// not compiled here).  The run functions restate each driver's run<Engine>()
// and initialisation.
#include <climits>
#include <cmath>
#include <cstdint>
#include <algorithm>
#include <vector>

#include "refgas.h"
#include "util.h"

namespace {

int g_iterationCount;  // bfs.cu:27, host path of bfs.cu:62-66

struct BFS {
  struct VertexData {
    int depth;
  };
  struct EdgeData {};
  typedef int GatherResult;
  static const int gatherZero = INT_MAX - 1;
  static int gatherReduce(const int &, const int &) { return 0; }
  static int gatherMap(const VertexData *, const VertexData *,
                       const EdgeData *) {
    return 0;
  }
  static bool apply(VertexData *vert, int) {
    if (vert->depth == -1) {
      vert->depth = g_iterationCount;
      return true;
    }
    return false;
  }
  static void scatter(const VertexData *, const VertexData *, EdgeData *) {}
};

struct SSSP {
  typedef int VertexData;
  typedef int EdgeData;
  typedef int GatherResult;
  static const int maxLength = 100000;
  static const int gatherZero = INT_MAX - maxLength;
  static int gatherReduce(const int &left, const int &right) {
    return std::min(left, right);
  }
  static int gatherMap(const VertexData *, const VertexData *srcDist,
                       const EdgeData *edgeLen) {
    return *srcDist + *edgeLen;
  }
  static bool apply(VertexData *curDist, GatherResult dist) {
    bool changed = dist < *curDist;
    *curDist = std::min(*curDist, dist);
    return changed;
  }
  static void scatter(const VertexData *, const VertexData *, EdgeData *) {}
};

struct PageRank {
  static constexpr float pageConst = 0.15f;
  static constexpr float tol = 0.01f;
  struct VertexData {
    float rank;
    int numOutEdges;
  };
  struct EdgeData {};
  typedef float GatherResult;
  static constexpr float gatherZero = 0.0f;
  static float gatherMap(const VertexData *, const VertexData *src,
                         const EdgeData *) {
    return src->rank / src->numOutEdges;
  }
  static float gatherReduce(const float &left, const float &right) {
    return left + right;
  }
  static bool apply(VertexData *vertexData, const float &gatherResult) {
    float newRank = pageConst + (1.0f - pageConst) * gatherResult;
    bool ret = std::fabs(newRank - vertexData->rank) >= tol;
    vertexData->rank = newRank;
    return ret;
  }
  static void scatter(const VertexData *, const VertexData *, EdgeData *) {}
};

struct CC {
  typedef int VertexData;
  struct EdgeData {};
  typedef int GatherResult;
  static const int gatherZero = INT_MAX;
  static int gatherReduce(const int &left, const int &right) {
    return std::min(left, right);
  }
  static int gatherMap(const VertexData *, const VertexData *srcLabel,
                       const EdgeData *) {
    return *srcLabel;
  }
  static bool apply(VertexData *curLabel, GatherResult label) {
    bool changed = label < *curLabel;
    *curLabel = std::min(*curLabel, label);
    return changed;
  }
  static void scatter(const VertexData *, const VertexData *, EdgeData *) {}
};

// Personalised PageRank (no counterpart in the reference): the PageRank
// struct above with the teleport 0.15 paid only at the source, run as fixed
// all-active sweeps -- the CPU side of the served stream's PPR queries.
struct PPR {
  struct VertexData {
    float rank;
    int numOutEdges;
    int isSource;
  };
  struct EdgeData {};
  typedef float GatherResult;
  static constexpr float gatherZero = 0.0f;
  static float gatherMap(const VertexData *, const VertexData *src,
                         const EdgeData *) {
    return src->rank / src->numOutEdges;
  }
  static float gatherReduce(const float &left, const float &right) {
    return left + right;
  }
  static bool apply(VertexData *v, const float &sum) {
    v->rank = (v->isSource ? 0.15f : 0.0f) + 0.85f * sum;
    return false;
  }
  static void scatter(const VertexData *, const VertexData *, EdgeData *) {}
};

// One engine of the reference bound to a graph once (setGraph outside any
// timed window) and reused query after query: the vertex array it retains
// (refgas.h:90) is re-initialised in place before each run.
struct Prepared {
  int kind = 0;  // 0 BFS, 1 SSSP, 2 PPR
  int nV = 0;
  std::vector<int> weights;
  std::vector<int> outdeg;
  std::vector<BFS::VertexData> bfs_vd;
  std::vector<int> sssp_vd;
  std::vector<PPR::VertexData> ppr_vd;
  GASEngineRef<BFS> bfs;
  GASEngineRef<SSSP> sssp;
  GASEngineRef<PPR> ppr;
};

// SSSP whose scatter writes its edge: hits counts the scatters of the
// edge, last the distance difference of its ends after the superstep's
// applies (the test program of tests/template/edge_trace.hip, here for the
// reference's own engine, whose run() calls scatter: refgas.h:178-192).
struct EdgeTrace {
  static const int kUnreached = INT_MAX - 100000;
  struct VertexData {
    int dist;
  };
  struct EdgeData {
    int w;
    int hits;
    int last;
  };
  typedef int GatherResult;
  static const int gatherZero = INT_MAX;
  static int gatherReduce(const int &a, const int &b) { return std::min(a, b); }
  static int gatherMap(const VertexData *, const VertexData *src,
                       const EdgeData *e) {
    return src->dist >= kUnreached ? INT_MAX : src->dist + e->w;
  }
  static bool apply(VertexData *v, GatherResult g) {
    if (g < v->dist) {
      v->dist = g;
      return true;
    }
    return false;
  }
  static void scatter(const VertexData *src, const VertexData *dst,
                      EdgeData *e) {
    e->hits += 1;
    e->last = src->dist - dst->dist;
  }
};

}  // namespace

extern "C" {

// bfs.cu:91-112; returns microseconds of the timed window (frontier init
// through results, as "Took" in the reference).
int64_t ref_bfs(int nV, int nE, const int *srcs, const int *dsts, int source,
                int *depth) {
  std::vector<BFS::VertexData> vd(nV);
  for (int i = 0; i < nV; ++i) vd[i].depth = -1;
  GASEngineRef<BFS> engine;
  engine.setGraph(nV, vd.data(), nE, 0, srcs, dsts);
  engine.setActive(source, source + 1);
  int iter = 0;
  g_iterationCount = iter;
  int64_t t0 = currentTime();
  while (engine.countActive()) {
    engine.gatherApply(false);
    engine.scatterActivate(false);
    engine.nextIter();
    g_iterationCount = ++iter;
  }
  engine.getResults();
  int64_t t1 = currentTime();
  for (int i = 0; i < nV; ++i) depth[i] = vd[i].depth;
  return t1 - t0;
}

// sssp.cu:67-82 with sssp.cu:358-362
int64_t ref_sssp(int nV, int nE, const int *srcs, const int *dsts,
                 const int *weights, int source, int *dist) {
  std::vector<int> vd(nV, SSSP::gatherZero);
  vd[source] = 0;
  std::vector<int> w(weights, weights + nE);
  GASEngineRef<SSSP> engine;
  engine.setGraph(nV, vd.data(), nE, w.data(), srcs, dsts);
  engine.setActive(0, nV);
  int64_t t0 = currentTime();
  engine.run();
  engine.getResults();
  int64_t t1 = currentTime();
  for (int i = 0; i < nV; ++i) dist[i] = vd[i];
  return t1 - t0;
}

// connected_component.cu:66-81 with :112-114
int64_t ref_cc(int nV, int nE, const int *srcs, const int *dsts, int *label) {
  std::vector<int> vd(nV);
  for (int i = 0; i < nV; ++i) vd[i] = i;
  GASEngineRef<CC> engine;
  engine.setGraph(nV, vd.data(), nE, 0, srcs, dsts);
  engine.setActive(0, nV);
  int64_t t0 = currentTime();
  engine.run();
  engine.getResults();
  int64_t t1 = currentTime();
  for (int i = 0; i < nV; ++i) label[i] = vd[i];
  return t1 - t0;
}

// pagerank.cu:85-101 with pagerank.cu:126-132.  iterations > 0 replaces
// run() by that many all-active sweeps (setActive(0,V); gatherApply()).
int64_t ref_pagerank(int nV, int nE, const int *srcs, const int *dsts,
                     int iterations, float *rank) {
  std::vector<int> srcOffsets(nV + 1);
  edgeListToCSR<int>(nV, nE, srcs, dsts, &srcOffsets[0], 0, 0);
  std::vector<PageRank::VertexData> vd(nV);
  for (int i = 0; i < nV; ++i) {
    vd[i].numOutEdges = srcOffsets[i + 1] - srcOffsets[i];
    vd[i].rank = PageRank::pageConst;
  }
  GASEngineRef<PageRank> engine;
  engine.setGraph(nV, vd.data(), nE, 0, srcs, dsts);
  int64_t t0 = currentTime();
  if (iterations > 0) {
    for (int it = 0; it < iterations; ++it) {
      engine.setActive(0, nV);
      engine.gatherApply();
    }
  } else {
    engine.setActive(0, nV);
    engine.run();
  }
  engine.getResults();
  int64_t t1 = currentTime();
  for (int i = 0; i < nV; ++i) rank[i] = vd[i].rank;
  return t1 - t0;
}

// pagerank.cu:85-101 with pagerank.cu:126-132, the tolerance-driven run:
// refgas.h:215-222's run() loop (gatherApply, scatterActivate, nextIter
// while countActive) with its supersteps counted; returns the supersteps.
int ref_pagerank_run(int nV, int nE, const int *srcs, const int *dsts,
                     float *rank) {
  std::vector<int> srcOffsets(nV + 1);
  edgeListToCSR<int>(nV, nE, srcs, dsts, &srcOffsets[0], 0, 0);
  std::vector<PageRank::VertexData> vd(nV);
  for (int i = 0; i < nV; ++i) {
    vd[i].numOutEdges = srcOffsets[i + 1] - srcOffsets[i];
    vd[i].rank = PageRank::pageConst;
  }
  GASEngineRef<PageRank> engine;
  engine.setGraph(nV, vd.data(), nE, 0, srcs, dsts);
  engine.setActive(0, nV);
  int steps = 0;
  while (engine.countActive()) {
    engine.gatherApply();
    engine.scatterActivate();
    engine.nextIter();
    ++steps;
  }
  engine.getResults();
  for (int i = 0; i < nV; ++i) rank[i] = vd[i].rank;
  return steps;
}

// The reference engine's run() (refgas.h:215-222, scatter included) on
// EdgeTrace, every vertex active first; returns the supersteps run.
int ref_edge_trace(int nV, int nE, const int *srcs, const int *dsts,
                   const int *w, int source, int *dist, int *hits, int *last) {
  std::vector<EdgeTrace::VertexData> vd(nV);
  for (int i = 0; i < nV; ++i) vd[i].dist = i == source ? 0 : EdgeTrace::kUnreached;
  std::vector<EdgeTrace::EdgeData> ed(nE);
  for (int i = 0; i < nE; ++i) ed[i] = {w[i], 0, 0};
  GASEngineRef<EdgeTrace> engine;
  engine.setGraph(nV, vd.data(), nE, ed.data(), srcs, dsts);
  engine.setActive(0, nV);
  int steps = 0;
  while (engine.countActive()) {
    engine.gatherApply();
    engine.scatterActivate();
    engine.nextIter();
    ++steps;
  }
  engine.getResults();
  for (int i = 0; i < nV; ++i) dist[i] = vd[i].dist;
  for (int i = 0; i < nE; ++i) {
    hits[i] = ed[i].hits;
    last[i] = ed[i].last;
  }
  return steps;
}

// A prepared engine of `kind` (0 BFS, 1 SSSP, 2 personalised PageRank)
// on the graph; weights are needed for SSSP only.
void *ref_prepare(int kind, int nV, int nE, const int *srcs, const int *dsts,
                  const int *weights) {
  Prepared *p = new Prepared;
  p->kind = kind;
  p->nV = nV;
  if (kind == 0) {
    p->bfs_vd.assign(nV, BFS::VertexData{-1});
    p->bfs.setGraph(nV, p->bfs_vd.data(), nE, 0, srcs, dsts);
  } else if (kind == 1) {
    p->weights.assign(weights, weights + nE);
    p->sssp_vd.assign(nV, SSSP::gatherZero);
    p->sssp.setGraph(nV, p->sssp_vd.data(), nE, p->weights.data(), srcs, dsts);
  } else {
    std::vector<int> off(nV + 1);
    edgeListToCSR<int>(nV, nE, srcs, dsts, &off[0], 0, 0);
    p->outdeg.resize(nV);
    for (int i = 0; i < nV; ++i) p->outdeg[i] = off[i + 1] - off[i];
    p->ppr_vd.resize(nV);
    p->ppr.setGraph(nV, p->ppr_vd.data(), nE, 0, srcs, dsts);
  }
  return p;
}

// One query on a prepared engine: the driver's initialisation and loop
// (bfs.cu:91-112; sssp.cu:67-82; `iterations` PPR sweeps); the result goes
// to out (int32 or float per vertex).  Returns microseconds of the
// reference's Took window (initialisation of the frontier through results).
int64_t ref_prepared_query(void *h, int source, int iterations, void *out) {
  Prepared *p = static_cast<Prepared *>(h);
  const int nV = p->nV;
  int64_t t0 = 0, t1 = 0;
  if (p->kind == 0) {
    for (auto &x : p->bfs_vd) x.depth = -1;
    t0 = currentTime();
    p->bfs.setActive(source, source + 1);
    int iter = 0;
    g_iterationCount = iter;
    while (p->bfs.countActive()) {
      p->bfs.gatherApply(false);
      p->bfs.scatterActivate(false);
      p->bfs.nextIter();
      g_iterationCount = ++iter;
    }
    p->bfs.getResults();
    t1 = currentTime();
    int *o = static_cast<int *>(out);
    for (int i = 0; i < nV; ++i) o[i] = p->bfs_vd[i].depth;
  } else if (p->kind == 1) {
    for (auto &x : p->sssp_vd) x = SSSP::gatherZero;
    p->sssp_vd[source] = 0;
    t0 = currentTime();
    p->sssp.setActive(0, nV);
    p->sssp.run();
    p->sssp.getResults();
    t1 = currentTime();
    std::copy(p->sssp_vd.begin(), p->sssp_vd.end(), static_cast<int *>(out));
  } else {
    for (int i = 0; i < nV; ++i)
      p->ppr_vd[i] = {i == source ? 0.15f : 0.0f, p->outdeg[i], i == source ? 1 : 0};
    t0 = currentTime();
    for (int it = 0; it < iterations; ++it) {
      p->ppr.setActive(0, nV);
      p->ppr.gatherApply();
    }
    p->ppr.getResults();
    t1 = currentTime();
    float *o = static_cast<float *>(out);
    for (int i = 0; i < nV; ++i) o[i] = p->ppr_vd[i].rank;
  }
  return t1 - t0;
}

void ref_prepared_free(void *h) { delete static_cast<Prepared *>(h); }

// util.h:79-110, exposed for layout checks.
void ref_edge_list_to_csr(int nV, int nE, const int *srcs, const int *dsts,
                          int *offsets, int *outDsts, int *sortIndices) {
  edgeListToCSR<int>(nV, nE, srcs, dsts, offsets, outDsts, sortIndices);
}

}  // extern "C"
