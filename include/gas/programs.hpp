// The four vertex programs, in the reference's program-struct shape
// (typedefs VertexData / EdgeData / GatherResult, constant gatherZero, and
// static gatherMap / gatherReduce / apply / scatter).
//
// One deliberate change: BFS reads its level from gas::iteration(), the
// superstep counter the engine keeps on the device, instead of the
// __constant__ variable the reference driver rewrites every level with
// cudaMemcpyToSymbol (bfs.cu:27-28, 62-66, 81-88).  The values are the same:
// the driver's counter also counts completed supersteps from 0.
#pragma once

#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>

#include "kernels.hpp"

namespace gas {

// bfs.cu:31-78
struct BFS {
  struct VertexData {
    int depth;
  };
  struct EdgeData {};
  typedef int GatherResult;
  static constexpr int gatherZero = INT_MAX - 1;

  __host__ __device__ static int gatherReduce(const int &, const int &) {
    return 0;
  }
  __host__ __device__ static int gatherMap(const VertexData *,
                                           const VertexData *,
                                           const EdgeData *) {
    return 0;
  }
  __host__ __device__ static bool apply(VertexData *vert, int) {
    if (vert->depth == -1) {
      vert->depth = gas::iteration();
      return true;
    }
    return false;
  }
  __host__ __device__ static void scatter(const VertexData *,
                                          const VertexData *, EdgeData *) {}
};

// sssp.cu:25-64, plus the engine's pushMin trait (whole runs may push the
// changed distances instead of pulling; same values every superstep)
struct SSSP {
  typedef int VertexData;
  typedef int EdgeData;
  typedef int GatherResult;
  static constexpr int maxLength = 100000;
  static constexpr int gatherZero = INT_MAX - maxLength;
  static constexpr bool pushMin = true;

  // an unreached vertex's pushes (gatherZero + w) lower nothing
  __host__ __device__ static bool pushSkip(const VertexData &d) {
    return d >= gatherZero;
  }

  __host__ __device__ static int gatherReduce(const int &left,
                                              const int &right) {
    return left < right ? left : right;
  }
  __host__ __device__ static int gatherMap(const VertexData *,
                                           const VertexData *srcDist,
                                           const EdgeData *edgeLen) {
    return *srcDist + *edgeLen;
  }
  __host__ __device__ static bool apply(VertexData *curDist,
                                        GatherResult dist) {
    bool changed = dist < *curDist;
    *curDist = dist < *curDist ? dist : *curDist;
    return changed;
  }
  __host__ __device__ static void scatter(const VertexData *,
                                          const VertexData *, EdgeData *) {}
};

// pagerank.cu:27-73, plus three opt-in traits of this engine: compensated
// summation of the gather, gatherMap reading only the source (so the engine
// can keep rank / numOutEdges per vertex), and an FP64 residual
// |new - old| per vertex.
struct PageRank {
  static constexpr float pageConst = 0.15f;
  static constexpr float tol = 0.01f;
  static constexpr bool compensatedSum = true;
  static constexpr bool gatherSourceOnly = true;

  struct VertexData {
    float rank;
    int numOutEdges;
  };
  struct EdgeData {};
  typedef float GatherResult;
  static constexpr float gatherZero = 0.0f;

  __host__ __device__ static float gatherMap(const VertexData *,
                                             const VertexData *src,
                                             const EdgeData *) {
    return src->rank / src->numOutEdges;
  }
  __host__ __device__ static float gatherReduce(const float &left,
                                                const float &right) {
    return left + right;
  }
  __host__ __device__ static bool apply(VertexData *vertexData,
                                        const float &gatherResult) {
    float newRank = pageConst + (1.0f - pageConst) * gatherResult;
    bool ret = fabsf(newRank - vertexData->rank) >= tol;
    vertexData->rank = newRank;
    return ret;
  }
  __host__ __device__ static double residual(const VertexData &before,
                                             const VertexData &after) {
    return fabs((double)after.rank - (double)before.rank);
  }
  __host__ __device__ static void scatter(const VertexData *,
                                          const VertexData *, EdgeData *) {}
};

// Personalised PageRank from one source: the PageRank program above with
// the teleport term 0.15 paid only at the source,
//   r_0 = 0.15 e_s,  r_{k+1}(v) = 0.15 [v == s] + 0.85 sum_{u->v} r_k(u)/outdeg(u),
// run as a fixed number of all-active sweeps.  No counterpart in the
// reference; it is the one-query-at-a-time baseline of the batched PPR.
struct PersonalizedPageRank {
  static constexpr bool compensatedSum = true;
  static constexpr bool gatherSourceOnly = true;
  struct VertexData {
    float rank;
    int numOutEdges;
    int isSource;
  };
  struct EdgeData {};
  typedef float GatherResult;
  static constexpr float gatherZero = 0.0f;

  __host__ __device__ static float gatherMap(const VertexData *,
                                             const VertexData *src,
                                             const EdgeData *) {
    return src->rank / src->numOutEdges;
  }
  __host__ __device__ static float gatherReduce(const float &left,
                                                const float &right) {
    return left + right;
  }
  __host__ __device__ static bool apply(VertexData *v, const float &sum) {
    v->rank = (v->isSource ? 0.15f : 0.0f) + 0.85f * sum;
    return false;
  }
  __host__ __device__ static void scatter(const VertexData *,
                                          const VertexData *, EdgeData *) {}
};

// connected_component.cu:25-63, plus the engine's pushMin trait
struct CC {
  typedef int VertexData;
  struct EdgeData {};
  typedef int GatherResult;
  static constexpr int gatherZero = INT_MAX;
  static constexpr bool pushMin = true;

  __host__ __device__ static int gatherReduce(const int &left,
                                              const int &right) {
    return left < right ? left : right;
  }
  __host__ __device__ static int gatherMap(const VertexData *,
                                           const VertexData *srcLabel,
                                           const EdgeData *) {
    return *srcLabel;
  }
  __host__ __device__ static bool apply(VertexData *curLabel,
                                        GatherResult label) {
    bool changed = label < *curLabel;
    *curLabel = label < *curLabel ? label : *curLabel;
    return changed;
  }
  __host__ __device__ static void scatter(const VertexData *,
                                          const VertexData *, EdgeData *) {}
};

}  // namespace gas
